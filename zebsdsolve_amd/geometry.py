"""Where the detector is, and where each pattern's source point was (src/geometry.zig:1-107).

The sample frame has its origin at the source point of the map's reference cell; the detector
maps (row, col, 0) in pixels to x = R (row, col, 0) + t with R = R0(tilts) . Exp(rot). The
pattern centre (Bruker convention) gives t = R (0.5 - pcy rows, 0.5 - pcx cols, pcz rows).
A pattern recorded with the beam at s sees the detector at t - s.
"""
from dataclasses import dataclass, field

import numpy as np

from . import rot


@dataclass
class Geometry:
    pc: tuple = (0.5, 0.5, 0.5)
    sample_tilt_deg: float = 70.0
    detector_tilt_deg: float = 0.0
    rot: tuple = (0.0, 0.0, 0.0)
    px_um: float = 0.0
    step_x_um: float = 1.0
    step_y_um: float = 1.0
    extra: dict = field(default_factory=dict)

    def nominal(self):
        """R0: the Bruker construction (src/geometry.zig:56-62)."""
        a = np.radians(self.detector_tilt_deg - self.sample_tilt_deg)
        s, c = np.sin(a), np.cos(a)
        return np.array([[-s, 0, c], [0, 1, 0], [-c, 0, -s]])

    def rotation(self):
        return self.nominal() @ rot.to_mat(rot.so3_exp(np.asarray(self.rot, np.float64)))

    def translation_for(self, pc, rows, cols):
        """The detector position seen from a source at pattern-centre coordinates `pc`
        (src/geometry.zig:71-75)."""
        return self.rotation() @ np.array([0.5 - pc[1] * rows, 0.5 - pc[0] * cols, pc[2] * rows])

    def translation(self, rows, cols):
        return self.translation_for(self.pc, rows, cols)

    @staticmethod
    def pc_of(R, t, rows, cols):
        """The pattern centre a pose implies (src/geometry.zig:85-89)."""
        p = np.asarray(R).T @ np.asarray(t)
        return np.array([(0.5 - p[1]) / cols, (0.5 - p[0]) / rows, p[2] / rows])

    def pixel_angle(self, rows):
        """Small-angle width of one pixel at the pattern centre, radians (src/geometry.zig:93-95)."""
        return 1.0 / max(self.pc[2] * rows, 1e-9)

    def scan_offset(self, row, col, ref_row, ref_col, x_sign=-1, y_sign=-1, bin_=1.0):
        """Beam offset of map cell (row, col) from the reference cell, in pixels (sample frame);
        zero when the pixel size is unknown (src/geometry.zig:99-107). Vectorised."""
        row = np.asarray(row, np.float64)
        col = np.asarray(col, np.float64)
        if not (self.px_um > 0):
            return np.zeros(np.broadcast(row, col).shape + (3,))
        k = 1.0 / (self.px_um * bin_)
        return np.stack([y_sign * (row - ref_row) * self.step_y_um * k,
                         x_sign * (col - ref_col) * self.step_x_um * k,
                         np.zeros(np.broadcast(row, col).shape)], -1)

    def as_dict(self):
        return {"pc": list(self.pc), "sample_tilt_deg": self.sample_tilt_deg, "detector_tilt_deg": self.detector_tilt_deg,
                "rot": list(self.rot), "px_um": self.px_um, "step_x_um": self.step_x_um, "step_y_um": self.step_y_um,
                "plane_kind": "none"}


def frame_of(geo, h, w, mask_on=True, mask_radius=1.0):
    """The detector's attitude and circular mask as the kernels take them (src/engine.zig:1444-1448,
    1567-1579): columns of R, pattern shape, mask centre (row, col) and squared radius."""
    R = geo.rotation()
    if mask_on:
        r = mask_radius * 0.5 * min(h, w)
        mask = (0.5 * (h - 1), 0.5 * (w - 1), r * r)
    else:
        mask = (0.0, 0.0, 0.0)
    return {"c0": R[:, 0].copy(), "c1": R[:, 1].copy(), "c2": R[:, 2].copy(), "h": int(h), "w": int(w),
            "mask": np.array(mask, np.float32).astype(np.float64)}
