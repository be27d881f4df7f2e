"""Options of the indexing path, by the reference's names and defaults (src/settings.zig).

`apply` rejects unknown keys and values out of range atomically, as Settings.apply/check does
(src/settings.zig:404-463). Keys of the reference that belong to stages not implemented here
(NLPAR, dictionary and PCA search, binning, the fits) are accepted only at the values that leave
those stages out, and say so otherwise.
"""
import copy
import math

DEFAULTS = {
    # ── a toy scan (synthesize) ──
    "toy_seed": 1, "toy_rows": 48, "toy_cols": 64, "toy_grains": 24, "toy_spread_deg": 0.0, "toy_blank": 0.0,
    "toy_height": 60, "toy_width": 60, "toy_contrast": 1.0, "toy_background_sigma": 0.0, "toy_counts": 20.0,
    "toy_noise_seed": 0,
    # ── scan ──
    "scan_x_sign": -1, "scan_y_sign": -1,
    # ── resolution ──
    "binning": 1.0, "full_resolution": "none",
    # ── averaging and pattern filters, in the order applied (NLPAR is not implemented: only
    # its "off" value) ──
    "nlpar": False, "static_background": "average", "static_patterns": 1024, "dynamic_background": "subtract",
    "dynamic_sigma": 0.15, "clahe": False, "clahe_grid": 4, "clahe_bins": 64, "clahe_clip": 40.0,
    "bandpass": False, "bp_fine": 1.0, "bp_coarse": 8.0,
    # ── mask ──
    "mask": True, "mask_radius": 1.0,
    # ── master ──
    "pixel_angle_deg": 0.0, "master_smooth": 0.3,
    # ── spherical indexing ──
    "bandwidth": 64, "method": "spherical",
    # ── refinement ──
    "refine": True, "refine_iters": 8, "refine_step": 2.0,
    # (not in the reference) refine a batch's items in the order of the master region they look up,
    # so that items refined side by side share the caches; the records are the same either way
    "refine_order": True,
    # ── phases ──
    "phase_by": "nmi", "nmi_bins": 16,
    # ── the correlation's GEMM (not in the reference, which computes in float32): "fp32",
    # or "fp16" / "bf16": 16-bit products with float32 accumulation (hi + lo products for the
    # degrees from split_degree on) and the GEMM's output kept in the same 16-bit type, the
    # transform along a in float32; phases with fewer than 32 points along c stay "fp32" ──
    "correlation_precision": "fp32", "split_degree": 64,
    # ── the matrix-core correlation searches b only up to pi / 2 when the master has a 2-fold
    # axis along y (its correlation repeats: C(a, b, c) = C(a + pi, pi - b, pi - c)); this
    # build's own, the orientations found are the same up to the master's symmetry ──
    "correlation_symmetry": True,
    # ── how the work is cut up ──
    "batch": 256, "tile_mb": 256, "refine_pixels": 64, "wg_fft": 64, "wg_refine": 64,
    # ── sweeps and acceptance ──
    "sweeps": 1, "sweep_bandwidth": 32, "sweep_method": "same", "accept_nmi": 0.0, "accept_z": 15.0, "accept_step_deg": 0.2,
}

# values a key may take where only some of the reference's are implemented
_ONLY = {
    "nlpar": (False,), "static_background": ("average", "file", "none"), "dynamic_background": ("subtract", "divide", "none"),
    "method": ("spherical",), "binning": (1, 1.0), "full_resolution": ("none",),
    "sweep_method": ("same", "spherical"),
}


class SettingError(ValueError):
    pass


def check(s):
    def bad(cond, key):
        if cond:
            raise SettingError(f"setting {key}={s[key]!r} is out of range")
    bad(not (8 <= s["bandwidth"] <= 256), "bandwidth")
    bad(not (2 <= s["nmi_bins"] <= 32), "nmi_bins")
    bad(not (1 <= s["batch"] <= 65536), "batch")
    bad(not (8 <= s["tile_mb"] <= 65536), "tile_mb")
    bad(not (s["mask_radius"] > 0), "mask_radius")
    bad(not (s["master_smooth"] >= 0), "master_smooth")
    bad(not (s["pixel_angle_deg"] >= 0), "pixel_angle_deg")
    bad(s["refine_iters"] < 0 or s["refine_iters"] > 1000, "refine_iters")
    bad(not (s["refine_step"] > 0), "refine_step")
    bad(s["phase_by"] not in ("nmi", "ncc"), "phase_by")
    bad(not (1 <= s["toy_rows"] <= 4096 and 1 <= s["toy_cols"] <= 4096 and s["toy_grains"] >= 1), "toy_rows")
    bad(not (8 <= s["toy_height"] <= 1024 and 8 <= s["toy_width"] <= 1024), "toy_height")
    bad(not (s["toy_spread_deg"] >= 0 and 0 <= s["toy_blank"] <= 1 and 0 < s["toy_contrast"] <= 1 and s["toy_background_sigma"] >= 0), "toy_contrast")
    bad(not (0 <= s["toy_counts"] <= 1e6), "toy_counts")
    bad(abs(s["scan_x_sign"]) != 1 or abs(s["scan_y_sign"]) != 1, "scan_x_sign")
    bad(not (0 <= s["accept_nmi"] <= 1) or not (s["accept_step_deg"] > 0), "accept_nmi")
    bad(not (1 <= s["sweeps"] <= 6), "sweeps")
    bad(not (s["static_patterns"] >= 1), "static_patterns")
    bad(not (s["dynamic_sigma"] > 0), "dynamic_sigma")
    bad(not (1 <= s["clahe_grid"] <= 16 and 2 <= s["clahe_bins"] <= 256 and s["clahe_clip"] > 0), "clahe_grid")
    bad(not (s["bp_fine"] >= 0 and s["bp_coarse"] > s["bp_fine"]), "bp_fine")
    bad(not (8 <= s["sweep_bandwidth"] <= 256), "sweep_bandwidth")
    bad(s["correlation_precision"] not in ("fp32", "fp16", "bf16"), "correlation_precision")
    bad(not (0 <= s["split_degree"] <= 256), "split_degree")
    bad(not isinstance(s["correlation_symmetry"], bool), "correlation_symmetry")
    bad(not isinstance(s["refine_order"], bool), "refine_order")
    for k in ("wg_fft", "wg_refine"):
        bad(not (16 <= s[k] <= 256 and (s[k] & (s[k] - 1)) == 0), k)
    for k, allowed in _ONLY.items():
        if s[k] not in allowed:
            raise SettingError(f"setting {k}={s[k]!r} selects a stage not implemented here (supported: {allowed[0]!r})")


class Settings:
    def __init__(self):
        self.values = copy.deepcopy(DEFAULTS)

    def apply(self, **kw):
        nxt = dict(self.values)
        for k, v in kw.items():
            if k not in DEFAULTS:
                raise SettingError(f"unknown setting {k!r}")
            d = DEFAULTS[k]
            if isinstance(d, bool):
                if not isinstance(v, (bool, int)):
                    raise SettingError(f"setting {k} takes a boolean")
                v = bool(v)
            elif isinstance(d, int):
                if isinstance(v, float) and v.is_integer():
                    v = int(v)
                if not isinstance(v, int) or isinstance(v, bool):
                    raise SettingError(f"setting {k} takes an integer")
            elif isinstance(d, float):
                if not isinstance(v, (int, float)) or isinstance(v, bool):
                    raise SettingError(f"setting {k} takes a number")
                v = float(v)
            elif isinstance(d, str) and not isinstance(v, str):
                raise SettingError(f"setting {k} takes a string")
            nxt[k] = v
        check(nxt)
        self.values = nxt

    def __getitem__(self, k):
        return self.values[k]

    def sweep_bandwidth_of(self, i):
        """Bandwidth of sweep i of `sweeps`: sweep_bandwidth first, `bandwidth` last, equal ratios
        between, rounded (src/settings.zig:463-483)."""
        s = self.values
        if s["sweeps"] <= 1 or i + 1 >= s["sweeps"]:
            return s["bandwidth"]
        f = i / (s["sweeps"] - 1)
        a, b = float(s["sweep_bandwidth"]), float(s["bandwidth"])
        return int(math.floor(a * (b / a) ** f + 0.5))

    def as_dict(self):
        return dict(self.values)
