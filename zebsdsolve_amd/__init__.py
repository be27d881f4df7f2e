"""zebsdsolve_amd: the EBSD indexing hot path of zebsdsolve on one AMD MI355X (gfx950).

The surface mirrors the reference's Python package (python/zebsdsolve/__init__.py, api.py) for
the indexing path::

    import zebsdsolve_amd as ze

    res = ze.index(patterns, ["Ni_master_20kV.npz"], pc=(0.5, 0.5, 0.55), bandwidth=64)
    res.quaternions, res.ncc, res.phase

`Solver` is the engine driven a step at a time: set patterns (or synthesize a toy scan), set the
geometry, add phases, index. Every kernel runs from libzebsd_hip.so (hand-written HIP for gfx950)
on torch's current stream; there is no CPU path.

Conventions (as the reference's): quaternions (w, x, y, z) map sample directions to crystal
directions; Bunge angles in radians; map cells numbered row-major.
"""
import numpy as np

from . import _native
from . import rot as _rot
from . import sh
from . import sym as _sym
from .api import Result, index
from .settings import SettingError

__version__ = "0.1.0"

METHODS = ("spherical",)


class Error(RuntimeError):
    """An engine call failed; the message says which and why."""


def _hash_u32(x):
    """PCG-style integer hash (detector.wgsl:284-288), on uint32 arrays."""
    x = np.asarray(x, np.uint64) & 0xFFFFFFFF
    v = (x * 747796405 + 2891336453) & 0xFFFFFFFF
    w = (((v >> ((v >> 28) + 4)) ^ v) * 277803737) & 0xFFFFFFFF
    return ((w >> 22) ^ w) & 0xFFFFFFFF


def _unit_rand(x):
    return (_hash_u32(x) >> 8).astype(np.float64) / 16777216.0


def decoys(n, draw):
    """Uniform random orientations (Shoemake) for items 0..n-1 of draw `draw`, a different one
    per item and draw (detector.wgsl decoy, 296-308)."""
    b = np.arange(n, dtype=np.uint64)
    s = _hash_u32(b + 1) ^ ((draw * 0x9E3779B9) & 0xFFFFFFFF)
    u1 = _unit_rand(s)
    a2 = _unit_rand(s + 1) * 2 * np.pi
    a3 = _unit_rand(s + 2) * 2 * np.pi
    ra, rb = np.sqrt(1 - u1), np.sqrt(u1)
    return np.stack([ra * np.sin(a2), ra * np.cos(a2), rb * np.sin(a3), rb * np.cos(a3)], -1).astype(np.float32)


class Solver:
    """One engine on one GPU. Use as a context manager or call close()."""

    def __init__(self):
        from . import engine
        self._e = engine.Engine()
        self._chance = None
        self._chance_key = None
        self._res = None
        # a step's batch launched ahead of the step (index_step), and the state it belongs to
        self._pending = None
        self._gen = 0
        self._idx = None
        self._step_cells = 16384  # cells index_step takes at least (tune climbs it)
        self._events = []
        self._cancel = False

    def close(self):
        self._e = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ── settings and geometry ──

    def configure(self, **settings):
        """Set options by name (zebsdsolve_amd/settings.py lists them with the reference's names)."""
        self._changed()
        try:
            self._e.settings.apply(**settings)
        except SettingError as ex:
            raise Error(f"configure: {ex}") from None
        return self

    @property
    def settings(self):
        return self._e.settings.as_dict()

    @property
    def geometry(self):
        return self._e.geo.as_dict()

    def set_geometry(self, **fields):
        """Set geometry fields: pc (3), sample_tilt_deg, detector_tilt_deg, rot (3), px_um,
        step_x_um, step_y_um."""
        self._changed()
        g = self._e.geo
        for k, v in fields.items():
            if k in ("pc", "rot"):
                v = tuple(float(x) for x in v)
                if len(v) != 3:
                    raise Error(f"geometry field {k} takes three numbers")
                setattr(g, k, v)
            elif k in ("sample_tilt_deg", "detector_tilt_deg", "px_um", "step_x_um", "step_y_um"):
                setattr(g, k, float(v))
            else:
                raise Error(f"geometry has no field {k!r}")
        return self

    # ── scan ──

    def set_patterns(self, patterns):
        """Patterns (rows, cols, h, w) or (n, h, w): uint8, uint16 or float32, a numpy array or a
        device tensor (kept on the device)."""
        self._changed()
        self._opened = None
        self._e.file_background = None
        try:
            self._e.set_patterns(patterns)
        except ValueError as ex:
            raise Error(str(ex)) from None
        self._res = None
        return self

    def open(self, path):
        """Open a scan file: kikuchipy's h5ebsd (src/scan.zig:640-700): the map of patterns, its
        shape, steps and tilt, the pattern centre (the mean of the stored per-pattern ones), the
        detector pixel (its size times the binning) and the stored static background (used when
        `static_background` is "file"). Other formats are refused."""
        self._changed()
        from . import h5
        try:
            with h5.File(str(path)) as f:
                root = f["/"]
                scans = [k for k in root.keys() if k.startswith("Scan")]
                if not scans or "EBSD" not in root[scans[0]]:
                    raise Error(f"open: {path}: not an h5ebsd scan (no Scan */EBSD)")
                ebsd = root[scans[0]]["EBSD"]
                hdr, data = ebsd["Header"], ebsd["Data"]
                if "pcx" not in hdr or "patterns" not in data:
                    raise Error(f"open: {path}: an h5ebsd layout this reader does not take (kikuchipy's is read)")

                def num(g, k, default=np.nan):
                    return float(np.asarray(g[k].read()).reshape(-1)[0]) if k in g else default

                rows, cols = int(num(hdr, "n_rows")), int(num(hdr, "n_columns"))
                pats = data["patterns"].read()
                n, h, w = pats.shape[-3:] if pats.ndim >= 3 else (0, 0, 0)
                if rows * cols != n:
                    rows, cols = 1, n
                pc = tuple(float(np.mean(hdr[k].read())) for k in ("pcx", "pcy", "pcz"))
                geo = {"pc": pc, "sample_tilt_deg": num(hdr, "sample_tilt", 70.0), "detector_tilt_deg": num(hdr, "elevation_angle", 0.0)}
                px, binning = num(hdr, "detector_pixel_size", 0.0), num(hdr, "binning", 1.0)
                if px > 1:  # kikuchipy writes 1 when the size is unknown
                    geo["px_um"] = px * (binning if binning > 0 else 1)
                for k, name in (("step_x_um", "step_x"), ("step_y_um", "step_y")):
                    v = num(hdr, name)
                    if np.isfinite(v) and v > 0:
                        geo[k] = v
                bg = hdr["static_background"].read().astype(np.float32) if "static_background" in hdr else None
        except h5.H5Error as ex:
            raise Error(f"open: {ex}") from None
        self.set_patterns(np.ascontiguousarray(pats.reshape(rows, cols, h, w)))
        self.set_geometry(**geo)
        self._e.file_background = bg
        self._opened = {"path": str(path), "format": "kikuchipy h5ebsd", "has_pc": True, "has_background": bg is not None}
        return self

    @property
    def scan(self):
        if self._e.shape is None:
            raise Error("no scan")
        rows, cols, h, w = self._e.shape
        p = self._e.patterns
        dt = str(p.dtype).replace("torch.", "")
        info = {"rows": rows, "cols": cols, "pattern_shape": (h, w), "dtype": dt, "format": "patterns", "has_background": False,
                "has_pc": False}
        info.update({k: v for k, v in (getattr(self, "_opened", None) or {}).items() if k != "path"})
        return info

    def pattern(self, cell, processed=True):
        """One pattern as float32 (h, w): as it is indexed (filtered and normalized) or, with
        processed=False, as stored; "background": the static background subtracted from every
        pattern (an error when none is)."""
        e = self._e
        if processed == "background":
            if e.settings["static_background"] == "none":
                raise Error("pattern: static_background subtracts none")
            return e.static_background().cpu().numpy()
        h, w = e.pattern_hw()
        if not processed:
            flat = e.patterns.reshape(-1, h, w)[int(cell)]
            return (flat.float().cpu().numpy() if hasattr(flat, "cpu") else np.asarray(flat, np.float32)).reshape(h, w)
        raw = e._raw_for(np.array([int(cell)]))
        import torch
        from .engine import stream_ptr
        out = torch.empty((1, h, w), dtype=torch.float32, device=raw.device)
        e.prep_into(raw, 1, h, w, out, stream_ptr())
        return out[0].cpu().numpy()

    def shape(self, full=False):
        return self._e.pattern_hw()

    @property
    def pattern_shape(self):
        return self.shape()

    # ── phases ──

    def add_phase(self, path):
        """Load a master pattern (ebsdsim .npz); returns its index."""
        self._changed()
        try:
            k = self._e.add_phase(str(path))
        except (OSError, ValueError) as ex:
            raise Error(f"add_phase: {ex}") from None
        m = self._e.phases[k].master
        self._log(f"phase {k}: {m.name} ({m.side} px, point group {m.point_group}); orientations reduced by its "
                  f"{len(m.rotations)} rotations")
        return k

    def clear_phases(self):
        self._changed()
        self._e.phases = []

    @property
    def phases(self):
        out = []
        for p in self._e.phases:
            m = p.master
            out.append({"name": m.name, "side": m.side, "point_group": m.point_group, "laue": m.laue,
                        "rotations": len(m.rotations), "voltage_kv": m.voltage_kv, "lattice": m.lattice})
        return out

    def master(self, phase=0):
        """The phase's master as indexing compares patterns with it (smoothed for the pattern size
        and geometry): float32 [2, side, side]."""
        _, (_, _, host) = self._e.baked(phase)
        return host.copy()

    def simulate(self, quaternions, cells=None, phase=0, full=False):
        """Simulated patterns (n, h, w) of a phase at the given orientations, each as map cell
        cells[k] would record it."""
        q = np.asarray(quaternions, np.float32).reshape(-1, 4)
        cells = np.zeros(len(q), np.int64) if cells is None else np.asarray(cells, np.int64).reshape(-1)
        if len(cells) != len(q):
            raise Error("simulate: one cell per quaternion")
        return self._e.simulate(phase, cells, q)

    def synthesize(self, **settings):
        """A toy scan from the phases loaded (settings toy_*), in place of any scan: grains by the
        nearest of toy_grains random points, each a random orientation and phase, every cell
        simulated under the present geometry, scaled to darkest pixel 0 and mean 1, blended with a
        background, and Poisson-counted at toy_counts (src/engine.zig:2771-2898). Returns the
        answer: quaternions (rows, cols, 4), phase (rows, cols; -1 blank), grain, geometry.
        The random streams are numpy's (the reference draws from Zig's)."""
        self._changed()
        if settings:
            self.configure(**settings)
        s = self.settings
        if not self._e.phases:
            raise Error("synthesize: no phases")
        rows, cols, h, w = s["toy_rows"], s["toy_cols"], s["toy_height"], s["toy_width"]
        n = rows * cols
        self._e.set_patterns(np.zeros((rows, cols, h, w), np.float32))
        r = np.random.default_rng(s["toy_seed"])
        ng = min(s["toy_grains"], n)
        gx = r.random(ng) * cols
        gy = r.random(ng) * rows
        gq = _rot.qstd(r.normal(size=(ng, 4)))
        gp = r.integers(0, len(self._e.phases), ng)
        yy, xx = np.meshgrid(np.arange(rows) + 0.5, np.arange(cols) + 0.5, indexing="ij")
        d = (gx[None, None, :] - xx[..., None]) ** 2 + (gy[None, None, :] - yy[..., None]) ** 2
        grain = np.argmin(d, axis=-1).reshape(-1)
        q = gq[grain]
        spread = np.radians(s["toy_spread_deg"]) / np.sqrt(3.0)
        if spread > 0:
            q = _rot.qstd(_rot.qmul(_rot.so3_exp(spread * r.normal(size=(n, 3))), q))
        phase = gp[grain].astype(np.int32)
        if s["toy_blank"] > 0:
            phase = np.where(r.random(n) < s["toy_blank"], -1, phase).astype(np.int32)
        data = np.zeros((n, h, w), np.float64)
        self._cancel = False
        made = 0
        for p in range(len(self._e.phases)):
            sel = np.nonzero(phase == p)[0]
            for c0 in range(0, len(sel), 16384):
                if self._cancel:
                    raise Error("synthesize: cancelled")
                cs = sel[c0:c0 + 16384]
                data[cs] = self._e.simulate(p, cs, q[cs].astype(np.float32))
                made += len(cs)
                self._progress("synthesize", made, int((phase >= 0).sum()))
        crystals = phase >= 0
        mean = data[crystals].mean() if crystals.any() else 0.0
        low = data[crystals].min() if crystals.any() else 0.0
        gain = 1.0 / (mean - low) if crystals.any() and mean > low else 0.0
        back = np.ones((h, w))
        if s["toy_background_sigma"] > 0:
            sg = s["toy_background_sigma"] * w
            by, bx = np.meshgrid(np.arange(h) + 0.5 - 0.5 * h, np.arange(w) + 0.5 - 0.5 * w, indexing="ij")
            back = np.exp(-(bx * bx + by * by) / (2 * sg * sg))
            back *= h * w / back.sum()
        signal = np.where((crystals & (gain > 0))[:, None, None], (data - low) * gain, 1.0)
        inten = (1 - s["toy_contrast"]) * back[None] + s["toy_contrast"] * signal
        rn = np.random.default_rng(s["toy_noise_seed"] if s["toy_noise_seed"] else (s["toy_seed"] ^ 0x9E3779B97F4A7C15))
        out = rn.poisson(s["toy_counts"] * inten).astype(np.float32) if s["toy_counts"] > 0 else inten.astype(np.float32)
        self._e.set_patterns(out.reshape(rows, cols, h, w))
        self._res = None
        self._log(f"toy scan: {rows} × {cols} patterns of {h} × {w}, {ng} grains of {len(self._e.phases)} phase(s), "
                  f"{int((phase < 0).sum())} blank cells, contrast {s['toy_contrast']:.2f}, {s['toy_counts']:.4f} electrons a pixel, "
                  f"seed {s['toy_seed']}")
        return {"quaternions": q.astype(np.float32).reshape(rows, cols, 4), "phase": phase.reshape(rows, cols),
                "grain": grain.astype(np.int32).reshape(rows, cols), "geometry": self.geometry}

    # ── chance and acceptance ──

    def _calibrate(self):
        """What NMI chance gives per phase: about 256 patterns of the map on a lattice, each at 8
        random orientations, median and 1.4826 x median deviation (src/engine.zig:2202-2292)."""
        e = self._e
        # every setter drops the calibration (_changed); the key also guards changes made to the
        # engine underneath this surface
        key = (e.shape, repr(sorted(e.settings.as_dict().items())), repr(sorted(e.geo.as_dict().items())),
               tuple(id(p) for p in e.phases), id(e.patterns))
        if self._chance is not None and self._chance_key == key:
            return self._chance
        rows, cols = e.shape[0], e.shape[1]
        region = rows * cols
        n = min(region, 256)
        cells = np.array([(2 * j + 1) * region // (2 * n) for j in range(n)], np.int64)
        med, spread = [], []
        for pi in range(len(e.phases)):
            vals = []
            for d in range(8):
                vals.append(e.nmi_at(pi, cells, decoys(n, d)))
            c = np.sort(np.concatenate(vals).astype(np.float32))
            m = c[len(c) // 2]
            dev = np.sort(np.abs(c - m))
            med.append(float(m))
            spread.append(max(1.4826 * float(dev[len(c) // 2]), 1e-6))
            az = e.settings["accept_z"]
            self._log(f"chance, phase {pi}: NMI {med[-1]:.4f} ± {spread[-1]:.4f} over {n} patterns at random orientations; "
                      f"z {az:.0f} is NMI {med[-1] + az * spread[-1]:.4f}")
        self._chance = (np.array(med, np.float32), np.array(spread, np.float32))
        self._chance_key = key
        return self._chance

    def chance(self, phase=0):
        med, spread = self._calibrate()
        return float(med[phase]), float(spread[phase])

    # ── indexing ──

    def index_cells(self, cells, method=None, bandwidth=None, dict_grid=None):
        """Index map cells without storing anything: arrays (phases, n[, 4]) of quaternions (raw,
        not reduced), ncc, nmi, step (radians of the last improving refinement step) and z (NMI
        above chance in chance's standard deviations) (python/zebsdsolve/__init__.py:666-683)."""
        if method not in (None, "spherical"):
            raise Error(f"index_cells: method {method!r} is not implemented here (spherical only)")
        if not self._e.phases:
            raise Error("index_cells: no phases")
        return self._out_of(self._e.index_cells(np.asarray(cells, np.int64).reshape(-1), bandwidth))

    def _out_of(self, rec):
        med, spread = self._calibrate()
        z = (rec[..., 5] - med[:, None]) / spread[:, None]
        return {"quaternions": rec[..., :4].copy(), "ncc": rec[..., 4].copy(), "nmi": rec[..., 5].copy(), "step": rec[..., 6].copy(), "z": z}

    def _changed(self):
        """Settings, geometry, patterns or phases changed: a batch launched ahead is stale, and so
        is the chance calibration (every score depends on them; the reference keys its
        calibration on every setting, the geometry and the scan's and phases' serials,
        src/engine.zig:2202-2240)."""
        self._drop_pending()
        self._chance = None
        self._chance_key = None

    def _drop_pending(self):
        self._gen = getattr(self, "_gen", 0) + 1
        self._pending = None

    # ── events and cancelling ──

    def _log(self, text):
        self._events.append({"kind": "log", "text": text})

    def _progress(self, stage, done, total, **more):
        self._events.append({"kind": "progress", "stage": stage, "done": int(done), "total": int(total), **more})

    def events(self):
        """Progress events queued since the last call, as a list of dicts: {"kind": "log",
        "text": ...} and {"kind": "progress", "stage": ..., "done": ..., "total": ...}
        (python/zebsdsolve/__init__.py:114-123; src/events.zig:1-4)."""
        out, self._events = self._events, []
        return out

    def cancel(self):
        """Ask a running indexing, tuning or toy-scan synthesis to stop at its next step: index
        keeps what it has (index_step returns 0), tune and synthesize raise Error
        (src/exports.zig:62-65; checked between steps as src/engine.zig:2429)."""
        self._cancel = True

    def _sweep_bandwidth(self, i):
        return self._e.settings.sweep_bandwidth_of(i)

    def index_begin(self):
        """Start indexing the region; index_step then does a few batches at a time. Clears a
        cancel() asked for before (src/engine.zig:2335-2348)."""
        self._drop_pending()
        self._cancel = False
        if not self._e.phases:
            raise Error("index: no phases")
        rows, cols = self._e.shape[0], self._e.shape[1]
        n = rows * cols
        self._res = {"quat": np.zeros((n, 4), np.float32), "ncc": np.zeros(n, np.float32), "nmi": np.zeros(n, np.float32),
                     "phase": np.full(n, -1, np.int32), "z": np.zeros(n, np.float32), "sweep": np.full(n, -1, np.int32),
                     "settled": np.zeros(n, bool)}
        self._todo = np.arange(n, dtype=np.int64)
        self._done = 0
        self._accepted = 0
        self._sweep = 0

    def index_step(self):
        """Index the next few batches of the sweep in progress; when the sweep ends, the cells
        without an accepted match become the next sweep's, at its bandwidth. Returns the cells
        left to visit (0: done) (src/engine.zig:2426-2465)."""
        if self._res is None:
            self.index_begin()
        left = len(self._todo) - self._done
        if left == 0 or self._cancel:
            return 0
        s = self.settings
        n = self._step_size(left)
        cells = self._todo[self._done:self._done + n]
        L = self._sweep_bandwidth(self._sweep)
        pend, self._pending = self._pending, None
        if pend is not None and pend[0] == (self._gen, self._sweep, self._done, n):
            handle = pend[1]
        else:
            handle = self._e.index_cells_begin(cells, L)
        # the sweep's next step goes to the GPU before this one is read back and landed on the
        # host, so the device does not wait for the host between steps
        nd = self._done + n
        if nd < len(self._todo):
            n2 = self._step_size(len(self._todo) - nd)
            self._pending = ((self._gen, self._sweep, nd, n2), self._e.index_cells_begin(self._todo[nd:nd + n2], L))
        out = self._out_of(self._e.index_cells_finish(handle))
        self._land(cells, out)
        self._done += n
        self._progress("index", self._done, len(self._todo), sweep=self._sweep, sweeps=s["sweeps"])
        if self._done == len(self._todo) and self._sweep + 1 < s["sweeps"]:
            # the next sweep takes what this one left
            keep = self._todo[self._res["sweep"][self._todo] < 0]
            self._log(f"sweep {self._sweep + 1} of {s['sweeps']} (spherical, bandwidth {L}): {len(self._todo) - len(keep)} of "
                      f"{len(self._todo)} patterns accepted, {len(keep)} left")
            if len(keep):
                self._todo = keep
                self._done = 0
                self._sweep += 1
        return len(self._todo) - self._done

    def _step_size(self, left):
        s = self.settings
        n = min(left, max(s["batch"], 1) * max(2, min(16, 4096 // max(s["batch"], 1))) if s["batch"] < 4096 else left)
        return max(n, min(left, self._step_cells))

    def _land(self, cells, out):
        """Each cell keeps the better match, the accepted ones their sweep (src/engine.zig:2390-2419)."""
        s = self.settings
        res = self._res
        settle = np.radians(s["accept_step_deg"])
        score = out["nmi"] if s["phase_by"] == "nmi" else out["ncc"]
        n_ph = score.shape[0]
        p = np.zeros(score.shape[1], np.int64)
        best = np.full(score.shape[1], -np.inf, np.float32)
        for k in range(n_ph):
            better = score[k] > best
            p = np.where(better, k, p)
            best = np.where(better, score[k], best)
        ar = np.arange(len(cells))
        nmi, z, step = out["nmi"][p, ar], out["z"][p, ar], out["step"][p, ar]
        good = nmi >= s["accept_nmi"] if s["accept_nmi"] > 0 else z >= s["accept_z"]
        take = (res["phase"][cells] < 0) | (good & (step <= settle)) | (z > res["z"][cells])
        for k in range(n_ph):
            sel = take & (p == k)
            if not sel.any():
                continue
            q = _sym.to_fz(out["quaternions"][k, ar[sel]].astype(np.float64), self._e.phases[k].master.rotations)
            c = cells[sel]
            res["quat"][c] = q.astype(np.float32)
            res["ncc"][c] = out["ncc"][k, ar[sel]]
            res["nmi"][c] = nmi[sel]
            res["z"][c] = z[sel]
            res["phase"][c] = k
            res["settled"][c] = step[sel] <= settle
        kept_good = res["nmi"][cells] >= s["accept_nmi"] if s["accept_nmi"] > 0 else res["z"][cells] >= s["accept_z"]
        newly = (res["sweep"][cells] < 0) & kept_good & res["settled"][cells]
        res["sweep"][cells[newly]] = self._sweep
        self._accepted += int(newly.sum())

    @property
    def index_status(self):
        total = 0 if self._res is None else len(self._todo)
        region = 0 if self._res is None else len(self._res["phase"])
        sw = getattr(self, "_sweep", 0)
        return {"sweep": sw, "sweeps": self.settings["sweeps"], "bandwidth": self._sweep_bandwidth(sw), "dict_grid": 0,
                "done": getattr(self, "_done", 0), "total": total, "region": region, "accepted": getattr(self, "_accepted", 0),
                "method": "spherical"}

    @property
    def sweep_plan(self):
        """What each sweep searches with, in order: a list of (bandwidth, dictionary grid,
        method) (python/zebsdsolve/__init__.py:748-755; src/settings.zig:465-483). The
        dictionary methods are not implemented here, so the grid is 0 and the method spherical."""
        return [(int(self._sweep_bandwidth(i)), 0, "spherical") for i in range(int(self.settings["sweeps"]))]

    def index(self, progress=None):
        """Index the whole map a few batches at a time; progress(status) after each step, False
        from it stops. Returns results()."""
        self.index_begin()
        while True:
            left = self.index_step()
            if progress is not None and progress(self.index_status) is False:
                break
            if left == 0:
                break
        return self.results()

    def results(self):
        """quaternions (rows, cols, 4) in each phase's fundamental zone, euler (Bunge, radians),
        ncc, nmi, phase (-1: not indexed), z, sweep (-1: no accepted match)."""
        if self._res is None:
            raise Error("results: nothing indexed")
        rows, cols = self._e.shape[0], self._e.shape[1]
        r = self._res
        return {"quaternions": r["quat"].reshape(rows, cols, 4).copy(),
                "euler": _rot.to_bunge(r["quat"].astype(np.float64)).astype(np.float32).reshape(rows, cols, 3),
                "ncc": r["ncc"].reshape(rows, cols).copy(), "nmi": r["nmi"].reshape(rows, cols).copy(),
                "phase": r["phase"].reshape(rows, cols).copy(), "z": r["z"].reshape(rows, cols).copy(),
                "sweep": r["sweep"].reshape(rows, cols).copy()}

    # ── symmetry ──

    def reduce(self, quaternions, phase=0):
        """Orientations as the results store them: (quaternions (n, 4) float32, euler (n, 3)
        float64), each quaternion's representative in the phase's fundamental zone and its Bunge
        angles in radians (python/zebsdsolve/__init__.py:907-916)."""
        if not 0 <= int(phase) < len(self._e.phases):
            raise Error(f"reduce: no phase {phase}")
        q = np.asarray(quaternions, np.float32).reshape(-1, 4).astype(np.float64)
        fz = _sym.to_fz(q, self._e.phases[int(phase)].master.rotations)
        return fz.astype(np.float32), _rot.to_bunge(fz)

    def disorientation(self, a, b, phase=0):
        """Disorientation angles (radians) between orientations a and b of one phase."""
        a = np.asarray(a, np.float64).reshape(-1, 4)
        b = np.asarray(b, np.float64).reshape(-1, 4)
        return _sym.disorientation(a, b, self._e.phases[phase].master.rotations)

    # ── timing ──

    def tune(self, progress=None):
        """Fit how the work is cut up to this machine and this scan (python/zebsdsolve/__init__.py
        tune): a sample of the region is indexed and timed, and each of the engine's knobs in
        turn climbs from the value it has, doubled while that is faster, else halved while that
        is: the cells indexed a step (`index_step`, 16 384), the correlation's workspace a
        launch may take (the engine's V budget, 4 GiB) and the streams a step is spread over
        (`Engine.index_lanes`, 1 to 4), the items a coefficient and refinement launch takes
        (`Engine.batch_items`, 16 384); then the library's launch-shape knobs (zb_knob_set,
        process-wide), each value of a small set tried in turn: the refinement's launch form
        (`refine_queue` 0, 1, 2) and the ring pairs a block of the sphere stage transforms
        (`rings_group` 4 to 32, 0 the library's default). The sample is sized by the bandwidth
        (8 192 patterns up to 32, 4 096 up to 64, 1 024 up to 128, 128 beyond) and timed the
        way index_step runs it. A move is tried only when it changes the sample's work (a step
        of at least the sample, or a budget that leaves the chunks as they are, is the same
        work), is kept only when it is faster by more than 3 % (timing noise), and a trial that
        fails (out of memory, say) is a loss, its value put back (the reference scores a failed
        trial the same way, python/zebsdsolve/__init__.py:705-710). Neither the orientations
        found nor which matches are accepted depend on these knobs. The reference's own knobs
        (batch, refine_pixels, wg_refine, wg_fft) cut up its WebGPU dispatches and mean nothing
        here. Returns the knobs before and after and a summary; `progress(trials_done)` is
        called after each timed trial, and returning False from it stops (the best so far
        stays); cancel() stops it with an Error."""
        import time

        import torch

        from .engine import LANE_MIN_ITEMS

        if not self._e.phases:
            raise Error("tune: no phases")
        self._cancel = False
        e = self._e
        L = int(self.settings["bandwidth"])
        rows, cols = e.shape[0], e.shape[1]
        want = 8192 if L <= 32 else 4096 if L <= 64 else 1024 if L <= 128 else 128
        cells = np.arange(min(want, rows * cols), dtype=np.int64)
        ns = len(cells)

        def steps_of(v):
            # the step sizes index_step would cut the sample into
            return tuple(min(v, ns - c0) for c0 in range(0, ns, v))

        def chunks_of(mb):
            return tuple(e.corr_chunk(k, L, mb << 20) for k in steps_of(self._step_cells))

        def lanes_of(v):
            return tuple(max(1, min(v, k // LANE_MIN_ITEMS)) for k in steps_of(self._step_cells))

        def batches_of(v):
            # the items a coefficient and refinement launch takes, per step (run_batch's rule)
            return tuple(max(e.corr_chunk(k, L), min(k, v)) for k in steps_of(self._step_cells))

        # name: (get, put, lowest, highest, the work a value gives the sample)
        knobs = {"step_cells": (lambda: self._step_cells, lambda v: setattr(self, "_step_cells", v), 256, 1 << 20, steps_of),
                 "spec_budget_mb": (lambda: e.spec_budget >> 20, lambda v: setattr(e, "spec_budget", v << 20), 64, 1 << 16, chunks_of),
                 "index_lanes": (lambda: e.index_lanes, lambda v: setattr(e, "index_lanes", v), 1, 4, lanes_of),
                 "batch_items": (lambda: e.batch_items, lambda v: setattr(e, "batch_items", v), 1024, 1 << 16, batches_of)}
        before = {k: g() for k, (g, _, _, _, _) in knobs.items()}
        done = 0

        def trial():
            nonlocal done
            if self._cancel:
                raise Error("tune: cancelled")
            try:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for c0 in range(0, ns, self._step_cells):
                    e.index_cells(cells[c0:c0 + self._step_cells], L)
                torch.cuda.synchronize()
                t = time.perf_counter() - t0
            except (RuntimeError, _native.NativeError):
                # a value this machine cannot run loses
                torch.cuda.synchronize()
                t = float("inf")
            done += 1
            self._progress("tune", done, 20)
            go = progress(done) if progress is not None else True
            return t, go is not False

        trial()  # tables, volumes and workspaces, outside the timings
        best, go = trial()
        if not np.isfinite(best):
            raise Error("tune: the sample does not run with the present settings")
        for name, (get, put, lo, hi, work) in knobs.items():
            if not go:
                break
            for factor in (2, 0.5):
                moved = False
                while go:
                    cur = get()
                    nxt = int(cur * factor)
                    # values past the range, or that give the sample the same work, are skipped
                    while lo <= nxt <= hi and nxt != cur and work(nxt) == work(cur):
                        nxt = int(nxt * factor)
                    if not (lo <= nxt <= hi) or nxt == cur:
                        break
                    put(nxt)
                    t, go = trial()
                    if t < 0.97 * best:
                        best, moved = t, True
                    else:
                        put(cur)
                        break
                if moved:
                    break
        # the library's launch-shape knobs (zb_knob_set): each value of a small set tried in
        # turn, kept when faster by the same margin; values that give the same launches as the
        # present one are skipped
        lib = _native.load()
        n = sh.fft_size(L)

        def group_of(v):
            # k_rings' ring pairs a block, as the library derives them from the knob
            g = v if v > 0 else 2048 // (2 * n)
            g = 1 << (max(g, 1).bit_length() - 1)
            return max(1, min(g, 8192 // (2 * n), n // 2))

        choices = {"refine_queue": ((0, 1, 2), lambda v: v), "rings_group": ((0, 4, 8, 16, 32), group_of)}
        for name, (values, launch) in choices.items():
            before[name] = lib.zb_knob_get(name.encode())
            for v in values:
                if not go:
                    break
                cur = lib.zb_knob_get(name.encode())
                if launch(v) == launch(cur):
                    continue
                lib.zb_knob_set(name.encode(), v)
                t, go = trial()
                if t < 0.97 * best:
                    best = t
                else:
                    lib.zb_knob_set(name.encode(), cur)
        after = {k: g() for k, (g, _, _, _, _) in knobs.items()}
        after.update({name: lib.zb_knob_get(name.encode()) for name in choices})
        summary = f"tuned on {ns} patterns at L={L}: {ns / best:.0f} patterns/s in {done} trials; " + \
                  ", ".join(f"{k} {before[k]} -> {after[k]}" for k in after)
        self._log(summary)
        return {"before": before, "after": after, "summary": summary}

    def tune_schedule(self, bandwidths=(16, 24, 32, 48), tol_deg=0.1, progress=None):
        """Pick the bandwidth schedule of the sweeps on a sample of the region: every schedule
        of two sweeps that starts at one of `bandwidths` below `bandwidth` (the cells the cheap
        first sweep does not accept go to the second at `bandwidth`) is timed against the one
        sweep at `bandwidth`, the way index() runs it (src/engine.zig:2426-2465), and the fastest
        whose results agree with the one sweep's is kept in the settings (`sweeps`,
        `sweep_bandwidth`): as many cells accepted (at most 0.5 % of the sample fewer), and of
        the cells both accept, at most 0.5 % in another phase or more than tol_deg apart. Unlike
        tune(), this changes results (within that agreement), so it is a separate call.
        Returns {"schedules": [(sweeps, sweep_bandwidth, seconds, agreed)], "chosen": (sweeps,
        sweep_bandwidth), "summary"}; progress(trials_done) after each trial, False stops."""
        import time

        import torch

        if not self._e.phases:
            raise Error("tune_schedule: no phases")
        s = self.settings
        L = int(s["bandwidth"])
        rows, cols = self._e.shape[0], self._e.shape[1]
        want = 8192 if L <= 32 else 4096 if L <= 64 else 1024 if L <= 128 else 128
        region = rows * cols
        ns = min(want, region)
        # cells spread over the region
        sample = np.array([(2 * j + 1) * region // (2 * ns) for j in range(ns)], np.int64)
        before = (int(s["sweeps"]), int(s["sweep_bandwidth"]))
        self._cancel = False

        def run(sweeps, sb):
            self.configure(sweeps=sweeps, sweep_bandwidth=sb)
            self._calibrate()  # chance's NMI, outside the timing (the same for every schedule)
            self.index_begin()
            self._todo = sample.copy()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            while self.index_step():
                pass
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            if self._cancel:
                raise Error("tune_schedule: cancelled")
            r = self._res
            return t, {k: r[k][sample].copy() for k in ("quat", "phase", "sweep")}

        cands = [(2, int(b)) for b in sorted(set(bandwidths)) if 8 <= b < L]
        out, done = [], 0
        try:
            run(1, before[1])  # tables at every bandwidth tried, outside the timings
            for sw, sb in cands:
                run(sw, sb)
            t_one, ref = run(1, before[1])
            out.append((1, before[1], t_one, True))
            best, chosen = t_one, (1, before[1])
            acc_ref = ref["sweep"] >= 0
            for sw, sb in cands:
                t, got = run(sw, sb)
                acc = got["sweep"] >= 0
                both = acc & acc_ref
                bad = both & (got["phase"] != ref["phase"])
                same = both & ~bad
                far = np.zeros(ns, bool)
                for k in range(len(self._e.phases)):
                    sel = same & (ref["phase"] == k)
                    if sel.any():
                        far[sel] = np.degrees(self.disorientation(got["quat"][sel], ref["quat"][sel], phase=k)) > tol_deg
                agreed = acc.sum() >= acc_ref.sum() - 0.005 * ns and (bad.sum() + far.sum()) <= 0.005 * max(int(both.sum()), 1)
                out.append((sw, sb, t, bool(agreed)))
                if agreed and t < 0.97 * best:
                    best, chosen = t, (sw, sb)
                done += 1
                if progress is not None and progress(done) is False:
                    break
        finally:
            self._res = None
            self._todo = np.zeros(0, np.int64)
        self.configure(sweeps=chosen[0], sweep_bandwidth=chosen[1])
        summary = f"schedule on {ns} patterns at L={L}: " + ", ".join(
            f"{sw} sweep{'s' if sw > 1 else ''} from {sb if sw > 1 else L}: {ns / t:.0f}/s{'' if ok else ' (disagrees)'}" for sw, sb, t, ok in out) + \
            f"; chosen {chosen[0]} sweep{'s' if chosen[0] > 1 else ''}" + (f" from {chosen[1]}" if chosen[0] > 1 else "")
        self._log(summary)
        return {"schedules": out, "chosen": chosen, "summary": summary}

    def profile(self, on=True):
        """Time every kernel launch on the GPU from now on (events around each); returns True
        (python/zebsdsolve/__init__.py profile)."""
        _native.profile(on)
        return True

    def profile_report(self):
        """GPU time per kernel since profile(True), slowest first, as text: a total line
        "ms count dispatches  total", then "ms count  kernel" per kernel (the reference's
        layout, src/gpu/gpu.zig:291-308)."""
        rows = sorted(_native.profile_report().items(), key=lambda kv: -kv[1][1])
        total = sum(ms for _, (_, ms) in rows)
        calls = sum(c for _, (c, _) in rows)
        out = [f"{total:9.3f} ms {calls:6d} dispatches  total"]
        out += [f"{ms:9.3f} ms {c:6d}  {name}" for name, (c, ms) in rows]
        return "\n".join(out) + "\n"


__all__ = ["Solver", "Result", "index", "Error", "decoys", "__version__"]
