"""The indexing engine on one GPU: device-resident tables and masters, and the batch pipeline
normalize -> harmonic coefficients -> SO(3) correlation -> Gauss-Newton refinement + NMI ->
records, all through libzebsd_hip.so on torch's current stream (src/engine.zig:1793-2137).

PyTorch provides device memory and the stream; every kernel is the library's.
"""
import numpy as np
import torch

from . import _native as nat
from . import master as master_mod
from . import tables as tables_mod
from .geometry import Geometry, frame_of
from .settings import Settings

RECORD = 8
PIXEL_KIND = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}
TORCH_PIXEL_KIND = {torch.uint8: 0, torch.uint16: 1, torch.float32: 2} if hasattr(torch, "uint16") else {torch.uint8: 0, torch.float32: 2}
# device memory the correlation's spectra may take per chunk of items
SPEC_BUDGET = 4 << 30
# host scans up to this size are kept page-locked as well (Engine._pin)
PIN_BYTES = 8 << 30
# items whose coefficients and refinement go in one launch (the correlation takes them in
# SPEC_BUDGET chunks)
BATCH_ITEMS = 16384
# the fewest items a stream of index_cells takes
LANE_MIN_ITEMS = 2048


def device():
    if not torch.cuda.is_available():
        raise nat.NativeError("no GPU: the indexing path runs only on the GPU (libzebsd_hip.so), there is no CPU fallback")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(v):
    return v if v else None


def _dev(a, dev, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(dev)


class DeviceBand:
    """A bandwidth's tables on the device, with the zb_band struct pointing at them."""

    def __init__(self, L, dev):
        self.host = tables_mod.BandHost(L)
        h = self.host
        self.ring = _dev(h.ring, dev)
        self.wplm = _dev(h.wplm, dev)
        self.dq = _dev(h.dq, dev)
        self.tw = _dev(h.tw, dev)
        self.c = nat.Band(L, h.n, self.ring.data_ptr(), self.wplm.data_ptr(), self.dq.data_ptr(), self.tw.data_ptr())


class DevicePhase:
    """A phase's weight tables at a bandwidth (for one baked master), and its denominator volume
    per detector geometry."""

    def __init__(self, band, side, sheets, dev, rotations=()):
        self.host = tables_mod.PhaseHost(band.host, side, sheets)
        # a 2-fold axis along y among the master's rotations: its correlation repeats past
        # b = pi / 2 (C(a, b, c) = C(a + pi, pi - b, pi - c)), so the search can stop there
        self.y2 = any(abs(r[0]) < 1e-4 and abs(abs(r[2]) - 1) < 1e-4 for r in np.asarray(rotations, np.float64).reshape(-1, 4))
        h = self.host
        self.w_f = _dev(h.w_f, dev)
        self.w_f2 = _dev(h.w_f2, dev)
        self.c = nat.Phase(h.fold, h.nc, h.mu_max, h.l_step, self.w_f.data_ptr(), self.w_f2.data_ptr())
        self.rden = {}
        self.band = band
        self.dev = dev
        self._vt = [None, None]
        self._vth = [None, None]
        self._vt16 = {}
        self.vrden = {}

    def vtable_of(self, k):
        """The GEMM table (zb_vtable, device buffer) of the master (k = 0) or of its square
        (k = 1, for the denominators), built on first use."""
        if self._vt[k] is None:
            t = tables_mod.VTableHost(self.band.host, self.host, (self.host.hm.f, self.host.hm.f2)[k])
            q = _dev(t.tiled, self.dev)
            t.tiled = None  # on the device now
            self._vt[k] = (nat.VTable(t.L, t.n, t.nc, t.l_step, t.slots, t.cols, q.data_ptr()), q)
            self._vth[k] = t
        return self._vt[k]

    def vtables(self):
        """The GEMM tables of the master and of its square."""
        return [self.vtable_of(0), self.vtable_of(1)]

    def big_tables(self):
        """Whether the FP32 table passes 1 GB (GaN at L = 96: 2.4 GB), so that the denominators
        come from the slab path instead of a second table of the master's square."""
        L, n, nc = self.band.host.L, self.band.host.n, self.host.nc
        slots = (L + self.host.l_step - 1) // self.host.l_step
        cols, S, _ = tables_mod.vgemm_shape(nc, slots)
        return (n // 2 + 1) * L * S * 2 * cols * 4 > (1 << 30)

    def vtable16(self, kind, split_degree):
        """The master's GEMM table with 16-bit products (kind 1 FP16, 2 BF16; hi + lo products for
        degrees from split_degree on), built on first use."""
        key = (kind, split_degree)
        if key not in self._vt16:
            vt, q = self.vtable_of(0)
            t = self._vth[0]
            split = tables_mod.split_slot(split_degree, t.l_step)
            q16 = torch.from_numpy(t.table16(kind, split).view(np.int16)).to(self.dev)
            v16 = nat.VTable(t.L, t.n, t.nc, t.l_step, t.slots, t.cols, q.data_ptr(), kind, split, q16.data_ptr())
            self._vt16[key] = (v16, q, q16)
        return self._vt16[key][0]


def lattice(rows, cols, want):
    """About `want` cells spread evenly over a rows x cols map, row-major (src/engine.zig:1325-1341):
    every cell when the map has no more."""
    gy, gx = rows, cols
    if rows * cols > want:
        gy = min(max(int(np.floor(np.sqrt(want * rows / cols) + 0.5)), 1), rows)
        gx = min(max(want // gy, 1), cols)
    r = np.arange(gy, dtype=np.int64)[:, None]
    c = np.arange(gx, dtype=np.int64)[None, :]
    return (((2 * r + 1) * rows // (2 * gy)) * cols + (2 * c + 1) * cols // (2 * gx)).reshape(-1)


def mfma_correlation(L, n, nc=None, slots=None):
    """Whether the matrix-core correlation takes this shape: its transform along a is held in
    registers (n <= 128 one lane per column with nc <= 64 and at most 64 degree slots; n = 256
    two lanes per column with 32 <= nc <= 128 and at most 96 slots)."""
    if n > 256:
        return False
    if n == 256:
        return (nc is None or 32 <= nc <= 128) and (slots is None or slots <= 96)
    return (nc is None or nc <= 64) and (slots is None or slots <= 64)


def correlation_part(n, chunk):
    """Items a correlation launch takes when n items are cut into launches of at most `chunk`
    (the workspace's): the fewest launches, of equal whole 128-item GEMM tiles where the cap
    allows, so that no launch is a small remainder that leaves most of the GPU idle."""
    parts = -(-n // chunk)
    return min(chunk, -(-(-(-n // parts)) // 128) * 128)


class Lane:
    """A HIP stream (default priority) with scratch buffers of its own, so that consecutive
    batches run concurrently: one batch's GEMM on the matrix cores beside another's refinement
    on the vector units and gathers. Buffers only grow and are used only on this stream."""

    def __init__(self, dev):
        self.dev = dev
        self.stream = torch.cuda.Stream(device=dev)
        self.bufs = {}

    def buf(self, name, shape, dtype):
        numel = int(np.prod(shape))
        b = self.bufs.get(name)
        if b is None or b.numel() < numel or b.dtype != dtype:
            with torch.cuda.stream(self.stream):
                b = torch.empty((max(numel, 1),), dtype=dtype, device=self.dev)
            self.bufs[name] = b
        return b[:numel].view(shape)


class Phase:
    def __init__(self, m):
        self.master = m
        self.baked = {}     # bake key -> (sheets_dev [2, s, s], grad_dev [2, s, s, 4], sheets_host)
        self.baked_h = {}   # bake key -> grad_dev in half precision [2, s, s, 2, 4]: each texel and the one below (zb_master_pack)
        self.bands = {}     # (L, bake key) -> DevicePhase


class Engine:
    def __init__(self):
        self.settings = Settings()
        self.geo = Geometry()
        self.phases = []
        self.patterns = None      # numpy (rows, cols, h, w) or a device tensor
        self.shape = None         # (rows, cols, h, w)
        self.bands = {}
        self.stamp = 0
        self.dev = None
        # "mfma": the correlation as a GEMM on the matrix cores (n <= 256, at most 64 degree slots);
        # "slab": the spectrum + slab transforms (every bandwidth, and the fallback where the GEMM does not fit)
        self.correlator = "mfma"
        # bytes of the correlation's workspace (V) a launch may take: it cuts a batch into
        # correlation chunks (Solver.tune climbs it)
        self.spec_budget = SPEC_BUDGET
        # items a launch of the coefficients and the refinement takes (Solver.tune climbs it)
        self.batch_items = BATCH_ITEMS
        self._lanes = []
        # streams a large index_cells request is spread over (1: one; more measured no faster
        # through Solver.index, whose steps each wait for their records)
        self.index_lanes = 1
        self._bg = None           # (key, static background on the device)
        self.file_background = None  # [h, w] the scan file's static background (static_background "file")

    # ── state ──

    def _device(self):
        if self.dev is None:
            self.dev = device()
            nat.load()
        return self.dev

    def set_patterns(self, patterns):
        """(rows, cols, h, w) or (n, h, w) of uint8, uint16 or float32, as a numpy array or a
        device tensor (kept where it is)."""
        self._bg = None
        if isinstance(patterns, torch.Tensor):
            a = patterns
            if a.dim() == 3:
                a = a[None]
            if a.dim() != 4:
                raise ValueError("patterns must be (rows, cols, h, w) or (n, h, w)")
            if a.dtype not in TORCH_PIXEL_KIND:
                a = a.float()
            self.patterns = a.contiguous()
        else:
            a = np.asarray(patterns)
            if a.ndim == 3:
                a = a[None]
            if a.ndim != 4:
                raise ValueError("patterns must be (rows, cols, h, w) or (n, h, w)")
            if a.dtype not in PIXEL_KIND:
                a = a.astype(np.float32)
            self.patterns = np.ascontiguousarray(a)
        if getattr(self, "_pinned", None) is not None:
            # the previous scan's page-locked copy goes back to the system, not to torch's cache
            self._pinned = None
            if torch.cuda.is_available():
                torch.cuda.synchronize()
                torch._C._host_emptyCache()
        self._pinned = self._pin(self.patterns)
        self.shape = tuple(int(x) for x in self.patterns.shape)
        if self.shape[2] < 4 or self.shape[3] < 4:
            raise ValueError("patterns smaller than 4 x 4")

    @staticmethod
    def _pin(a):
        """A host scan (numpy) copied once into page-locked memory, up to PIN_BYTES: the steps'
        uploads then go by DMA at the link's rate, asynchronously, where a pageable upload is
        staged by the HIP runtime at about a third of it (and waits for the stream). None for a
        device tensor, without a GPU, or past the limit (then the pageable upload)."""
        if isinstance(a, torch.Tensor) or a.nbytes > PIN_BYTES or not torch.cuda.is_available():
            return None
        t = torch.empty(a.shape, dtype=torch.from_numpy(a[:0]).dtype, pin_memory=True)
        t.copy_(torch.from_numpy(a))
        return t

    def add_phase(self, path):
        self.phases.append(Phase(master_mod.load(path)))
        return len(self.phases) - 1

    def add_master(self, m):
        self.phases.append(Phase(m))
        return len(self.phases) - 1

    # ── geometry ──

    def pattern_hw(self):
        if self.shape is None:
            raise ValueError("no patterns")
        return self.shape[2], self.shape[3]

    def frame(self):
        h, w = self.pattern_hw()
        f = frame_of(self.geo, h, w, self.settings["mask"], self.settings["mask_radius"])
        c = nat.Frame()
        c.c0[:] = [float(x) for x in f["c0"]]
        c.c1[:] = [float(x) for x in f["c1"]]
        c.c2[:] = [float(x) for x in f["c2"]]
        c.pat_h, c.pat_w = h, w
        c.mask[:] = [float(x) for x in f["mask"]]
        return c

    def mask_tuple(self):
        h, w = self.pattern_hw()
        return tuple(float(x) for x in frame_of(self.geo, h, w, self.settings["mask"], self.settings["mask_radius"])["mask"])

    def translation(self):
        h, w = self.pattern_hw()
        return self.geo.translation(float(h), float(w))

    def positions(self, cells):
        """Detector position t (pixels) per map cell: the reference cell's less the beam's travel
        (src/engine.zig:1588-1613)."""
        rows, cols = self.shape[0], self.shape[1]
        cells = np.asarray(cells, np.int64)
        r = cells // cols
        c = cells % cols
        off = self.geo.scan_offset(r, c, 0.5 * (rows - 1), 0.5 * (cols - 1), self.settings["scan_x_sign"], self.settings["scan_y_sign"], 1.0)
        return (self.translation()[None, :] - off).astype(np.float32)

    def pixel_angle(self):
        h, _ = self.pattern_hw()
        if self.settings["pixel_angle_deg"] > 0:
            return np.radians(self.settings["pixel_angle_deg"])
        return self.geo.pixel_angle(float(h))

    # ── masters and tables ──

    def bake_key(self, p):
        """What the phase's master is filtered for (src/engine.zig:1674-1697): the smoothing of a
        pattern pixel, and the patterns' band pass and CLAHE at their sizes on the sphere."""
        s = self.settings
        px = master_mod.bake_px(p.master.side, self.pixel_angle())
        if px == 0.0:
            return (0.0, None, None)
        smooth = s["master_smooth"] * px
        bp = (s["bp_fine"] * px, s["bp_coarse"] * px) if s["bandpass"] else None
        cl = None
        if s["clahe"]:
            # tiles as wide on the sphere as the patterns': pattern height / grid pixels
            tile = self.pattern_hw()[0] / s["clahe_grid"] * px
            cl = (int(np.clip(np.round(p.master.side / tile), 1, 64)), int(s["clahe_bins"]), float(s["clahe_clip"]))
        return (smooth if (smooth > 0.05 or bp or cl) else 0.0, bp, cl)

    def _sphere_blur(self, cur, side, sigma):
        """The master's sheets smoothed on the sphere by sigma sheet pixels (passes of doubling
        width, src/prep.zig:215-237); a new buffer when any pass runs."""
        lib = nat.load()
        passes = master_mod.blur_passes(sigma)
        if not passes:
            return cur
        cur = cur.clone()
        tmp = torch.empty_like(cur)
        px_rad = np.sqrt(2 * np.pi) / side
        for s in passes:
            nat.check(lib.zb_master_blur(cur.data_ptr(), tmp.data_ptr(), side, float(s * px_rad), stream_ptr()), "master blur")
            cur, tmp = tmp, cur
        return cur

    def baked(self, pi):
        """The phase's master filtered for this pattern size and geometry (src/engine.zig:1716-1788):
        smoothed on the sphere by what a pattern pixel averages; with CLAHE, equalized as two
        images and averaged over the master's point group; with the band pass, the difference of
        two smoothings on the sphere; renormalized over both sheets; and its gradient field."""
        p = self.phases[pi]
        key = self.bake_key(p)
        if key in p.baked:
            return key, p.baked[key]
        dev = self._device()
        lib = nat.load()
        m = p.master
        side = m.side
        cur = _dev(m.raw, dev).reshape(-1)
        host = m.raw
        smooth, bp, cl = key
        if smooth > 0.0:
            fine = bp[0] if bp else 0.0
            if cl:
                cur = self._sphere_blur(cur, side, smooth)  # (a buffer of its own either way)
                ws = torch.empty((int(lib.zb_clahe_workspace(2, cl[0], cl[1])),), dtype=torch.uint8, device=dev)
                nat.check(lib.zb_clahe(cur.data_ptr(), 2, side, side, cl[0], cl[1], cl[2], ws.data_ptr(), ws.numel(), stream_ptr()), "master clahe")
                if len(m.ops) > 1:
                    ops = torch.from_numpy(np.asarray(m.ops, np.float32).reshape(-1)).to(dev)
                    out = torch.empty_like(cur)
                    nat.check(lib.zb_master_symmetrize(cur.data_ptr(), out.data_ptr(), side, ops.data_ptr(), len(m.ops), stream_ptr()),
                              "master symmetrize")
                    cur = out
            else:
                # what the pixel averages and the band pass's fine Gaussian compose into one
                fine = float(np.sqrt(smooth * smooth + fine * fine))
                if not bp:
                    cur = self._sphere_blur(cur, side, smooth)
            if bp:
                # the difference of two smoothings on the sphere
                low = self._sphere_blur(cur, side, bp[1])
                cur = self._sphere_blur(cur, side, fine) - low
            host = master_mod.normalize_sheets(cur.cpu().numpy().astype(np.float64).reshape(2, side, side))
            cur = _dev(host, dev).reshape(-1)
        grad = torch.empty((2 * side * side * 4,), dtype=torch.float32, device=dev)
        nat.check(lib.zb_master_gradient(cur.data_ptr(), side, grad.data_ptr(), stream_ptr()), "master gradient")
        p.baked[key] = (cur, grad, np.asarray(host, np.float32).reshape(2, side, side))
        return key, p.baked[key]

    def grad_half(self, pi):
        """The baked gradient field in half precision, which zb_refine's Gauss-Newton passes read."""
        p = self.phases[pi]
        key, (_, grad, _) = self.baked(pi)
        if key not in p.baked_h:
            gh = torch.empty((2 * p.master.side * p.master.side * 8,), dtype=torch.float16, device=grad.device)
            nat.check(nat.load().zb_master_pack(grad.data_ptr(), p.master.side, gh.data_ptr(), stream_ptr()), "master pack")
            p.baked_h[key] = gh
        return p.baked_h[key]

    def band(self, L):
        if L not in self.bands:
            self.bands[L] = DeviceBand(L, self._device())
        return self.bands[L]

    def phase_band(self, pi, L):
        key, (_, _, host) = self.baked(pi)
        p = self.phases[pi]
        k = (L, key)
        if k not in p.bands:
            p.bands[k] = DevicePhase(self.band(L), p.master.side, host, self._device(), p.master.rotations)
        return p.bands[k]

    def _window(self, band):
        """Harmonic coefficients and statistics of the reference cell's window (ones under the
        mask)."""
        lib = nat.load()
        fr = self.frame()
        t = self.translation().astype(np.float32)
        dev = self._device()
        win = torch.tensor([[t[0], t[1], t[2], 0.0]], dtype=torch.float32, device=dev)
        w_alm = torch.empty((1, band.host.pairs, 2), dtype=torch.float32, device=dev)
        w_stats = torch.empty((1, 4), dtype=torch.float32, device=dev)
        ws = torch.empty((max(lib.zb_sh_workspace(band.c, 1), 1),), dtype=torch.uint8, device=dev)
        nat.check(lib.zb_sh_coefficients(band.c, fr, None, win.data_ptr(), 1, 4, ws.data_ptr(), ws.numel(), w_alm.data_ptr(),
                                         w_stats.data_ptr(), stream_ptr()), "window coefficients")
        return w_alm, w_stats

    def _geo_key(self):
        fr = self.frame()
        t = self.translation().astype(np.float32)
        return (tuple(fr.c0), tuple(fr.c1), tuple(fr.c2), fr.pat_h, fr.pat_w, tuple(fr.mask), tuple(t))

    def rden(self, pi, L):
        """The denominator volume of the normalized correlation for the reference cell's detector
        position (src/engine.zig:1839-1849), in the slab path's order."""
        lib = nat.load()
        band = self.band(L)
        ph = self.phase_band(pi, L)
        key = self._geo_key()
        if key in ph.rden:
            return ph.rden[key]
        w_alm, w_stats = self._window(band)
        dev = self._device()
        n = band.host.n
        rden = torch.empty((n // 2 + 1) * n * ph.host.nc, dtype=torch.float32, device=dev)
        ws2 = torch.empty((nat.check(lib.zb_so3_workspace(band.c, ph.c, 1), "so3 workspace"),), dtype=torch.uint8, device=dev)
        nat.check(lib.zb_so3_denominators(band.c, ph.c, w_alm.data_ptr(), w_stats.data_ptr(), 1, ws2.data_ptr(), ws2.numel(),
                                          rden.data_ptr(), stream_ptr()), "denominators")
        ph.rden[key] = rden
        return rden

    def vrden(self, pi, L):
        """The same volume for the matrix-core path, in natural order [j][a][c]."""
        lib = nat.load()
        band = self.band(L)
        ph = self.phase_band(pi, L)
        key = self._geo_key()
        if key in ph.vrden:
            return ph.vrden[key]
        n, nc = band.host.n, ph.host.nc
        if ph.big_tables():
            # the slab path's volume (no table of the master's square), put in natural order:
            # slab position (p, cq, k) holds a = 2 bitrev(p) + k at c = bitrev(cq)
            nb, h = n // 2 + 1, n // 2
            slab = self.rden(pi, L).view(nb, h, nc, 2).permute(0, 2, 1, 3)
            bc = torch.from_numpy(tables_mod.bitrev(np.arange(nc), int(np.log2(nc)))).to(slab.device)
            bh = torch.from_numpy(tables_mod.bitrev(np.arange(h), int(np.log2(h)))).to(slab.device)
            rden = slab[:, bc][:, :, bh].reshape(-1).contiguous()
            ph.vrden[key] = rden
            return rden
        w_alm, w_stats = self._window(band)
        dev = self._device()
        (tf, _), (tf2, _) = ph.vtables()
        rden = torch.empty((n // 2 + 1) * n * nc, dtype=torch.float32, device=dev)
        ws = torch.empty((nat.check(lib.zb_vcorr_workspace(tf, 1, 1), "vcorr workspace"),), dtype=torch.uint8, device=dev)
        nat.check(lib.zb_vcorr_denominators(tf, tf2, w_alm.data_ptr(), w_stats.data_ptr(), 1, ws.data_ptr(), ws.numel(), rden.data_ptr(),
                                            stream_ptr()), "vcorr denominators")
        ph.vrden[key] = rden
        return rden

    def lanes(self, k=2):
        """k streams to spread batches over (run_batch(..., lane=...))."""
        dev = self._device()
        while len(self._lanes) < k:
            self._lanes.append(Lane(dev))
        return self._lanes[:k]

    def corr_vtable(self, ph):
        """The GEMM table the correlation multiplies with, at the precision the settings name."""
        kind = {"fp32": 0, "fp16": 1, "bf16": 2}[self.settings["correlation_precision"]]
        # the 16-bit form pairs 32 complex columns a group, for up to 64 degree slots and 64
        # columns: otherwise the FP32 form
        slots = (ph.host.hm.L + ph.host.l_step - 1) // ph.host.l_step
        if kind == 0 or not (32 <= ph.host.nc <= 64 and slots <= 64):
            vt = ph.vtable_of(0)[0]
        else:
            vt = ph.vtable16(kind, int(self.settings["split_degree"]))
        # b searched to pi / 2 only where the master's 2-fold axis along y repeats the rest (a
        # copy of the struct: the phase's own tables stay as built)
        vt = nat.VTable.from_buffer_copy(vt)
        vt.b_half = 1 if (self.settings["correlation_symmetry"] and ph.y2) else 0
        return vt

    def sh_flags(self):
        """zb_sh_coefficients' flags for patterns: centred, unit norm, and the Legendre sums'
        products at the correlation's precision (16-bit with hi + lo from split_degree)."""
        kind = self.settings["correlation_precision"]
        f = nat.FLAG_CENTER | nat.FLAG_NORM
        if kind == "fp16":
            f |= nat.FLAG_FP16 | nat.flag_split(min(int(self.settings["split_degree"]), 255))
        elif kind == "bf16":
            f |= nat.FLAG_BF16 | nat.flag_split(min(int(self.settings["split_degree"]), 255))
        return f

    def use_mfma(self, L, pi=None):
        """The matrix-core correlation for phase pi at L (every phase when pi is None)."""
        if self.correlator != "mfma":
            return False
        n = self.band(L).host.n
        phs = range(len(self.phases)) if pi is None else [pi]
        for k in phs:
            h = self.phase_band(k, L).host
            if not mfma_correlation(L, n, h.nc, (L + h.l_step - 1) // h.l_step):
                return False
        return mfma_correlation(L, n)

    def corr_chunk(self, n, L=None, budget=None):
        """Items per correlation launch for a batch of n at L under a workspace budget (bytes;
        the engine's spec_budget by default), as run_batch cuts it."""
        lib = nat.load()
        L = int(L or self.settings["bandwidth"])
        band = self.band(L)
        per = 1
        for pi in range(len(self.phases)):
            ph = self.phase_band(pi, L)
            if self.use_mfma(L, pi):
                ws = int(lib.zb_vcorr_workspace(self.corr_vtable(ph), 4096, 0))
            else:
                ws = int(lib.zb_so3_workspace(band.c, ph.c, 4096))
            per = max(per, ws // 4096 + 1)
        budget = self.spec_budget if budget is None else budget
        return max(1, min(int(n), int(budget) // per))

    # ── the batch pipeline ──

    def refine_opts(self):
        s = self.settings
        o = nat.RefineOpts()
        o.iters = int(s["refine_iters"]) if s["refine"] else 0
        o.max_step = float(np.radians(s["refine_step"]))
        o.scale0, o.scale_min, o.scale_max = 1.0, 0.02, 1e4
        o.nmi_bins = int(s["nmi_bins"])
        self.stamp = self.stamp % (1 << 24) + 1
        o.stamp = self.stamp
        return o

    # ── the pattern filters before the path ──

    def static_background(self):
        """The static background: the mean pattern over a lattice of `static_patterns` cells of
        the map, summed in float64 (src/engine.zig:1344-1385), on the device."""
        rows, cols, h, w = self.shape
        kind = self.settings["static_background"]
        key = (id(self.patterns), self.shape, int(self.settings["static_patterns"]), kind, id(self.file_background))
        if kind == "file":
            if self.file_background is None:
                raise ValueError('static_background "file": the scan stores no background')
            if self._bg is None or self._bg[0] != key:
                self._bg = (key, torch.as_tensor(np.asarray(self.file_background, np.float32)).reshape(h, w).to(self._device()).contiguous())
            return self._bg[1]
        if self._bg is None or self._bg[0] != key:
            cells = lattice(rows, cols, int(self.settings["static_patterns"]))
            flat = self.patterns.reshape(rows * cols, h, w)
            if isinstance(flat, torch.Tensor):
                sel = flat[torch.as_tensor(cells, device=flat.device)].to(torch.float64)
                tot = sel.sum(0)
            else:
                tot = torch.from_numpy(np.asarray(flat)[cells].astype(np.float64).sum(0))
            avg = (tot * (1.0 / len(cells))).to(torch.float32).to(self._device()).contiguous()
            self._bg = (key, avg)
        return self._bg[1]

    def prep_into(self, raw, n, h, w, out, st, lane=None):
        """Raw patterns [n, h, w] (device) to the normalized ones in `out`: the pattern filters the
        settings ask for (zb_prep_work), else the normalization alone (zb_normalize). On a lane's
        stream the filters' workspace is the lane's own."""
        lib = nat.load()
        s = self.settings
        mask = (nat.ctypes.c_float * 3)(*self.mask_tuple())
        kind = TORCH_PIXEL_KIND[raw.dtype]
        if s["static_background"] == "none" and s["dynamic_background"] == "none" and not s["clahe"] and not s["bandpass"]:
            nat.check(lib.zb_normalize(raw.data_ptr(), kind, n, h, w, mask, out.data_ptr(), st), "normalize")
            return
        o = nat.PrepOpts()
        bg = self.static_background() if s["static_background"] != "none" else None
        o.background = bg.data_ptr() if bg is not None else None
        o.dynamic = {"none": 0, "subtract": 1, "divide": 2}[s["dynamic_background"]]
        o.dynamic_sigma = float(s["dynamic_sigma"]) * w
        o.clahe = int(bool(s["clahe"]))
        o.clahe_grid, o.clahe_bins, o.clahe_clip = int(s["clahe_grid"]), int(s["clahe_bins"]), float(s["clahe_clip"])
        o.bandpass = int(bool(s["bandpass"]))
        o.bp_fine, o.bp_coarse = float(s["bp_fine"]), float(s["bp_coarse"])  # pixels of the pattern as stored (binning 1)
        # patterns past one block's LDS (with these filters) go through a workspace in global
        # memory, kept for the next batch of the same shape (one per stream: a lane's is its own)
        need = int(lib.zb_prep_workspace(n, h, w, nat.ctypes.byref(o)))
        if need < 0:
            raise nat.NativeError(f"prep: {nat.last_error()}")
        ws, wsz = None, 0
        if need > 0:
            if lane is not None:
                cur = lane.buf("prep_ws", (need,), torch.uint8)
            else:
                cur = getattr(self, "_prep_ws", None)
                if cur is None or cur.numel() < need or cur.device != out.device:
                    cur = self._prep_ws = torch.empty((need,), dtype=torch.uint8, device=out.device)
            ws, wsz = cur.data_ptr(), cur.numel()
        nat.check(lib.zb_prep_work(raw.data_ptr(), kind, n, h, w, mask, nat.ctypes.byref(o), ws, wsz, out.data_ptr(), st), "prep")

    @staticmethod
    def _pack_items(chunks, items):
        """The items of several chunks as one [sum n_i, 4] tensor, each pattern index moved past
        the chunks before its own (on the current stream)."""
        counts = [int(c.shape[0]) for c in chunks]
        sizes = torch.tensor([int(it.shape[0]) for it in items], device=items[0].device)
        starts = torch.tensor(np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.float32), device=items[0].device)
        out = torch.cat([it.to(torch.float32) for it in items])
        out[:, 3] += torch.repeat_interleave(starts, sizes)
        return out

    def run_batch(self, raw, items, L=None, workspace=None, lane=None):
        """The pipeline on device-resident inputs: raw [n, h, w] (uint8/uint16/float32 tensor) and
        items [n, 4] float32 (t, index into raw). Returns records [n_phases, n, 8] on the device
        (and the stamp they carry). Nothing is read back.

        With `lane` (from lanes()), the batch runs on that lane's stream with its buffers, after
        what the current stream has enqueued so far, and the call returns without joining it:
        the caller synchronizes before reading the records.

        Uneven chunks of a scan as they arrive (lists: raw_i [n_i, h, w], items_i with indices
        into raw_i) are indexed as one batch, each stage one launch over all of them (the
        refinement and the coefficients at most batch_items items a launch): the records are
        [n_phases, sum n_i, 8], chunk after chunk, each the record its item gets alone (no record
        depends on the batch it is in)."""
        lib = nat.load()
        dev = self._device()
        s = self.settings
        L = int(L or s["bandwidth"])
        chunks = list(raw) if isinstance(raw, (list, tuple)) else [raw]
        if isinstance(items, (list, tuple)):
            if len(items) != len(chunks):
                raise ValueError("run_batch: one items tensor per chunk")
            items = self._pack_items(chunks, items)
        n = int(items.shape[0])
        h, w = int(chunks[0].shape[-2]), int(chunks[0].shape[-1])
        if any(c.dtype != chunks[0].dtype or tuple(c.shape[-2:]) != (h, w) for c in chunks):
            raise ValueError("run_batch: chunks of one pixel type and pattern size")
        n_raw = sum(int(c.shape[0]) for c in chunks)
        band = self.band(L)
        fr = self.frame()
        mask = (nat.ctypes.c_float * 3)(*self.mask_tuple())
        n_ph = len(self.phases)
        opts = self.refine_opts()
        # per-phase tables (built on the current stream the first time)
        phs = [self.phase_band(pi, L) for pi in range(n_ph)]
        # the correlation per phase: on the matrix cores where its shape allows, else the VALU form
        mfma = [self.use_mfma(L, pi) for pi in range(n_ph)]
        rdens = [self.vrden(pi, L) if mfma[pi] else self.rden(pi, L) for pi in range(n_ph)]
        vts = [self.corr_vtable(ph) if mfma[pi] else None for pi, ph in enumerate(phs)]
        corr_ws = [(lambda k, vt=vts[pi]: int(lib.zb_vcorr_workspace(vt, k, 0))) if mfma[pi] else
                   (lambda k, ph=ph: int(lib.zb_so3_workspace(band.c, ph.c, k))) for pi, ph in enumerate(phs)]
        sheets = [self.baked(pi)[1][0] for pi in range(n_ph)]
        grads_h = [self.grad_half(pi) for pi in range(n_ph)]
        # the correlation's workspace bounds its chunk (bytes per item at scale; the matrix-core
        # path pads items); the coefficients and the refinement take the whole batch at once, so
        # that an uneven batch is one launch of each, not one per correlation chunk
        per_item = max(f(4096) for f in corr_ws) // 4096 + 1
        chunk = max(1, min(n, self.spec_budget // max(per_item, 1)))
        big = max(chunk, min(n, self.batch_items))
        pairs = band.host.pairs
        need = max(int(lib.zb_sh_workspace(band.c, big)), max(f(chunk) for f in corr_ws))
        if lane is None:
            st = stream_ptr()
            pats = torch.empty((n_raw, h, w), dtype=torch.float32, device=dev)
            records = torch.empty((n_ph, n, RECORD), dtype=torch.float32, device=dev)
            alm = torch.empty((big, pairs, 2), dtype=torch.float32, device=dev)
            stats = torch.empty((big, 4), dtype=torch.float32, device=dev)
            peaks = torch.empty((n_ph, big, 8), dtype=torch.float32, device=dev)
            order = torch.empty((n_ph, big), dtype=torch.int32, device=dev)
            qwork = torch.empty((nat.REFINE_QUEUE_BYTES,), dtype=torch.uint8, device=dev)
            if workspace is None or workspace.numel() < need:
                workspace = torch.empty((need,), dtype=torch.uint8, device=dev)
        else:
            lane.stream.wait_stream(torch.cuda.current_stream(dev))
            st = lane.stream.cuda_stream
            pats = lane.buf("pats", (n_raw, h, w), torch.float32)
            alm = lane.buf("alm", (big, pairs, 2), torch.float32)
            stats = lane.buf("stats", (big, 4), torch.float32)
            peaks = lane.buf("peaks", (n_ph, big, 8), torch.float32)
            order = lane.buf("order", (n_ph, big), torch.int32)
            qwork = lane.buf("qwork", (nat.REFINE_QUEUE_BYTES,), torch.uint8)
            workspace = lane.buf("work", (need,), torch.uint8)
            with torch.cuda.stream(lane.stream):
                records = torch.empty((n_ph, n, RECORD), dtype=torch.float32, device=dev)
            for c in chunks:
                c.record_stream(lane.stream)
            items.record_stream(lane.stream)
        off = 0
        for c in chunks:
            self.prep_into(c, int(c.shape[0]), h, w, pats[off:], st, lane)
            off += int(c.shape[0])
        sh_flags = self.sh_flags()
        for b0 in range(0, n, big):
            nbt = min(big, n - b0)
            it = items[b0:b0 + nbt]
            nat.check(lib.zb_sh_coefficients(band.c, fr, pats.data_ptr(), it.data_ptr(), nbt, sh_flags, workspace.data_ptr(), workspace.numel(),
                                             alm.data_ptr(), stats.data_ptr(), st), "coefficients")
            even = correlation_part(nbt, chunk)
            for pi, ph in enumerate(phs):
                for c0 in range(0, nbt, even):
                    nt = min(even, nbt - c0)
                    a_c, p_c = alm[c0:c0 + nt], peaks[pi, c0:c0 + nt]
                    if mfma[pi]:
                        nat.check(lib.zb_vcorr_correlate(vts[pi], band.tw.data_ptr(), a_c.data_ptr(), nt, rdens[pi].data_ptr(), nt,
                                                         workspace.data_ptr(), workspace.numel(), p_c.data_ptr(), st), "vcorr correlate")
                    else:
                        nat.check(lib.zb_so3_correlate(band.c, ph.c, a_c.data_ptr(), nt, rdens[pi].data_ptr(), nt, workspace.data_ptr(),
                                                       workspace.numel(), p_c.data_ptr(), st), "correlate")
                if s["refine_order"]:
                    # items side by side on an XCD look up nearby regions of the master
                    nat.check(lib.zb_refine_order(fr, it.data_ptr(), peaks[pi].data_ptr(), 8, nbt, order[pi].data_ptr(), st), "refine order")
            # every phase's refinement of the batch in one launch (up to four phases a launch; its
            # form, a block per item or a persistent grid on device-side queues, is the library's
            # refine_queue knob)
            for p0 in range(0, n_ph, nat.REFINE_MAX_PHASES):
                grp = range(p0, min(n_ph, p0 + nat.REFINE_MAX_PHASES))
                arr = (nat.RefinePhase * len(grp))(*[
                    nat.RefinePhase(sheets[pi].data_ptr(), grads_h[pi].data_ptr(), self.phases[pi].master.side, peaks[pi].data_ptr(), 8,
                                    order[pi].data_ptr() if s["refine_order"] else None, records[pi].data_ptr()) for pi in grp])
                nat.check(lib.zb_refine_phases(arr, len(grp), fr, pats.data_ptr(), it.data_ptr(), nbt, opts, b0, qwork.data_ptr(),
                                               qwork.numel(), st), "refine phases")
        return records, opts.stamp

    def graph(self, n, h, w, dtype=torch.float32, L=None):
        """run_batch for batches of n patterns h x w captured once as a HIP graph (GraphBatch):
        a replay launches the whole batch (prep, coefficients, correlation, refinement, NMI,
        records) with one call."""
        return GraphBatch(self, n, h, w, dtype, L)

    def _raw_for(self, cells):
        dev = self._device()
        rows, cols, h, w = self.shape
        flat = self.patterns.reshape(rows * cols, h, w)
        cells = np.asarray(cells, np.int64)
        # a run of consecutive cells (the usual request) is a slice: no gather copy
        run = len(cells) > 0 and cells[-1] - cells[0] == len(cells) - 1 and bool(np.all(np.diff(cells) == 1))
        pinned = getattr(self, "_pinned", None)
        if run and pinned is not None:
            # page-locked and never written again: the upload goes on a copy stream of its own,
            # beside the kernels of the step before, and the current stream waits for it alone
            up = self._upload_stream(dev)
            cur = torch.cuda.current_stream(dev)
            with torch.cuda.stream(up):
                out = pinned.reshape(rows * cols, h, w)[int(cells[0]):int(cells[-1]) + 1].to(dev, non_blocking=True)
            cur.wait_stream(up)
            out.record_stream(cur)
            return out
        if isinstance(flat, torch.Tensor):
            part = flat[int(cells[0]):int(cells[-1]) + 1] if run else flat[torch.as_tensor(cells, device=flat.device)]
            return part.to(dev).contiguous()
        part = flat[cells[0]:cells[-1] + 1] if run else flat[cells]
        return torch.from_numpy(np.ascontiguousarray(part)).to(dev)

    def index_cells(self, cells, L=None):
        """Records [n_phases, n, 8] (q, ncc, nmi, step, stamp) for the given map cells; the
        stamp is checked (src/engine.zig:2127-2137)."""
        return self.index_cells_finish(self.index_cells_begin(cells, L))

    def index_cells_begin(self, cells, L=None):
        """index_cells' launches, without waiting for them: a handle for index_cells_finish."""
        if not self.phases:
            raise ValueError("no phases")
        cells = np.asarray(cells, np.int64).reshape(-1)
        n = len(cells)
        if n == 0:
            return [], 0, None
        raw = self._raw_for(cells)
        t = self.positions(cells)
        dev = self._device()
        # a large request is cut into consecutive parts on streams of their own, so that one
        # part's correlation runs beside another's refinement (each cell's record does not depend
        # on the batch it is in)
        k = max(1, min(self.index_lanes, n // LANE_MIN_ITEMS))
        lanes = self.lanes(k) if k > 1 else [None]
        bounds = np.linspace(0, n, k + 1).astype(np.int64)
        parts = []
        for i in range(k):
            a, b = int(bounds[i]), int(bounds[i + 1])
            items = np.concatenate([t[a:b], np.arange(b - a, dtype=np.float32)[:, None]], 1)
            items_d = self._items_to_device(items, dev)
            parts.append(self.run_batch(raw[a:b], items_d, L, lane=lanes[i]))
        # the points where the request's records are complete (this stream's, and each lane's
        # stream's): the read-back waits for them alone, not for what is launched after them
        done = torch.cuda.Event()
        done.record()
        events = [done]
        for ln in lanes:
            if ln is not None:
                ev = torch.cuda.Event()
                ev.record(ln.stream)
                events.append(ev)
        return parts, k, events

    def _items_to_device(self, items, dev):
        """items [n, 4] float32 on the device without the host waiting for the stream's earlier
        kernels (a pageable upload would): through one of two pinned buffers allocated once and
        used in turn, each refilled only after its previous upload has completed, the upload on
        the copy stream."""
        if torch.device(dev).type != "cuda":
            return torch.from_numpy(items).to(dev)
        self._stage_k = 1 - getattr(self, "_stage_k", 0)
        stage = getattr(self, "_stage", None) or [[None, None], [None, None]]
        self._stage = stage
        buf, ev = stage[self._stage_k]
        flat = np.ascontiguousarray(items, np.float32).reshape(-1)
        if ev is not None:
            ev.synchronize()
        if buf is None or buf.numel() < flat.size:
            buf = torch.empty((max(flat.size, 1 << 16),), dtype=torch.float32, pin_memory=True)
        # (a numpy copy: torch's CPU copy hands 64K elements to its thread pool, which measured
        # 5 ms a call beside the GPU's runtime threads, against 4 us here)
        buf.numpy()[:flat.size] = flat
        # on the copy stream (beside the kernels queued before), the current stream waiting for it
        up = self._upload_stream(dev)
        cur = torch.cuda.current_stream(dev)
        with torch.cuda.stream(up):
            out = torch.empty(items.shape, dtype=torch.float32, device=dev)
            out.view(-1).copy_(buf[:flat.size], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(up)
        cur.wait_stream(up)
        out.record_stream(cur)
        stage[self._stage_k] = [buf, ev]
        return out

    def _upload_stream(self, dev):
        """The stream host uploads go on, so that they run beside the kernels queued before them."""
        up = getattr(self, "_upload", None)
        if up is None or up.device != torch.device(dev):
            up = self._upload = torch.cuda.Stream(device=dev)
        return up

    def index_cells_finish(self, handle):
        """The records of an index_cells_begin, read back and stamp-checked."""
        parts, k, events = handle
        if k == 0:
            return np.zeros((len(self.phases), 0, RECORD), np.float32)
        # read back on a stream of its own, after the request's events: a request launched
        # since (Solver.index_step launches the next step first) keeps running meanwhile
        rb = getattr(self, "_readback", None)
        if rb is None:
            rb = self._readback = torch.cuda.Stream(device=self._device())
        for ev in events:
            rb.wait_event(ev)
        outs = []
        for rec, stamp in parts:
            rec.record_stream(rb)
            with torch.cuda.stream(rb):
                o = rec.cpu().numpy()
            if not np.all(o[..., 7] == np.float32(stamp)):
                raise RuntimeError("DeviceSkippedBatch: records without this batch's stamp")
            outs.append(o)
        return outs[0] if k == 1 else np.concatenate(outs, axis=1)

    def nmi_at(self, pi, cells, quats):
        """NMI of each cell's pattern with phase pi's master at the given orientation (no
        refinement): zb_refine with no iterations. Returns float32 [n]."""
        lib = nat.load()
        dev = self._device()
        cells = np.asarray(cells, np.int64).reshape(-1)
        n = len(cells)
        raw = self._raw_for(cells)
        h, w = self.pattern_hw()
        pats = torch.empty((n, h, w), dtype=torch.float32, device=dev)
        mask = (nat.ctypes.c_float * 3)(*self.mask_tuple())
        st = stream_ptr()
        self.prep_into(raw, n, h, w, pats, st)
        t = self.positions(cells)
        items = torch.from_numpy(np.concatenate([t, np.arange(n, dtype=np.float32)[:, None]], 1)).to(dev)
        seeds = torch.from_numpy(np.ascontiguousarray(np.asarray(quats, np.float32).reshape(n, 4))).to(dev)
        o = self.refine_opts()
        o.iters = 0
        rec = torch.empty((n, RECORD), dtype=torch.float32, device=dev)
        _, (sheets, _, _) = self.baked(pi)
        nat.check(lib.zb_refine(sheets.data_ptr(), self.grad_half(pi).data_ptr(), self.phases[pi].master.side, self.frame(), pats.data_ptr(), items.data_ptr(),
                                seeds.data_ptr(), 4, n, o, rec.data_ptr(), 0, st), "nmi")
        return rec[:, 5].cpu().numpy()

    def simulate(self, pi, cells, quats):
        """Simulated patterns of phase pi at the given orientations, as cells would record them
        (src/engine.zig:2753-2769). Returns float32 (n, h, w)."""
        lib = nat.load()
        dev = self._device()
        h, w = self.pattern_hw()
        cells = np.asarray(cells, np.int64).reshape(-1)
        quats = np.asarray(quats, np.float32).reshape(-1, 4)
        t = self.positions(cells)
        poses = np.concatenate([t, np.zeros((len(cells), 1), np.float32), quats], 1).astype(np.float32)
        _, (sheets, _, _) = self.baked(pi)
        pd = torch.from_numpy(poses).to(dev)
        out = torch.empty((len(cells), h, w), dtype=torch.float32, device=dev)
        nat.check(lib.zb_project(sheets.data_ptr(), self.phases[pi].master.side, self.frame(), pd.data_ptr(), len(cells), out.data_ptr(),
                                 stream_ptr()), "project")
        return out.cpu().numpy()


class GraphBatch:
    """One batch shape of Engine.run_batch as a HIP graph: the inputs are copied into the graph's
    own buffers, the captured launches replay on the current stream, the records land in the
    graph's own buffer (valid until the next replay). Built by one eager run (which makes every
    table and volume the batch needs) and one capture; the library's entry points allocate,
    clear (with kernels, never a memset node) and synchronize nothing, so all of the batch is
    in the graph. The records carry the stamp of the capture: a replay's records are checked
    against it (Engine.index_cells' guard does not apply here)."""

    def __init__(self, engine, n, h, w, dtype=torch.float32, L=None):
        dev = engine._device()
        self.raw = torch.zeros((n, h, w), dtype=dtype, device=dev)
        self.items = torch.zeros((n, 4), dtype=torch.float32, device=dev)
        self.items[:, 3] = torch.arange(n, dtype=torch.float32, device=dev)
        engine.run_batch(self.raw, self.items, L)  # tables, volumes, workspaces
        torch.cuda.synchronize(dev)
        self.g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.g):
            self.records, self.stamp = engine.run_batch(self.raw, self.items, L)
        torch.cuda.synchronize(dev)
        # the filters' workspace the replays write into (the engine may later replace its own)
        self._prep_ws = getattr(engine, "_prep_ws", None)

    def __call__(self, raw, items):
        """records [phases, n, 8] for raw [n, h, w] and items [n, 4] (device tensors)."""
        self.raw.copy_(raw)
        self.items.copy_(items)
        self.g.replay()
        return self.records
