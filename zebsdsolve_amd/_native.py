"""ctypes binding of libzebsd_hip.so (include/zebsd.h), the in-tree build of csrc/.

The library is required: there is no CPU fallback. `load()` raises when it is missing, with the
command that builds it.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libzebsd_hip.so")

f32p = ctypes.POINTER(ctypes.c_float)
vp = ctypes.c_void_p
i32 = ctypes.c_int32
i64 = ctypes.c_int64


class Frame(ctypes.Structure):
    _fields_ = [("c0", ctypes.c_float * 3), ("c1", ctypes.c_float * 3), ("c2", ctypes.c_float * 3),
                ("pat_h", i32), ("pat_w", i32), ("mask", ctypes.c_float * 3)]


class Band(ctypes.Structure):
    _fields_ = [("L", i32), ("n", i32), ("ring", vp), ("wplm", vp), ("dq", vp), ("twiddle", vp)]


class Phase(ctypes.Structure):
    _fields_ = [("fold", i32), ("nc", i32), ("mu_max", i32), ("l_step", i32), ("w_f", vp), ("w_f2", vp)]


class VTable(ctypes.Structure):
    _fields_ = [("L", i32), ("n", i32), ("nc", i32), ("l_step", i32), ("slots", i32), ("cols", i32), ("q", vp),
                ("lp_kind", i32), ("lp_split", i32), ("qlp", vp), ("b_half", i32)]


class RefineOpts(ctypes.Structure):
    _fields_ = [("iters", i32), ("max_step", ctypes.c_float), ("scale0", ctypes.c_float), ("scale_min", ctypes.c_float),
                ("scale_max", ctypes.c_float), ("nmi_bins", i32), ("stamp", ctypes.c_uint32), ("order", vp)]


class RefinePhase(ctypes.Structure):  # zb_refine_phase
    _fields_ = [("master", vp), ("mgrad_h", vp), ("side", i32), ("seeds", vp), ("seed_stride", i32), ("order", vp), ("records", vp)]


REFINE_QUEUE_BYTES = 1024  # ZB_REFINE_QUEUE_BYTES
REFINE_MAX_PHASES = 4


class PrepOpts(ctypes.Structure):  # zb_prep_opts
    _fields_ = [("background", ctypes.c_void_p), ("dynamic", i32), ("dynamic_sigma", ctypes.c_float),
                ("clahe", i32), ("clahe_grid", i32), ("clahe_bins", i32), ("clahe_clip", ctypes.c_float),
                ("bandpass", i32), ("bp_fine", ctypes.c_float), ("bp_coarse", ctypes.c_float)]


_SIGS = {
    "zb_version": ([], ctypes.c_char_p),
    "zb_last_error": ([ctypes.c_char_p, i32], i32),
    "zb_normalize": ([vp, i32, i32, i32, i32, f32p, vp, vp], i32),
    "zb_prep": ([vp, i32, i32, i32, i32, f32p, ctypes.POINTER(PrepOpts), vp, vp], i32),
    "zb_prep_workspace": ([i32, i32, i32, ctypes.POINTER(PrepOpts)], i64),
    "zb_prep_work": ([vp, i32, i32, i32, i32, f32p, ctypes.POINTER(PrepOpts), vp, i64, vp, vp], i32),
    "zb_sh_workspace": ([ctypes.POINTER(Band), i32], i64),
    "zb_sh_coefficients": ([ctypes.POINTER(Band), ctypes.POINTER(Frame), vp, vp, i32, i32, vp, i64, vp, vp, vp], i32),
    "zb_sh_coefficients_noalloc": ([ctypes.POINTER(Band), ctypes.POINTER(Frame), vp, vp, i32, i32, vp, vp, vp], i32),
    "zb_so3_workspace": ([ctypes.POINTER(Band), ctypes.POINTER(Phase), i32], i64),
    "zb_so3_correlate": ([ctypes.POINTER(Band), ctypes.POINTER(Phase), vp, i32, vp, i32, vp, i64, vp, vp], i32),
    "zb_so3_denominators": ([ctypes.POINTER(Band), ctypes.POINTER(Phase), vp, vp, i32, vp, i64, vp, vp], i32),
    "zb_vtable_shape": ([i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)], i32),
    "zb_vtable16_shape": ([i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)], i32),
    "zb_vcorr_workspace": ([ctypes.POINTER(VTable), i32, i32], i64),
    "zb_vcorr_correlate": ([ctypes.POINTER(VTable), vp, vp, i32, vp, i32, vp, i64, vp, vp], i32),
    "zb_vcorr_denominators": ([ctypes.POINTER(VTable), ctypes.POINTER(VTable), vp, vp, i32, vp, i64, vp, vp], i32),
    "zb_refine": ([vp, vp, i32, ctypes.POINTER(Frame), vp, vp, vp, i32, i32, ctypes.POINTER(RefineOpts), vp, i32, vp], i32),
    "zb_refine_order": ([ctypes.POINTER(Frame), vp, vp, i32, i32, vp, vp], i32),
    "zb_refine_phases": ([ctypes.POINTER(RefinePhase), i32, ctypes.POINTER(Frame), vp, vp, i32, ctypes.POINTER(RefineOpts), i32, vp, i64, vp], i32),
    "zb_knob_set": ([ctypes.c_char_p, i32], i32),
    "zb_knob_get": ([ctypes.c_char_p], i32),
    "zb_local6": ([vp, i32, vp, vp], i32),
    "zb_sums_active": ([vp, i32, ctypes.POINTER(Frame), vp, vp, i32, i32, vp, vp, vp], i32),
    "zb_fit6_step": ([vp, vp, vp, i32, i32, ctypes.c_double, ctypes.c_double, vp, vp, vp, vp], i32),
    "zb_fit6_keep": ([vp, vp, i32, ctypes.c_double, vp, vp, vp, vp, vp, vp], i32),
    "zb_sums": ([vp, i32, ctypes.POINTER(Frame), vp, vp, i32, i32, vp, vp], i32),
    "zb_project": ([vp, i32, ctypes.POINTER(Frame), vp, i32, vp, vp], i32),
    "zb_master_blur": ([vp, vp, i32, ctypes.c_float, vp], i32),
    "zb_clahe_workspace": ([i32, i32, i32], i64),
    "zb_clahe": ([vp, i32, i32, i32, i32, i32, ctypes.c_float, vp, i64, vp], i32),
    "zb_master_symmetrize": ([vp, vp, i32, vp, i32, vp], i32),
    "zb_fit6_propose": ([vp, vp, vp, i32, i32, ctypes.c_double, ctypes.c_double, vp, vp, vp, vp], i32),
    "zb_fit6_accept": ([vp, vp, i32, ctypes.c_double, vp, vp, vp, vp, vp, vp], i32),
    "zb_master_gradient": ([vp, i32, vp, vp], i32),
    "zb_master_pack": ([vp, i32, vp, vp], i32),
    "zb_profile": ([i32], i32),
    "zb_profile_report": ([ctypes.c_char_p, i32], i32),
    "zb_profile_select": ([ctypes.c_char_p], i32),
}

EXPORTS = tuple(_SIGS)

# zb_sh_coefficients' flags (include/zebsd.h)
FLAG_CENTER, FLAG_NORM, FLAG_WINDOW, FLAG_FP16, FLAG_BF16 = 1, 2, 4, 8, 16


def flag_split(degree):
    """The flag bits naming the first degree whose Legendre products take hi + lo."""
    return (int(degree) & 0xFF) << 16

_lib = None


class NativeError(RuntimeError):
    """A call into libzebsd_hip.so failed; the message is the library's."""


def load(path=LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NativeError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                          "or `python -m zebsdsolve_amd._build` (there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def last_error():
    lib = load()
    buf = ctypes.create_string_buffer(1024)
    n = lib.zb_last_error(buf, 1024)
    return buf.raw[:min(n, 1023)].decode(errors="replace")


def check(rc, what):
    if rc < 0:
        raise NativeError(f"{what}: {last_error() or rc}")
    return rc


def profile(on):
    load().zb_profile(1 if on else 0)


def profile_select(names=None):
    """Time only the named kernels while profiling is on (None: every kernel)."""
    load().zb_profile_select(",".join(names).encode() if names else None)


def profile_report():
    """{kernel: (launches, total ms)} of the launches recorded since profile(True)."""
    lib = load()
    n = lib.zb_profile_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib.zb_profile_report(buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.split()
        out[name] = (int(cnt), float(ms))
    return out
