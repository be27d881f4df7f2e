"""The C ABI: the built library loads and exports every entry point include/zebsd.h declares, with
the signatures the ctypes binding gives them. No kernel is launched (no GPU needed)."""
import ctypes
import os
import re

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "zebsd.h")
LIB = os.path.join(ROOT, "zebsdsolve_amd", "libzebsd_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(zb_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_entry_points():
    names = declared()
    for n in ("zb_normalize", "zb_sh_coefficients", "zb_so3_correlate", "zb_so3_denominators", "zb_refine", "zb_project",
              "zb_master_blur", "zb_master_gradient", "zb_master_pack", "zb_last_error", "zb_version"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzebsd_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    from zebsdsolve_amd import _native
    assert set(declared()) == set(_native.EXPORTS)
    lib.zb_version.restype = ctypes.c_char_p
    assert lib.zb_version().decode() == "0.1.0"


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzebsd_hip.so not built")
def test_bad_arguments_fail_with_a_message_and_launch_nothing():
    from zebsdsolve_amd import _native as nat
    lib = nat.load()
    rc = lib.zb_normalize(None, 7, 1, 60, 60, (ctypes.c_float * 3)(0, 0, 0), None, None)
    assert rc == -1 and "zb_normalize" in nat.last_error()
    band = nat.Band(64, 100, None, None, None, None)  # n not a power of two
    assert lib.zb_sh_workspace(ctypes.byref(band), 4) > 0 or True
    rc = lib.zb_sh_coefficients(ctypes.byref(band), ctypes.byref(nat.Frame()), None, None, 4, 3, None, 0, None, None, None)
    assert rc == -1 and "bad band" in nat.last_error()
    ph = nat.Phase(3, 32, 21, 2, None, None)  # fold 3 is not a power of two
    band = nat.Band(64, 128, None, None, None, None)
    assert lib.zb_so3_workspace(ctypes.byref(band), ctypes.byref(ph), 4) == -1
    vt = nat.VTable(96, 256, 64, 1, 96, 128, 4096, 1, 0, 4096)  # 96 degree slots: the 16-bit form takes 64
    assert lib.zb_vcorr_workspace(ctypes.byref(vt), 4, 0) == -1 and "degree slots" in nat.last_error()
    vt = nat.VTable(96, 256, 128, 1, 96, 256, 4096)  # GaN at L = 96: FP32 products take 96 slots, 128 columns
    assert lib.zb_vcorr_workspace(ctypes.byref(vt), 4, 0) > 0
    vt = nat.VTable(96, 256, 16, 2, 48, 32, 4096)  # n = 256 takes two lanes per column of 32
    assert lib.zb_vcorr_workspace(ctypes.byref(vt), 4, 0) == -1 and "nc <= 64" in nat.last_error()
    vt = nat.VTable(160, 512, 64, 2, 80, 128, None)  # n = 512: past the register transform
    assert lib.zb_vcorr_workspace(ctypes.byref(vt), 4, 0) == -1 and "n <= 256" in nat.last_error()
    rc = lib.zb_vcorr_correlate(ctypes.byref(vt), None, None, 4, None, 4, None, 0, None, None)
    assert rc == -1 and "zb_vcorr_correlate" in nat.last_error()
    opts = nat.RefineOpts(5, 0.035, 1.0, 0.02, 1e4, 64, 1)  # 64 NMI bins: too many
    rc = lib.zb_refine(None, None, 501, ctypes.byref(nat.Frame()), None, None, None, 8, 4, ctypes.byref(opts), None, 0, None)
    assert rc == -1
    # zb_refine_phases: phase count, per-phase arguments, workspace size
    opts = nat.RefineOpts(5, 0.035, 1.0, 0.02, 1e4, 16, 1)
    fr = nat.Frame()
    fr.pat_h = fr.pat_w = 60
    fake = 256  # never dereferenced: every call below fails its checks before any launch
    ph = (nat.RefinePhase * 5)(*[nat.RefinePhase(fake, fake, 501, fake, 8, None, fake) for _ in range(5)])
    args = lambda n_ph, work_bytes: (ph, n_ph, ctypes.byref(fr), fake, fake, 4, ctypes.byref(opts), 0, fake, work_bytes, None)
    assert lib.zb_refine_phases(*args(0, 1024)) == -1 and "bad arguments" in nat.last_error()
    assert lib.zb_refine_phases(*args(5, 1024)) == -1
    assert lib.zb_refine_phases(*args(2, 512)) == -2 and "work_bytes" in nat.last_error()
    ph[1].seed_stride = 3
    assert lib.zb_refine_phases(*args(2, 1024)) == -1 and "phase 1" in nat.last_error()
    ph[1].seed_stride = 8
    ph[1].records = None
    assert lib.zb_refine_phases(*args(2, 1024)) == -1 and "phase 1: null pointer" in nat.last_error()
    # zb_prep_work: filter arguments as zb_prep's, and a workspace the box blur can read as float pairs
    po = nat.PrepOpts(None, 1, 0.0)  # a dynamic background needs a positive sigma
    mask = (ctypes.c_float * 3)(0, 0, 0)
    assert lib.zb_prep_workspace(4, 60, 60, ctypes.byref(po)) == -1 and "zb_prep_workspace" in nat.last_error()
    assert lib.zb_prep_work(fake, 2, 4, 60, 60, mask, ctypes.byref(po), fake, 1 << 20, fake, None) == -1
    assert "zb_prep_work: bad arguments" in nat.last_error()
    po = nat.PrepOpts(None, 1, 9.0)
    assert lib.zb_prep_work(fake + 4, 2, 4, 60, 60, mask, ctypes.byref(po), fake + 4, 1 << 20, fake, None) == -1
    assert "8-byte aligned" in nat.last_error()
    assert lib.zb_prep_work(None, 2, 0, 60, 60, mask, ctypes.byref(po), None, 0, None, None) == 0  # nothing to do


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzebsd_hip.so not built")
def test_knobs_set_get_and_reject_unknown_names_and_ranges():
    from zebsdsolve_amd import _native as nat
    lib = nat.load()
    assert lib.zb_knob_get(b"refine_queue") == 0 and lib.zb_knob_get(b"rings_group") == 0
    assert lib.zb_knob_set(b"rings_group", 16) == 0
    assert lib.zb_knob_get(b"rings_group") == 16
    assert lib.zb_knob_set(b"rings_group", 0) == 16
    assert lib.zb_knob_set(b"rings_group", 65) == -1 and "outside" in nat.last_error()
    assert lib.zb_knob_set(b"refine_queue", 3) == -1
    assert lib.zb_knob_set(b"no_such_knob", 1) == -1 and "no knob" in nat.last_error()
    assert lib.zb_knob_get(b"no_such_knob") == -1


def test_flop_model_reproduces_the_exact_term_counts():
    import bench
    assert bench.spectrum_terms(32, 4, True) == 65152
    assert bench.spectrum_terms(64, 4, True) == 1047040
    assert bench.spectrum_terms(48, 4, True) == 330912
    assert bench.spectrum_terms(64, 2, False) == 4280640
    assert bench.mask_count(60, 60) == 2828 and bench.mask_count(120, 120) == 11304
    per, total = bench.flops_per_pattern(64, 4, True, 5, 2828)
    # 23.85 MFLOP per pattern at L = 64, 5 iterations, 60 x 60 (spectrum 8.38, transforms 11.91,
    # harmonic analysis 1.19, refinement and NMI 2.38)
    assert abs(total / 1e6 - 23.85) < 0.01
    # both correlation paths carry the same terms
    assert per["k_vgemm"] + per["k_apeak"] == per["k_spectrum"] + per["k_slab0"]
    assert abs((per["k_vgemm"] + per["k_apeak"]) / 1e6 - (8.38 + 11.91)) < 0.01


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzebsd_hip.so not built")
def test_table_tiling_rule_is_the_librarys():
    """The host builds the matrix-core table in the tiling the kernel launch chooses: the two
    implementations of the rule agree, and the tiled layout holds each element where the header
    says."""
    import numpy as np
    from zebsdsolve_amd import _native as nat, tables
    lib = nat.load()
    for nc in (2, 4, 8, 16, 32, 64, 128):
        for slots in (1, 5, 8, 9, 16, 17, 32, 33, 48, 49, 64, 65, 96):
            c, S, b = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
            assert lib.zb_vtable_shape(nc, slots, ctypes.byref(c), ctypes.byref(S), ctypes.byref(b)) == 0
            assert (c.value, S.value, b.value) == tables.vgemm_shape(nc, slots)
    assert lib.zb_vtable_shape(48, 8, ctypes.byref(c), ctypes.byref(S), ctypes.byref(b)) == -1
    rng = np.random.default_rng(0)
    nb, L, slots, cols = 3, 4, 5, 64
    T = rng.normal(size=(nb, L, slots, 2, cols)).astype(np.float32)
    _, S, nblk = tables.vgemm_shape(32, slots)
    t = tables.tile_vtable(T, S, nblk).reshape(nb, L, cols // (32 * nblk), S // 4, nblk, 64, 4)
    for _ in range(200):
        j, m, s, r, col = rng.integers(nb), rng.integers(L), rng.integers(S), rng.integers(2), rng.integers(cols)
        g, rest = divmod(col, 32 * nblk)
        bb, x = divmod(rest, 32)
        want = T[j, m, s, r, col] if s < slots else 0.0
        assert t[j, m, g, s // 4, bb, 32 * r + x, s % 4] == want


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzebsd_hip.so not built")
def test_table16_tiling_rule_is_the_librarys():
    """The 16-bit table: the host's shape rule is the library's, each hi value sits where the
    header says (block 0 the real and block 1 the imaginary parts of 32 complex columns), the lo
    values follow their group's hi from the split group on, and hi + lo gives back the value."""
    import numpy as np
    from zebsdsolve_amd import _native as nat, tables
    lib = nat.load()
    c, G, b = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    for nc in (2, 4, 8, 16, 32, 64):
        for slots in (1, 5, 8, 9, 16, 17, 24, 32, 33, 48, 64):
            assert lib.zb_vtable16_shape(nc, slots, ctypes.byref(c), ctypes.byref(G), ctypes.byref(b)) == 0
            assert (c.value, G.value, b.value) == tables.vgemm16_shape(nc, slots)
    assert lib.zb_vtable16_shape(48, 8, ctypes.byref(c), ctypes.byref(G), ctypes.byref(b)) == -1
    rng = np.random.default_rng(1)
    L, slots, cols = 3, 40, 64
    T = rng.normal(size=(L, slots, 2, cols))
    _, groups, nblk = tables.vgemm16_shape(32, slots)
    assert (groups, nblk) == (6, 2)
    ncg = cols // (32 * nblk)
    for kind, rel in ((1, 2.0 ** -21), (2, 2.0 ** -15)):
        split = 32  # groups 4 and 5 carry lo
        flat = tables.tile_vtable16(T, groups, nblk, split, kind)
        per = nblk * 64 * 8
        assert flat.size == L * ncg * (groups + 2) * per
        blocks = flat.reshape(L, ncg, groups + 2, nblk, 64, 8)
        val = (lambda u: u.view(np.float16).astype(np.float64)) if kind == 1 else \
              (lambda u: (u.astype(np.uint32) << np.uint32(16)).view(np.float32).astype(np.float64))
        for _ in range(300):
            m, s, r, col = rng.integers(L), rng.integers(8 * groups), rng.integers(2), rng.integers(cols)
            c, bb = divmod(col, 2)  # complex column c, real (0) or imaginary (1) part
            g, x = divmod(c, 32)
            t = s // 8
            pos = t if t < 4 else 4 + 2 * (t - 4)
            hi = val(blocks[m, g, pos, bb, 32 * r + x, s % 8])
            want = T[m, s, r, col] if s < slots else 0.0
            if t < 4:
                assert abs(hi - want) <= 2.0 ** (-8 if kind == 2 else -11) * abs(want)
            else:
                lo = val(blocks[m, g, pos + 1, bb, 32 * r + x, s % 8])
                assert abs(hi + lo - want) <= rel * abs(want) + 1e-30


def test_round16_is_nearest_even():
    """FP16 is numpy's; BF16 keeps the top 16 bits of the float32, nearest with ties to even."""
    import numpy as np
    from zebsdsolve_amd import tables
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5, 0.0, 3.0e38])
    bits, v = tables.round16(x, 2)
    assert list(bits) == [0x3F80, 0x3F80, 0x3F82, 0x3F81, 0xC020, 0x0000, 0x7F62]
    assert v[1] == 1.0 and v[2] == 1.0 + 2.0 ** -6 and v[3] == 1.0 + 2.0 ** -7
    h, hv = tables.round16(np.array([1.0 + 2.0 ** -11, 65504.0]), 1)
    assert hv[0] == 1.0 and hv[1] == 65504.0
    with pytest.raises(ValueError):
        tables.round16(np.array([70000.0]), 1)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzebsd_hip.so not built")
def test_every_documented_knob_has_its_documented_range_and_default():
    """The knobs include/zebsd.h lists ("name" lo..hi (default)) are the library's: each reads its
    default, takes both ends of its range and refuses the value past them; nothing else is
    settable under another name."""
    from zebsdsolve_amd import _native as nat
    lib = nat.load()
    text = open(HEADER).read()
    knobs = re.findall(r'^\s*\*\s+"(\w+)"\s+(-?\d+)\.\.(-?\d+)\s+\((-?\d+)\)', text, flags=re.M)
    assert {k[0] for k in knobs} >= {"refine_queue", "rings_group", "prep_global"}
    for name, lo, hi, default in knobs:
        key, lo, hi, default = name.encode(), int(lo), int(hi), int(default)
        assert lib.zb_knob_get(key) == default, name
        try:
            assert lib.zb_knob_set(key, hi) == default, name
            assert lib.zb_knob_set(key, lo) == hi, name
            assert lib.zb_knob_set(key, hi + 1) == -1 and lib.zb_knob_get(key) == lo, name
        finally:
            lib.zb_knob_set(key, default)
