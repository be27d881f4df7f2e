"""zb_sh_coefficients_noalloc (the entry point that allocates nothing: its workspace is the
library's static device array) against zb_sh_coefficients with a workspace of the caller's, two
calls in a row on the same device, the second with fewer items and another bandwidth's band, so
that what the first call left in the array cannot pass for the second's; past the array's limit
the call fails with the limit in the message."""
import numpy as np
import pytest

from tests import cases
from tests.conftest import NI, gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def test_noalloc_coefficients_match_the_workspace_entry_point():
    import torch
    from zebsdsolve_amd import _native as nat, engine
    from zebsdsolve_amd.engine import stream_ptr
    from zebsdsolve_amd.geometry import Geometry
    print("no-alloc coefficients", flush=True)
    e = engine.Engine()
    e.settings.apply(static_background="none", dynamic_background="none")
    e.geo = Geometry(pc=cases.PC, sample_tilt_deg=70.0)
    e.add_phase(NI)
    n0 = 300
    e.set_patterns(np.zeros((1, n0, 60, 60), np.float32))
    _, (_, _, host) = e.baked(0)
    side = e.phases[0].master.side
    pats, _ = cases.simulate(host, side, n0, seed=23, counts=400.0)
    lib = nat.load()
    raw = torch.from_numpy(pats).cuda()
    pn = torch.empty_like(raw)
    mask = (nat.ctypes.c_float * 3)(*e.mask_tuple())
    nat.check(lib.zb_normalize(raw.data_ptr(), 2, n0, 60, 60, mask, pn.data_ptr(), stream_ptr()), "normalize")
    t = e.positions(np.arange(n0))
    items = torch.from_numpy(np.concatenate([t, np.arange(n0, dtype=np.float32)[:, None]], 1)).cuda()
    fr = e.frame()
    for L, n in ((64, n0), (32, 77)):
        band = e.band(L)
        pairs = band.host.pairs
        a1 = torch.full((n, pairs, 2), np.nan, dtype=torch.float32, device="cuda")
        s1 = torch.full((n, 4), np.nan, dtype=torch.float32, device="cuda")
        ws = torch.empty((lib.zb_sh_workspace(band.c, n),), dtype=torch.uint8, device="cuda")
        nat.check(lib.zb_sh_coefficients(band.c, fr, pn.data_ptr(), items.data_ptr(), n, 3, ws.data_ptr(), ws.numel(), a1.data_ptr(),
                                         s1.data_ptr(), stream_ptr()), "coefficients")
        a2 = torch.full((n, pairs, 2), np.nan, dtype=torch.float32, device="cuda")
        s2 = torch.full((n, 4), np.nan, dtype=torch.float32, device="cuda")
        nat.check(lib.zb_sh_coefficients_noalloc(band.c, fr, pn.data_ptr(), items.data_ptr(), n, 3, a2.data_ptr(), s2.data_ptr(),
                                                 stream_ptr()), "coefficients noalloc")
        torch.cuda.synchronize()
        # the same kernels on the same inputs: bit for bit
        assert torch.equal(a1, a2) and torch.equal(s1, s2), L
        assert torch.isfinite(a2).all()
    # past the static array: a clear error, nothing launched
    band = e.band(64)
    big = 1024
    rc = lib.zb_sh_coefficients_noalloc(band.c, fr, pn.data_ptr(), items.data_ptr(), big, 3, a2.data_ptr(), s2.data_ptr(), stream_ptr())
    assert rc == -2 and "at most 1023 items" in nat.last_error()
