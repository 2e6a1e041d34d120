"""cans1d_mlw_m_noalloc (the entry point that allocates no device memory)
gives the same bits as cans1d_mlw_m, on consecutive calls that share its
static workspace, the second with a smaller grid; sizes beyond the workspace
are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import assert_bitwise, smooth_state_1d, stretched_grid, to_dev, to_host

pytestmark = pytest.mark.gpu

F1 = ("ro", "pr", "vx", "vy", "by")


def _call(entry, t, dx, dxm, dt, qav, gm):
    import torch
    from utrip_amd import cans
    ix = t["ro"].shape[0]
    ptrs = [C.c_void_p(t[k].data_ptr()) for k in ("ro", "pr", "vx", "vy", "by", "bx", "bxm")]
    rc = getattr(cans.lib(), entry)(*ptrs, C.c_double(dt), C.c_double(qav), C.c_double(gm),
                                    C.c_void_p(dx.data_ptr()), C.c_void_p(dxm.data_ptr()), C.c_int(ix),
                                    C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, cans.lib().cans_gpu_last_error().decode()


def test_noalloc_matches_allocating_entry_point(dev, oracle):
    import torch
    for ix, seed, dt in ((4099, 7, 1.1e-3), (600, 8, 2.3e-3)):
        print("CASE noalloc mlw1d ix", ix, flush=True)
        s = smooth_state_1d(ix, seed=seed)
        _, dx, dxm = stretched_grid(ix, seed=seed)
        a = {k: to_dev(v, dev) for k, v in s.items()}
        b = {k: to_dev(v, dev) for k, v in s.items()}
        ddx, ddxm = to_dev(dx, dev), to_dev(dxm, dev)
        _call("cans1d_mlw_m", a, ddx, ddxm, dt, 3.0, 5.0 / 3.0)
        _call("cans1d_mlw_m_noalloc", b, ddx, ddxm, dt, 3.0, 5.0 / 3.0)
        torch.cuda.synchronize()
        ref = {k: v.copy() for k, v in s.items()}
        oracle.mlw_m_1d(ref["ro"], ref["pr"], ref["vx"], ref["vy"], ref["by"], ref["bx"], ref["bxm"],
                        dt, 3.0, 5.0 / 3.0, dx, dxm)
        for k in F1:
            assert_bitwise(to_host(b[k]), to_host(a[k]), k)
            assert_bitwise(to_host(b[k]), ref[k], k)


def test_noalloc_refuses_oversize(dev):
    import torch
    from utrip_amd import cans
    ix = cans.NOALLOC_MAX_IX + 1
    t = {k: torch.zeros(ix, dtype=torch.float64, device=dev) for k in ("ro", "pr", "vx", "vy", "by", "bx", "bxm")}
    d = torch.ones(ix, dtype=torch.float64, device=dev)
    with pytest.raises(cans.CansError, match="static workspace"):
        cans.mlw_m_1d_noalloc(t["ro"], t["pr"], t["vx"], t["vy"], t["by"], t["bx"], t["bxm"], 1e-3, 3.0, 1.4,
                              d, d)
