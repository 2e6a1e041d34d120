"""Host-array mirrors of the Fortran drop-in layer in resident mode: a sync
copies back only the mirrors a routine changed (never a read-only input the
program has since changed on the host, never a released array), and leaving
resident mode hands the results to the host first."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BD_CNS = 5


def _lib():
    from utrip_amd import cans
    L = cans.lib()
    L.cans_gpu_map.restype = C.c_void_p
    L.cans_gpu_map.argtypes = [C.c_void_p, C.c_size_t, C.c_int]
    return L


def _ok(rc):
    from utrip_amd import cans
    assert rc == 0, cans.lib().cans_gpu_last_error().decode()


def _touch(L, a, q0):
    """A routine call on the mirror of a: bdcnsx low side, margin 2 (the
    first two entries become q0 on the device), then unmap with copy-out."""
    n = a.size
    d = L.cans_gpu_map(C.c_void_p(a.ctypes.data), n * 8, 1)
    assert d
    _ok(L.cans1d_bnd(C.c_int(BD_CNS), C.c_int(0), C.c_int(2), C.c_void_p(d), C.c_double(q0), C.c_void_p(0),
                     C.c_int(n), C.c_void_p(0)))
    _ok(L.cans_gpu_unmap(C.c_void_p(a.ctypes.data), 1))


def _read(L, a):
    """A read-only use of a (mapped with copy-in, unmapped without copy-out)."""
    d = L.cans_gpu_map(C.c_void_p(a.ctypes.data), a.size * 8, 1)
    assert d
    _ok(L.cans_gpu_unmap(C.c_void_p(a.ctypes.data), 0))


def test_resident_sync_copies_back_changed_mirrors_only(dev):
    print("CASE host mirrors: dirty tracking, forget, leaving resident mode", flush=True)
    L = _lib()
    n = 16
    inp = np.arange(n, dtype=np.float64)       # a read-only input
    out = np.arange(n, dtype=np.float64)       # an array a routine changes
    gone = np.arange(n, dtype=np.float64)      # changed, then released by the program
    _ok(L.cans_gpu_set_resident(1))
    try:
        _read(L, inp)
        _touch(L, out, 7.0)
        _touch(L, gone, 9.0)
        assert out[0] == 0.0                    # resident: nothing copied back yet
        inp[:] = 42.0                           # the program changes its input on the host
        _ok(L.cans_gpu_forget(C.c_void_p(gone.ctypes.data)))
        gone[:] = 5.0                           # memory the program reuses after the release
        _ok(L.cans_gpu_sync_host())
        assert (out[:2] == 7.0).all() and (out[2:] == np.arange(2, n)).all()
        assert (inp == 42.0).all()
        assert (gone == 5.0).all()
        # a second sync copies nothing (out is clean): host edits survive
        out[5] = -1.0
        _ok(L.cans_gpu_sync_host())
        assert out[5] == -1.0
        # a changed mirror is handed back when resident mode ends
        _ok(L.cans_gpu_invalidate(C.c_void_p(out.ctypes.data)))
        _touch(L, out, 3.0)
        assert out[0] == 7.0
    finally:
        _ok(L.cans_gpu_set_resident(0))
    assert out[0] == 3.0 and out[1] == 3.0 and out[5] == -1.0
    for a in (inp, out):
        _ok(L.cans_gpu_forget(C.c_void_p(a.ctypes.data)))


def test_strict_mode_copies_through(dev):
    L = _lib()
    a = np.arange(8, dtype=np.float64)
    _touch(L, a, 2.5)                          # strict: the call returns with the host updated
    assert (a[:2] == 2.5).all() and (a[2:] == np.arange(2, 8)).all()
    _ok(L.cans_gpu_forget(C.c_void_p(a.ctypes.data)))
