"""The allocation-free Cell2AtomList (mdp_hipCell2AtomListNoAlloc) against mdp_hipCell2AtomList through
the C ABI: bit-identical outputs on two calls in a row, the second smaller than the first (the shared
scratch keeps the last call's contents), and clear errors for sizes beyond its fixed scratch and for
cell ids outside [0, ncell)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P, I = ctypes.c_void_p, ctypes.c_int


def _fns():
    from mdp_amd import library
    lib = library().cdll
    out = []
    for name in ("mdp_hipCell2AtomList", "mdp_hipCell2AtomListNoAlloc"):
        f = getattr(lib, name)
        f.argtypes, f.restype = [P] * 5 + [I] * 2, None
        out.append(f)
    return lib, out


def _run(f, lib, cl, ncell):
    from mdp_amd.devarray import DevArray
    n = len(cl)
    dcl = DevArray.from_numpy(cl, np.int32)
    da = DevArray.from_numpy(np.full(max(n, 1), -7, np.int32), np.int32)
    ds = DevArray.from_numpy(np.full(ncell + 1, -7, np.int32), np.int32)
    dn = DevArray.from_numpy(np.full(max(n, 1), -7, np.int32), np.int32)
    f(da.ptr, ds.ptr, dn.ptr, dcl.ptr, None, n, ncell)
    err = lib.mdp_snap_last_error().decode()
    return err, da.numpy()[:n], ds.numpy(), dn.numpy()[:n]


def test_no_alloc_matches_allocating_entry_point():
    lib, (alloc, noalloc) = _fns()
    rng = np.random.default_rng(3)
    for natom, ncell in ((400000, 9261), (1500, 343)):      # the second call is smaller
        cl = rng.integers(0, ncell, natom).astype(np.int32)
        print(f"Cell2AtomList no-alloc: {natom} atoms, {ncell} cells", flush=True)
        e0, a0, s0, n0 = _run(alloc, lib, cl, ncell)
        e1, a1, s1, n1 = _run(noalloc, lib, cl, ncell)
        assert e0 == "" and e1 == ""
        assert np.array_equal(a0, a1) and np.array_equal(s0, s1) and np.array_equal(n0, n1)
        order = np.argsort(cl, kind="stable")
        assert np.array_equal(a1, order) and np.array_equal(n1, cl[order])
        assert s1[0] == 0 and s1[-1] == natom


def test_no_alloc_rejects_what_it_cannot_do():
    lib, (_, noalloc) = _fns()
    noalloc(None, None, None, None, None, 1 << 29, 1000)        # beyond the fixed scratch: no work is done
    assert "scratch" in lib.mdp_snap_last_error().decode()
    noalloc(None, None, None, None, None, 10, 0)
    assert "ncell" in lib.mdp_snap_last_error().decode()
    cl = np.array([0, 3, 1, 9, 2], np.int32)                     # 9 is outside [0, 4)
    err, *_ = _run(noalloc, lib, cl, 4)
    assert "outside" in err
    # a valid call afterwards starts from a cleared flag
    err, a, s, n = _run(noalloc, lib, np.array([3, 0, 3, 1], np.int32), 4)
    assert err == "" and a.tolist() == [1, 3, 0, 2] and s.tolist() == [0, 1, 2, 2, 4]
