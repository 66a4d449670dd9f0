"""Backend-3 entry points of the upstream C++ core dispatch (include/mdpsnap.h part 5, device arrays)
against the upstream C++ core itself: src/C++/Core/cpuCore.cpp compiled from its sources into
oracle/_ref/libcpucore_r2.so.  The SNAP force chain SnapComputeUi -> AddWself2Ui -> SnapComputeEi ->
SnapComputeYi -> SnapComputeFi runs on the same inputs in both, and every stage is compared: U after
Ui and after the self weight, energies, Y, forces and the SoA half/half virial, within 1e-10 of the max
norm.  The upstream core's energies and forces are also checked against the oracle's (same physics
through the C route), and its total virial against the oracle's.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, relerr
from mdp_amd import lattice, snapcpp

pytestmark = pytest.mark.gpu

R2LIB = os.path.join(ROOT, "oracle", "_ref", "libcpucore_r2.so")
RCUTFAC, RFAC0 = 4.67637, 0.99363
TOL = 1e-10

P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
SIGS = {
    "SnapComputeUi": [P] * 6 + [D] * 3 + [P] * 5 + [I] * 6,
    "AddWself2Ui": [P, P, D, P, P, P, P] + [I] * 6,
    "SnapComputeEi": [P] * 12 + [I] * 10,
    "SnapComputeYi": [P] * 12 + [I] * 8,
    "SnapComputeFi": [P] * 8 + [D] * 3 + [P] * 7 + [I] * 7,
}


def _bind(lib, prefix):
    out = {}
    for name, argt in SIGS.items():
        f = getattr(lib, prefix + name)
        f.argtypes = argt
        f.restype = None
        out[name] = f
    return out


def _run_chain(fn, ptr, s, p, nelem, out_of):
    """The upstream MD chain on prepared arrays: ptr(name) gives the pointer argument of an array,
    out_of(name) its current contents as numpy."""
    N, ijnum, iu = p["n"], p["ijnum"], s.idxu_max
    res = {}
    fn["SnapComputeUi"](ptr("Ur"), ptr("Ui"), ptr("rootpq"), ptr("rij"), ptr("wj"), ptr("rad"), s.rmin0, s.rfac0,
                        s.rcutfac, ptr("idxu_block"), ptr("map"), ptr("ai"), ptr("ti"), ptr("tj"), s.twojmax, iu, N,
                        ijnum, s.switchflag, s.chemflag)
    res["U_ui"] = (out_of("Ur"), out_of("Ui"))
    fn["AddWself2Ui"](ptr("Ur"), ptr("Ui"), s.wself, ptr("idxu_block"), ptr("type"), ptr("map"), ptr("ai"),
                      s.wselfallflag, s.chemflag, iu, nelem, s.twojmax, N)
    res["U"] = (out_of("Ur"), out_of("Ui"))
    fn["SnapComputeEi"](ptr("eatom"), ptr("Ur"), ptr("Ui"), ptr("cg"), ptr("bzero"), ptr("coeff"), ptr("ilist"),
                        ptr("map"), ptr("type"), ptr("idxb"), ptr("idxcg_block"), ptr("idxu_block"), s.twojmax,
                        s.idxb_max, iu, nelem, s.ncoeffall, s.bnormflag, s.bzeroflag, s.wselfallflag,
                        s.quadraticflag, N)
    res["eatom"] = out_of("eatom")
    fn["SnapComputeYi"](ptr("Yr"), ptr("Yi"), ptr("Ur"), ptr("Ui"), ptr("cg"), ptr("coeff"), ptr("map"), ptr("type"),
                        ptr("idxz"), ptr("idxb_block"), ptr("idxu_block"), ptr("idxcg_block"), s.twojmax,
                        s.idxb_max, iu, s.idxz_max, nelem, s.ncoeffall, s.bnormflag, N)
    res["Y"] = (out_of("Yr"), out_of("Yi"))
    fn["SnapComputeFi"](ptr("f"), ptr("v"), ptr("Yr"), ptr("Yi"), ptr("rootpq"), ptr("rij"), ptr("wj"), ptr("rad"),
                        s.rmin0, s.rfac0, s.rcutfac, ptr("idxu_block"), ptr("map"), ptr("ai"), ptr("ai"), ptr("aj"),
                        ptr("ti"), ptr("tj"), s.twojmax, iu, N, N, ijnum, s.switchflag, s.chemflag)
    res["f"] = out_of("f")
    res["v"] = out_of("v")
    return res


@pytest.mark.parametrize("J,ncell,seed", [(4, 3, 1), (8, 3, 2), (6, 4, 3)])
def test_backend3_force_chain_vs_upstream_core(oracle, J, ncell, seed):
    if not os.path.exists(R2LIB):
        pytest.skip("upstream C++ core checker not built (oracle/_ref)")
    from mdp_amd import library
    from mdp_amd.devarray import DevArray
    r2 = _bind(ctypes.CDLL(R2LIB), "r2_cpu")
    gpu = _bind(library().cdll, "mdp_hip")
    x, box = lattice.crystal("bcc", ncell, 3.316, 0.1, seed)
    param = [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0]
    s0 = snapcpp.initsna(oracle, param, [0.5], [1.0])
    coeff = lattice.snap_coefficients(s0.ncoeff, seed=seed)
    s = snapcpp.initsna(oracle, param, [0.5], [1.0], coeff)
    B = np.diag(box)
    p = snapcpp.pairs_from_box(x, np.ones(len(x), np.int32), B[0], B[1], B[2], [1, 1, 1], s)
    N, nelem, iu = p["n"], s.nelements, s.idxu_max
    host = dict(Ur=np.zeros(N * iu * nelem), Ui=np.zeros(N * iu * nelem), Yr=np.zeros(N * iu * nelem),
                Yi=np.zeros(N * iu * nelem), eatom=np.zeros(N), f=np.zeros(3 * N), v=np.zeros(6 * N),
                rootpq=s.rootpqarray, rij=np.ascontiguousarray(p["rij"], np.float64), wj=s.wjelem, rad=s.radelem,
                cg=s.cglist, bzero=s.bzero, coeff=s.coeffelem, idxu_block=s.idxu_block, idxb=s.idxb,
                idxz=s.idxz, idxb_block=s.idxb_block, idxcg_block=s.idxcg_block, map=s.map,
                ai=np.ascontiguousarray(p["ai"], np.int32), aj=np.ascontiguousarray(p["aj"], np.int32),
                ti=np.ascontiguousarray(p["ti"], np.int32), tj=np.ascontiguousarray(p["tj"], np.int32),
                type=p["atomtype"], ilist=p["ilist"])
    host = {k: np.ascontiguousarray(v) for k, v in host.items()}
    outputs = ("Ur", "Ui", "Yr", "Yi", "eatom", "f", "v")
    dev = {k: DevArray.from_numpy(v, v.dtype) for k, v in host.items() if k not in outputs}
    for k in outputs:
        dev[k] = DevArray.zeros(host[k].size)
    ref = _run_chain(r2, lambda k: host[k].ctypes.data, s, p, nelem, lambda k: host[k].copy())
    print(f"backend-3 chain J={J}: {N} atoms, {p['ijnum']} pairs", flush=True)
    got = _run_chain(gpu, lambda k: dev[k].ptr, s, p, nelem, lambda k: dev[k].numpy())
    assert library().cdll.mdp_snap_last_error().decode() == ""
    for key in ("U_ui", "U", "Y"):
        for a, b in zip(got[key], ref[key]):
            assert relerr(a, b) < TOL, key
    for key in ("eatom", "f", "v"):
        assert relerr(got[key], ref[key]) < TOL, key
    # the upstream C++ core against the oracle (C route): energies and forces; total virial
    e1, f1, v1 = snapcpp.snappotential_pairs(oracle, s, p)
    assert relerr(ref["eatom"], e1) < TOL
    assert relerr(ref["f"].reshape(N, 3), f1) < TOL
    vt = ref["v"].reshape(6, N).sum(axis=1)
    assert relerr(vt, v1.sum(axis=0)) < TOL
