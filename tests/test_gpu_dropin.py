"""GPU drop-in parity: the 23 upstream entry points of libmdpsnap against the golden vectors of the
upstream SNAP library and against the CPU oracle on the same inputs.

Tolerances: integer outputs (tables, pair lists) bit-exact; FP64 outputs within 1e-12 relative
(max-norm over the array) for U, Z, B, Y, E and 1e-10 for forces, virials and dB/dr.
"""
import numpy as np
import pytest

from conftest import golden_names, load_golden, relerr, sna_and_pairs
from mdp_amd import lattice, snapcpp

pytestmark = pytest.mark.gpu

NAMES = golden_names()
TOL_U = 1e-12
TOL_F = 1e-10


@pytest.fixture(scope="module")
def gpu():
    import mdp_amd
    return mdp_amd.library()


def _chain(lib, s, p):
    N, iu, nel = p["n"], s.idxu_max, s.nelements
    Ur, Ui = np.zeros(N * iu * nel), np.zeros(N * iu * nel)
    lib.cpuSnapComputeUi(Ur, Ui, s.rootpqarray, p["rij"], s.wjelem, s.radelem, s.rmin0, s.rfac0, s.rcutfac, s.map,
                         p["ai"], p["ti"], p["tj"], s.twojmax, iu, N, p["ijnum"], s.switchflag, s.chemflag)
    lib.cpuAddWself2Ui(Ur, Ui, s.wself, s.idxu_block, p["atomtype"], s.map, s.wselfallflag, s.chemflag, iu, nel,
                       s.twojmax, N)
    Yr, Yi = np.zeros(N * iu * nel), np.zeros(N * iu * nel)
    lib.cpuSnapComputeYi(Yr, Yi, Ur, Ui, s.cglist, s.coeffelem, s.map, p["atomtype"], s.idxz, s.idxb_block,
                         s.idxu_block, s.idxcg_block, s.twojmax, s.idxb_max, iu, s.idxz_max, nel, s.ncoeffall,
                         s.bnormflag, s.chemflag, N)
    return Ur, Ui, Yr, Yi


@pytest.mark.parametrize("name", NAMES)
def test_tables_and_counts(gpu, name):
    g = load_golden(name)
    s, _ = sna_and_pairs(gpu, g)
    for f in ("idxz", "idxb", "idxu_block", "idxcg_block", "idxz_block", "idxb_block", "idx_max", "rootpqarray"):
        assert np.array_equal(getattr(s, f), g[f]), f
    assert relerr(s.cglist, g["cglist"]) < 1e-15


@pytest.mark.parametrize("name", NAMES)
def test_force_path_vs_golden(gpu, name):
    import mdp_amd
    g = load_golden(name)
    s, p = sna_and_pairs(gpu, g)
    Ur, Ui, Yr, Yi = _chain(gpu, s, p)
    mdp_amd.check("force chain")
    assert relerr(Ur, g["Ur"]) < TOL_U and relerr(Ui, g["Ui"]) < TOL_U
    assert relerr(Yr, g["Yr"]) < TOL_U and relerr(Yi, g["Yi"]) < TOL_U
    e, f, v = snapcpp.snappotential_pairs(gpu, s, p)
    mdp_amd.check("snappotential")
    assert relerr(e, g["eatom"]) < TOL_U
    assert relerr(f, g["fatom"]) < TOL_F
    if str(g["force_source"]) == "cpuSnapComputeFi":
        assert relerr(v, g["vatom"]) < TOL_F
    else:
        assert relerr(v.sum(axis=0), g["vatom"][0]) < TOL_F


@pytest.mark.parametrize("name", NAMES)
def test_descriptor_path_vs_golden(gpu, oracle, name):
    import mdp_amd
    g = load_golden(name)
    s, p = sna_and_pairs(gpu, g)
    bi, bd, bv, ex = snapcpp.snapdescriptors_pairs(gpu, s, p)
    mdp_amd.check("snapdescriptors")
    assert relerr(ex["blist"], g["blist"]) < TOL_U
    assert relerr(bi, g["bi"]) < TOL_U
    assert relerr(bd, g["bd"]) < TOL_F
    assert relerr(bv, g["bv"]) < TOL_F
    so, _ = sna_and_pairs(oracle, g)
    _, _, _, eo = snapcpp.snapdescriptors_pairs(oracle, so, p)
    for k in ("Ur", "Ui", "zr", "zi", "dblist"):
        assert relerr(ex[k], eo[k]) < TOL_F, k


def _rand_case(J, seed, lib, ncell=2, sigma=0.12):
    x, box = lattice.crystal("bcc", ncell, 3.3, sigma, seed)
    t = np.ones(len(x), np.int32)
    param = [1, J, 4.67637, 0.99363, 0.0, 1, 1, 0, 0, 0, 0]
    s0 = snapcpp.initsna(lib, param, [0.5], [1.0])
    coeff = lattice.snap_coefficients(s0.ncoeff, seed=seed)
    s = snapcpp.initsna(lib, param, [0.5], [1.0], coeff)
    B = np.diag(box)
    return s, snapcpp.pairs_from_box(x, t, B[0], B[1], B[2], [1, 1, 1], s), coeff, param


@pytest.mark.parametrize("J", [2, 4, 6, 8])
def test_unsorted_pairs_and_accumulate(gpu, oracle, J):
    """Pairs in shuffled order; Utot/fatom/vatom/eatom pre-filled (upstream adds into them)."""
    s, p, coeff, param = _rand_case(J, 40 + J, gpu)
    so = snapcpp.initsna(oracle, param, [0.5], [1.0], coeff)
    perm = np.random.default_rng(J).permutation(p["ijnum"])
    q = dict(p)
    for k in ("ai", "aj", "ti", "tj"):
        q[k] = np.ascontiguousarray(p[k][perm])
    q["rij"] = np.ascontiguousarray(p["rij"].reshape(-1, 3)[perm].ravel())
    N, iu = p["n"], s.idxu_max
    rng = np.random.default_rng(7)
    base = rng.normal(size=N * iu)
    out = []
    for lib, ss in ((gpu, s), (oracle, so)):
        Ur, Ui = base.copy(), base[::-1].copy()
        lib.cpuSnapComputeUi(Ur, Ui, ss.rootpqarray, q["rij"], ss.wjelem, ss.radelem, ss.rmin0, ss.rfac0,
                             ss.rcutfac, ss.map, q["ai"], q["ti"], q["tj"], J, iu, N, q["ijnum"], 1, 0)
        f = np.full(3 * N, 0.25)
        v = np.full(6 * N, -0.5)
        lib.cpuSnapComputeFi(f, v, base, base[::-1].copy(), ss.rootpqarray, q["rij"], ss.wjelem, ss.radelem,
                             ss.rmin0, ss.rfac0, ss.rcutfac, ss.map, q["ai"], q["ai"], q["aj"], q["ti"], q["tj"],
                             J, iu, N, N, q["ijnum"], 1, 0)
        e = np.full(N, 3.0)
        lib.cpuSnapComputeEi(e, Ur, Ui, ss.cglist, ss.bzero, ss.coeffelem, p["ilist"], ss.map, p["atomtype"],
                             ss.idxb, ss.idxcg_block, ss.idxu_block, J, ss.idxb_max, iu, 1, ss.ncoeffall, 0, 1, 0, 0,
                             N)
        out.append((Ur, Ui, f, v, e))
    for a, b in zip(*out):
        assert relerr(a, b) < TOL_F


def test_pair_list_builders(gpu, oracle):
    """cpuNeighPairList/cpuNeighPairs on a CSR neighbour list with ghosts: bit-exact with the oracle."""
    from mdp_amd import neighbors
    x, box = lattice.crystal("bcc", 3, 3.316, 0.15, 3)
    n = len(x)
    B = np.diag(box)
    y, alist, nl, nn = neighbors.fullneighborlist(x, B[0], B[1], B[2], [1, 1, 1], 4.97637)
    t = np.ones(len(y), np.int32)
    rc = 4.67637
    rcutsq = np.full(4, rc * rc)
    ilist = np.arange(n, dtype=np.int32)
    yy = np.ascontiguousarray(y.ravel())
    res = []
    for lib in (gpu, oracle):
        pn = np.zeros(n, np.int32)
        pl = np.zeros(len(nl), np.int32)
        lib.cpuNeighPairList(pn, pl, yy, rcutsq, ilist, nl, nn, t, alist, n, 3, 1)
        ps = np.concatenate([[0], np.cumsum(pn)]).astype(np.int32)
        P = int(ps[-1])
        arrs = [np.zeros(3 * P)] + [np.zeros(P, np.int32) for _ in range(5)]
        lib.cpuNeighPairs(arrs[0], yy, *arrs[1:], pl, ps, ilist, t, alist, n, 3)
        res.append([pn, pl[:P]] + arrs)
    for a, b in zip(*res):
        assert np.array_equal(a, b)


def test_unsupported_twojmax_reports_error(gpu):
    import mdp_amd
    J = 5
    iu = gpu.idxucount(J)
    Ur, Ui = np.zeros(iu), np.zeros(iu)
    rij = np.array([1.0, 2.0, 0.5])
    one = np.ones(1, np.int32)
    gpu.cpuSnapComputeUi(Ur, Ui, np.zeros((J + 2) ** 2), rij, np.ones(2), np.full(2, 0.5), 0.0, 0.99, 4.67,
                         np.zeros(2, np.int32), np.zeros(1, np.int32), one, one, J, iu, 1, 1, 1, 0)
    assert "twojmax=5" in mdp_amd.last_error()
    assert not Ur.any()
