"""The CPU oracle against the golden vectors written by the upstream SNAP C library.

Fixtures come from tests/golden/make_golden.py (upstream cpuSnap.cpp compiled unmodified).  At
twojmax <= 6 the oracle reproduces them bit for bit; at twojmax = 8 everything except the upstream
force routine (which aborts there) is bit-exact, and forces are checked against the upstream
descriptor route.
"""
import numpy as np
import pytest

from conftest import golden_names, load_golden, relerr, sna_and_pairs
from mdp_amd import snapcpp

NAMES = golden_names()


@pytest.mark.parametrize("name", NAMES)
def test_tables(oracle, name):
    g = load_golden(name)
    s, _ = sna_and_pairs(oracle, g)
    for f in ("rootpqarray", "cglist", "idxz", "idxb", "idxu_block", "idxcg_block", "idxz_block",
              "idxb_block", "idx_max"):
        assert np.array_equal(getattr(s, f), g[f]), f


@pytest.mark.parametrize("name", NAMES)
def test_energy_force_virial(oracle, name):
    g = load_golden(name)
    s, p = sna_and_pairs(oracle, g)
    N, iu, nel = p["n"], s.idxu_max, s.nelements
    Ur, Ui = np.zeros(N * iu * nel), np.zeros(N * iu * nel)
    oracle.cpuSnapComputeUi(Ur, Ui, s.rootpqarray, p["rij"], s.wjelem, s.radelem, s.rmin0, s.rfac0,
                            s.rcutfac, s.map, p["ai"], p["ti"], p["tj"], s.twojmax, iu, N, p["ijnum"],
                            s.switchflag, s.chemflag)
    oracle.cpuAddWself2Ui(Ur, Ui, s.wself, s.idxu_block, p["atomtype"], s.map, s.wselfallflag,
                          s.chemflag, iu, nel, s.twojmax, N)
    assert np.array_equal(Ur, g["Ur"]) and np.array_equal(Ui, g["Ui"])
    e, f, v = snapcpp.snappotential_pairs(oracle, s, p)
    assert np.array_equal(e, g["eatom"])
    if str(g["force_source"]) == "cpuSnapComputeFi":
        assert np.array_equal(f, g["fatom"])
        assert np.array_equal(v, g["vatom"])
    else:
        assert relerr(f, g["fatom"]) < 1e-12
        assert relerr(v.sum(axis=0), g["vatom"][0]) < 1e-12


@pytest.mark.parametrize("name", NAMES)
def test_descriptors(oracle, name):
    g = load_golden(name)
    s, p = sna_and_pairs(oracle, g)
    bi, bd, bv, extra = snapcpp.snapdescriptors_pairs(oracle, s, p)
    assert np.array_equal(extra["blist"], g["blist"])
    assert np.array_equal(bi, g["bi"])
    assert np.array_equal(bd, g["bd"])
    assert np.array_equal(bv, g["bv"])


def test_against_compiled_reference_random(oracle, reflib):
    """Fresh seeded inputs (not in the fixtures) through both libraries, twojmax 2..8.  At twojmax=8
    the upstream force routine runs past its stack arrays (it is built without the stack protector);
    its outputs still equal the oracle's bit for bit with this compiler."""
    from mdp_amd import lattice
    for J, seed in ((2, 21), (5, 22), (6, 23), (8, 24)):
        x, box = lattice.crystal("bcc", 2, 3.3, 0.12, seed)
        t = np.ones(len(x), np.int32)
        param = [1, J, 4.67637, 0.99363, 0.0, 1, 1, 0, 0, 0, 0]
        s0 = snapcpp.initsna(oracle, param, [0.5], [1.0])
        coeff = lattice.snap_coefficients(s0.ncoeff, seed=seed)
        so = snapcpp.initsna(oracle, param, [0.5], [1.0], coeff)
        sr = snapcpp.initsna(reflib, param, [0.5], [1.0], coeff)
        B = np.diag(box)
        p = snapcpp.pairs_from_box(x, t, B[0], B[1], B[2], [1, 1, 1], so)
        a = snapcpp.snappotential_pairs(oracle, so, p)
        b = snapcpp.snappotential_pairs(reflib, sr, p)
        for u, w in zip(a, b):
            assert np.array_equal(u, w)
