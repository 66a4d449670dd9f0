"""ZBL reference pair potential on the GPU (include/mdpsnap.h part 6 and mdp_md_set_zbl).

* The backend-3 pair gradient and pair tallies against the upstream generated code's golden vectors
  (tests/golden/zbl_ta.npz) and against the oracle, within 1e-12 of the max norm (device exp/tanh against
  the host's libm).
* The MD engine with SNAP + ZBL (the Ta application's pair, snapTA06A.jl:62-66) against the oracle's SNAP
  plus the oracle's ZBL on a compressed BCC lattice where the ZBL core matters: energies, forces and
  virials within 1e-10; central finite differences of the engine's total energy against its forces; and
  NVE energy conservation with both terms.
"""
import ctypes
import os

import numpy as np
import pytest

import zbl_oracle
from conftest import GOLDEN, relerr
from mdp_amd import lattice, snapcpp

pytestmark = pytest.mark.gpu

RCUTFAC, RFAC0, MASS = 4.67637, 0.99363, 180.95
VP, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


def _lib():
    import mdp_amd
    c = mdp_amd.library().cdll
    c.mdp_hipPairaGradient.argtypes = [VP] * 12 + [I] * 7
    c.mdp_hipFullNeighPairDecomposition.argtypes = [VP] * 5 + [I, I]
    c.mdp_hipVirialPairTally.argtypes = [VP] * 3 + [D, VP, I, I, I]
    for f in ("mdp_hipPairaGradient", "mdp_hipFullNeighPairDecomposition", "mdp_hipVirialPairTally"):
        getattr(c, f).restype = None
    return c


def _err(c):
    return c.mdp_snap_last_error().decode()


def test_pair_gradient_and_tallies_vs_golden(oracle):
    from mdp_amd.devarray import DevArray
    c = _lib()
    gd = dict(np.load(os.path.join(GOLDEN, "zbl_ta.npz"), allow_pickle=False))
    ng = len(gd["u"])
    dx = DevArray.from_numpy(gd["xij"].ravel(), np.float64)
    du, dux = DevArray.zeros(ng), DevArray.zeros(3 * ng)
    dmu = DevArray.from_numpy(gd["mu"], np.float64)
    print(f"zbl pair gradient, {ng} pairs", flush=True)
    c.mdp_hipPairaGradient(du.ptr, dux.ptr, dx.ptr, None, None, None, None, None, None, dmu.ptr, None, None, 3, 0, 3,
                           0, 0, ng, 1)
    assert _err(c) == ""
    assert relerr(du.numpy(), gd["u"]) < 1e-12
    assert relerr(dux.numpy().reshape(-1, 3), gd["u_xij"]) < 1e-12
    # the whole route on the golden system: gradient, full-list decomposition, virial tally
    P, n = len(gd["ai"]), len(gd["x"])
    pij = DevArray.from_numpy(gd["pij"].ravel(), np.float64)
    pu, pux = DevArray.zeros(P), DevArray.zeros(3 * P)
    ai = DevArray.from_numpy(gd["ai"], np.int32)
    c.mdp_hipPairaGradient(pu.ptr, pux.ptr, pij.ptr, None, None, None, None, None, None, dmu.ptr, None, None, 3, 0, 3,
                           0, 0, P, 1)
    e, f, w = DevArray.zeros(n), DevArray.zeros(3 * n), DevArray.zeros(6 * n)
    c.mdp_hipFullNeighPairDecomposition(e.ptr, f.ptr, pu.ptr, pux.ptr, ai.ptr, P, 3)
    c.mdp_hipVirialPairTally(w.ptr, pux.ptr, pij.ptr, -0.5, ai.ptr, 3, n, P)
    assert _err(c) == ""
    assert relerr(e.numpy(), gd["e"]) < 1e-12
    assert relerr(-f.numpy().reshape(n, 3), gd["f"]) < 1e-12
    assert relerr(w.numpy().reshape(6, n), gd["vsoa"]) < 1e-12
    # the oracle on the same inputs agrees as well (accumulating outputs: start from nonzero contents)
    e0 = np.linspace(-1, 1, n)
    e2 = DevArray.from_numpy(e0, np.float64)
    c.mdp_hipFullNeighPairDecomposition(e2.ptr, f.ptr, pu.ptr, pux.ptr, ai.ptr, P, 3)
    assert relerr(e2.numpy() - e0, gd["e"]) < 1e-12


def test_pair_gradient_rejects_bad_arguments():
    c = _lib()
    c.mdp_hipPairaGradient(None, None, None, None, None, None, None, None, None, None, None, None, 3, 0, 3, 0, 0, 4, 2)
    assert "potnum" in _err(c)
    c.mdp_hipPairaGradient(None, None, None, None, None, None, None, None, None, None, None, None, 3, 0, 2, 0, 0, 4, 1)
    assert "mu" in _err(c)


def _snap_zbl_engine(oracle, J, x, box, coeff, v=None, dt=0.0005):
    from mdp_amd.md import SnapMD
    md = SnapMD(x, box, np.ones(len(x), np.int32), J, coeff, [0.5], [1.0], [MASS], v=v, dt=dt)
    md.set_zbl(4.0, 4.8, 73.0)
    return md


def _coeff(oracle, J, seed):
    s = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0])
    return lattice.snap_coefficients(s.ncoeff, seed=seed)


@pytest.mark.parametrize("J,ncell,a", [(6, 5, 3.0), (6, 6, 3.316), (8, 4, 2.95)])
def test_md_snap_plus_zbl_vs_oracle(oracle, J, ncell, a):
    x, box = lattice.crystal("bcc", ncell, a, 0.1, 50 + ncell)
    coeff = _coeff(oracle, J, 7)
    md = _snap_zbl_engine(oracle, J, x, box, coeff)
    print(f"SNAP+ZBL engine, J={J}, {len(x)} atoms, a={a}", flush=True)
    md.compute()
    st = md.state()
    s = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0], coeff)
    B = np.diag(box)
    p = snapcpp.pairs_from_box(x, np.ones(len(x), np.int32), B[0], B[1], B[2], [1, 1, 1], s)
    e, f, w = snapcpp.snappotential_pairs(oracle, s, p)
    ez, fz, wz = zbl_oracle.zbl_system(oracle, x, box)
    assert np.max(np.abs(fz)) > 1e-3 * np.max(np.abs(f))       # the pair term is not negligible here
    assert relerr(st["eatom"], e + ez) < 1e-10
    assert relerr(st["f"], f + fz) < 1e-10
    assert relerr(st["vatom"], w.reshape(-1, 6) + wz) < 1e-10


def test_md_snap_plus_zbl_finite_differences(oracle, monkeypatch):
    from mdp_amd.md import SnapMD
    monkeypatch.setenv("MDP_AUTOTUNE", "0")     # 24 small engines: default launch variants
    x, box = lattice.crystal("bcc", 4, 2.95, 0.12, 61)
    coeff = _coeff(oracle, 6, 61)
    md = _snap_zbl_engine(oracle, 6, x, box, coeff)
    md.compute()
    F = md.state()["f"]
    h = 1e-5
    rng = np.random.default_rng(2)
    for k in rng.choice(len(x), 4, replace=False):
        for c in range(3):
            es = []
            for sgn in (1, -1):
                xx = x.copy()
                xx[k, c] += sgn * h
                m = SnapMD(xx, box, np.ones(len(x), np.int32), 6, coeff, [0.5], [1.0], [MASS])
                m.set_zbl(4.0, 4.8, 73.0)
                m.compute()
                es.append(m.state()["eatom"].sum())
                m.close()
            fd = -(es[0] - es[1]) / (2 * h)
            assert abs(fd - F[k, c]) < 1e-6 * np.max(np.abs(F)), (k, c, fd, F[k, c])


def _energy_excursion(md, steps):
    e0 = sum(md.energy())
    dev = 0.0
    for _ in range(steps):
        md.run(1)
        dev = max(dev, abs(sum(md.energy()) - e0))
    return dev


def test_md_snap_plus_zbl_conserves_energy(oracle):
    x, box = lattice.crystal("bcc", 6, 3.316, 0.05, 71)
    v = lattice.maxwell_boltzmann(len(x), MASS, 300.0, seed=71)
    md = _snap_zbl_engine(oracle, 6, x, box, _coeff(oracle, 6, 71), v=v, dt=0.0005)
    md.compute()
    e0 = sum(md.energy())
    md.run(200)
    e1 = sum(md.energy())
    assert abs(e1 - e0) < 2e-4 * len(x) * 0.0259
    assert md.stats()["overflows"] == 0


def test_md_snap_plus_zbl_energy_error_scales_with_dt2(oracle):
    """Velocity Verlet's energy error over a fixed time span falls as dt^2: halving dt (and doubling the
    steps) divides the largest excursion of the total energy by about four.  A force inconsistent with
    the energy (a missing or mis-scaled pair term) leaves an error that does not shrink this way."""
    x, box = lattice.crystal("bcc", 5, 3.1, 0.08, 81)
    v = lattice.maxwell_boltzmann(len(x), MASS, 600.0, seed=81)
    coeff = _coeff(oracle, 6, 81)
    dev = []
    for dt, steps in ((0.001, 40), (0.0005, 80)):
        from mdp_amd.md import SnapMD
        md = SnapMD(x, box, np.ones(len(x), np.int32), 6, coeff, [0.5], [1.0], [MASS], v=v, dt=dt, rebuild_every=1)
        md.set_zbl(4.0, 4.8, 73.0)
        md.compute()
        dev.append(_energy_excursion(md, steps))
        md.close()
    ratio = dev[0] / dev[1]
    assert 3.0 < ratio < 5.0, (dev, ratio)
