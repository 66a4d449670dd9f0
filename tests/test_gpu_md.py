"""GPU MD engine: neighbour lists as canonical sets, first-step energy/force/virial parity with the CPU
oracle (small systems whole; the 128,000-atom twojmax=8 system on a seeded sample of atoms), and
size-independent properties of the dynamics (momentum and energy conservation).

Tolerances: neighbour sets bit-exact; energies, forces, virials within 1e-10 relative to the max
norm of the oracle array.
"""
import numpy as np
import pytest

from conftest import relerr
from mdp_amd import lattice, snapcpp

pytestmark = pytest.mark.gpu

RCUTFAC, RFAC0 = 4.67637, 0.99363
MASS = 180.95
TOL = 1e-10


def _engine(x, box, J, coeff, v=None, skin=0.3, every=10, dt=0.0005):
    from mdp_amd.md import SnapMD
    return SnapMD(x, box, np.ones(len(x), np.int32), J, coeff, [0.5], [1.0], [MASS], v=v, dt=dt, skin=skin,
                  rebuild_every=every)


def _coeff(oracle, J, seed):
    s = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0])
    return lattice.snap_coefficients(s.ncoeff, seed=seed)


def _oracle_bruteforce(oracle, x, box, rc):
    import ctypes
    n = len(x)
    off = np.zeros(n + 1, np.int32)
    maxnb = 128 * n
    nj = np.zeros(maxnb, np.int32)
    ns = np.zeros(3 * maxnb, np.int32)
    f = oracle.cdll.oracle_neighbor_bruteforce
    f.restype = ctypes.c_long
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C")
    f.argtypes = [dp, ctypes.c_int, dp, ctypes.c_double, ip, ip, ip, ctypes.c_long]
    cnt = f(np.ascontiguousarray(x), n, np.ascontiguousarray(box), rc, off, nj, ns, maxnb)
    assert cnt > 0
    return off, nj, ns.reshape(-1, 3)


@pytest.mark.parametrize("variant", ["0", "1"])
@pytest.mark.parametrize("ncell,a", [(6, 3.316), (2, 3.316), (5, 3.0)])
def test_neighbor_sets_bit_exact(oracle, monkeypatch, ncell, a, variant):
    """Neighbour lists of the cell-list kernels (0: eight lanes per atom through L1, 1: the per-wavefront
    LDS tile) as sorted sets against brute force (the 2-cell box uses the brute-force kernel)."""
    monkeypatch.setenv("MDP_NB_VARIANT", variant)
    x, box = lattice.crystal("bcc", ncell, a, 0.15, 100 + ncell)
    md = _engine(x, box, 4, _coeff(oracle, 4, 1))
    md.compute()
    off, nj, ns = md.neighbors()
    xw = md.state()["x"]
    assert np.array_equal(xw, x)   # already inside the box: wrapping leaves it unchanged
    o2, j2, s2 = _oracle_bruteforce(oracle, xw, box, RCUTFAC + 0.3)
    for i in range(len(x)):
        a_ = sorted(zip(nj[off[i]:off[i + 1]].tolist(), map(tuple, ns[off[i]:off[i + 1]].tolist())))
        b_ = sorted(zip(j2[o2[i]:o2[i + 1]].tolist(), map(tuple, s2[o2[i]:o2[i + 1]].tolist())))
        assert a_ == b_, i


def test_neighbor_variants_identical(oracle, monkeypatch):
    """Both cell-list kernels write the same lists, entry for entry (same candidate order), on a hot
    disordered system with ragged counts, and on a dilute gas with empty cells and isolated atoms."""
    rng = np.random.default_rng(5)
    x1, box1 = lattice.crystal("bcc", 9, 3.316, 0.35, 61)
    x2 = rng.uniform(0.0, 40.0, size=(400, 3))
    for x, box in ((x1, box1), (x2, np.full(3, 40.0))):
        lists = []
        for v in ("0", "1"):
            monkeypatch.setenv("MDP_NB_VARIANT", v)
            md = _engine(x, box, 4, _coeff(oracle, 4, 1))
            md.compute()
            assert md.tuning()["chosen"]["neighbor_build"] == int(v)
            lists.append(md.neighbors())
            md.close()
        for a_, b_ in zip(*lists):
            assert np.array_equal(a_, b_)


@pytest.mark.parametrize("J,ncell,seed", [(4, 6, 1), (6, 6, 2), (8, 6, 3), (8, 4, 4)])
def test_first_step_vs_oracle(oracle, J, ncell, seed):
    x, box = lattice.crystal("bcc", ncell, 3.316, 0.1, seed)
    coeff = _coeff(oracle, J, seed)
    md = _engine(x, box, J, coeff)
    md.compute()
    st = md.state()
    s = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0], coeff)
    B = np.diag(box)
    p = snapcpp.pairs_from_box(x, np.ones(len(x), np.int32), B[0], B[1], B[2], [1, 1, 1], s)
    e, f, v = snapcpp.snappotential_pairs(oracle, s, p)
    assert md.stats()["pairs"] == p["ijnum"]
    assert relerr(st["eatom"], e) < TOL
    assert relerr(st["f"], f) < TOL
    assert relerr(st["vatom"], v) < TOL


@pytest.mark.parametrize("variant", ["0", "2", "3", "4", "5", "6", "7", "8", "9", "10", "11", "12", "13"])
def test_every_dedr_variant_vs_oracle(oracle, monkeypatch, variant):
    """Each dE/dr kernel variant the autotuner may pick, pinned in turn (twojmax = 8, 3,456 atoms)."""
    monkeypatch.setenv("MDP_DEDR_VARIANT", variant)
    x, box = lattice.crystal("bcc", 12, 3.316, 0.1, 31)
    coeff = _coeff(oracle, 8, 31)
    md = _engine(x, box, 8, coeff)
    md.compute()
    st = md.state()
    s = snapcpp.initsna(oracle, [1, 8, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0], coeff)
    B = np.diag(box)
    p = snapcpp.pairs_from_box(x, np.ones(len(x), np.int32), B[0], B[1], B[2], [1, 1, 1], s)
    e, f, v = snapcpp.snappotential_pairs(oracle, s, p)
    assert relerr(st["eatom"], e) < TOL
    assert relerr(st["f"], f) < TOL
    assert relerr(st["vatom"], v) < TOL


def _first_step_vs_oracle(oracle, x, box, J, seed):
    coeff = _coeff(oracle, J, seed)
    md = _engine(x, box, J, coeff)
    md.compute()
    st = md.state()
    s = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0], coeff)
    B = np.diag(box)
    p = snapcpp.pairs_from_box(x, np.ones(len(x), np.int32), B[0], B[1], B[2], [1, 1, 1], s)
    e, f, v = snapcpp.snappotential_pairs(oracle, s, p)
    assert md.stats()["pairs"] == p["ijnum"]
    assert relerr(st["eatom"], e) < TOL
    assert relerr(st["f"], f) < TOL
    assert relerr(st["vatom"], v) < TOL
    return md


@pytest.mark.parametrize("variant", ["8", "12", "4", "112", "113", "115", "116"])
def test_every_yi_variant_vs_oracle(oracle, monkeypatch, variant):
    """compute_yi pinned to each twojmax-8 variant: 8, 12 or 4 waves per workgroup, or 12 waves with the
    larger layer's U rows shared across couplings (112), also with (re, im) adjacent in LDS (113), with the
    swap symmetry of the (j1, j1) couplings (115), and with the row index made wave-uniform (116)."""
    monkeypatch.setenv("MDP_YI_VARIANT", variant)
    x, box = lattice.crystal("bcc", 12, 3.316, 0.1, 51)
    md = _first_step_vs_oracle(oracle, x, box, 8, 51)
    assert md.tuning()["chosen"]["compute_yi_waves"] == int(variant)


@pytest.mark.parametrize("variant", ["0", "256", "512", "1024"])
@pytest.mark.parametrize("J,ncell,a,seed", [(8, 12, 3.316, 41), (4, 2, 3.316, 42), (6, 5, 3.0, 43)])
def test_every_ui_variant_vs_oracle(oracle, monkeypatch, variant, J, ncell, a, seed):
    """compute_ui pinned to the per-atom kernel (0) or to the LDS-sum kernel on 256/512/1024 workgroups:
    3,456 atoms (blocks of ~21 atoms), 16 atoms (one block, most workgroups idle), a compressed lattice
    (more pairs per atom)."""
    monkeypatch.setenv("MDP_UI_VARIANT", variant)
    x, box = lattice.crystal("bcc", ncell, a, 0.1, seed)
    md = _first_step_vs_oracle(oracle, x, box, J, seed)
    assert md.tuning()["chosen"]["compute_ui"] == int(variant)


@pytest.mark.parametrize("variant", ["0", "512"])
def test_ui_variants_sparse_gas(oracle, monkeypatch, variant):
    """A dilute random gas: ragged pair counts, many atoms with no pair at all (U is wself on the diagonal)."""
    monkeypatch.setenv("MDP_UI_VARIANT", variant)
    rng = np.random.default_rng(77)
    L = 40.0
    box = np.full(3, L)
    x = rng.uniform(0.0, L, size=(300, 3))
    # keep pairs at least 1.6 A apart, so the energies stay in a sane range
    keep = [0]
    for i in range(1, len(x)):
        d = x[keep] - x[i]
        d -= L * np.round(d / L)
        if np.min(np.sum(d * d, axis=1)) > 1.6 ** 2:
            keep.append(i)
    x = np.ascontiguousarray(x[keep])
    md = _first_step_vs_oracle(oracle, x, box, 8, 78)
    off, _, _ = md.neighbors()
    assert np.any(np.diff(off) == 0)


@pytest.mark.parametrize("J", [4, 8])
def test_first_step_two_types_vs_oracle(oracle, J):
    """Two atom types with their own radii and weights (pair cutoff (R_i + R_j) rcutfac, neighbour weight
    w_j): energies, forces and virials against the oracle.  With chemflag = 0 upstream applies the first
    coefficient block to every atom (cpuSnap.cpp:573, :924); the oracle gets a different second block to show
    that it is not read."""
    from mdp_amd.md import SnapMD
    x, box = lattice.crystal("bcc", 6, 3.316, 0.1, 71)
    types = np.where(np.random.default_rng(72).random(len(x)) < 0.35, 2, 1).astype(np.int32)
    rad, wt = [0.5, 0.44], [1.0, 0.73]
    param = [2, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0]
    s0 = snapcpp.initsna(oracle, [1, J] + param[2:], [0.5], [1.0])
    c1 = lattice.snap_coefficients(s0.ncoeff, seed=73)
    c2 = lattice.snap_coefficients(s0.ncoeff, seed=74)
    md = SnapMD(x, box, types, J, c1, rad, wt, [MASS, 92.9])
    md.compute()
    st = md.state()
    s = snapcpp.initsna(oracle, param, rad, wt, np.concatenate([c1, c2]))
    B = np.diag(box)
    p = snapcpp.pairs_from_box(x, types, B[0], B[1], B[2], [1, 1, 1], s)
    e, f, v = snapcpp.snappotential_pairs(oracle, s, p)
    assert md.stats()["pairs"] == p["ijnum"]
    assert relerr(st["eatom"], e) < TOL
    assert relerr(st["f"], f) < TOL
    assert relerr(st["vatom"], v) < TOL


def test_first_step_128k_twojmax8_sampled(oracle):
    """The headline system: BCC Ta 40^3 cells (128,000 atoms), sigma = 0.1 A, twojmax = 8."""
    import sampled_oracle
    x, box = lattice.crystal("bcc", 40, 3.316, 0.1, 2024)
    coeff = _coeff(oracle, 8, 2024)
    md = _engine(x, box, 8, coeff)
    md.compute()
    st = md.state()
    sample = np.random.default_rng(5).choice(len(x), size=24, replace=False)
    s = sampled_oracle.init(oracle, 8, coeff)
    e, f, v = sampled_oracle.sampled_eval(oracle, s, x, box, sample, RCUTFAC)
    assert relerr(st["eatom"][sample], e) < TOL
    assert relerr(st["f"][sample], f) < TOL
    assert relerr(st["vatom"][sample], v) < TOL
    # size-independent checks on the whole system
    assert np.max(np.abs(st["f"].sum(axis=0))) < 1e-9 * np.max(np.abs(st["f"]))
    assert np.all(np.isfinite(st["f"]))


def test_nve_conservation(oracle):
    x, box = lattice.crystal("bcc", 6, 3.316, 0.05, 9)
    v = lattice.maxwell_boltzmann(len(x), MASS, 300.0, seed=9)
    md = _engine(x, box, 6, _coeff(oracle, 6, 9), v=v, dt=0.0005)
    md.compute()
    e0 = sum(md.energy())
    p0 = v.sum(axis=0) * MASS
    md.run(200)
    e1 = sum(md.energy())
    st = md.state()
    assert abs(e1 - e0) < 2e-4 * len(x) * 0.0259      # drift << kT per atom
    assert np.max(np.abs(st["v"].sum(axis=0) * MASS - p0)) < 1e-8
    stats = md.stats()
    assert stats["rebuilds"] == 1 + 20 and stats["overflows"] == 0


def test_short_trajectory_vs_oracle_integration(oracle):
    """Ten velocity-Verlet steps driven by oracle forces on the host against the GPU engine."""
    from mdp_amd.lattice import FTM2V_METAL
    J, dt, steps = 4, 0.0005, 10
    x, box = lattice.crystal("bcc", 4, 3.316, 0.05, 11)
    v = lattice.maxwell_boltzmann(len(x), MASS, 300.0, seed=11)
    coeff = _coeff(oracle, J, 11)
    md = _engine(x, box, J, coeff, v=v, dt=dt, every=1)
    md.compute()
    md.run(steps)
    g = md.state()
    s = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0], coeff)
    B = np.diag(box)
    t = np.ones(len(x), np.int32)

    def force(xx):
        p = snapcpp.pairs_from_box(np.mod(xx, box), t, B[0], B[1], B[2], [1, 1, 1], s)
        return snapcpp.snappotential_pairs(oracle, s, p)[1]

    dtf = 0.5 * dt * FTM2V_METAL / MASS
    xh, vh = x.copy(), v.copy()
    f = force(xh)
    for _ in range(steps):
        vh += dtf * f
        xh += dt * vh
        f = force(xh)
        vh += dtf * f
    d = np.mod(g["x"] - xh + 0.5 * box, box) - 0.5 * box
    assert np.max(np.abs(d)) < 1e-9
    assert relerr(g["v"], vh) < 1e-9


def test_twojmax8_energy_error_scales_with_dt2(oracle):
    """twojmax = 8 on the headline lattice (sigma = 0.1 A, 300 K): velocity Verlet's largest total-energy
    excursion over 25 fs falls about fourfold when dt is halved, so the forces are the gradient of the
    energy the engine reports over many steps (an inconsistent force term would not scale this way)."""
    x, box = lattice.crystal("bcc", 5, 3.316, 0.1, 91)
    v = lattice.maxwell_boltzmann(len(x), MASS, 300.0, seed=91)
    coeff = _coeff(oracle, 8, 91)
    dev = []
    for dt, steps in ((0.001, 25), (0.0005, 50)):
        md = _engine(x, box, 8, coeff, v=v, dt=dt, every=1)
        md.compute()
        e0 = sum(md.energy())
        d = 0.0
        for _ in range(steps):
            md.run(1)
            d = max(d, abs(sum(md.energy()) - e0))
        dev.append(d)
        md.close()
    ratio = dev[0] / dev[1]
    assert 3.0 < ratio < 5.0, (dev, ratio)


def test_tungsten_potential_first_step(oracle):
    """The upstream tungsten potential (snapW.coeff with snapW.m:34-53: BCC a = 3.1803 A, rcutfac 4.8123,
    twojmax 8): first-step energies, forces and virials against the oracle; the per-atom energy of the
    SNAP part alone is about -15.7 eV on this lattice (oracle)."""
    import os
    from mdp_amd import surrogate as S
    from mdp_amd.md import SnapMD
    pot = S.tungsten(os.path.join(os.path.dirname(__file__), "golden", "snapW.coeff"))
    x, box = lattice.crystal("bcc", 5, pot.a0, 0.05, 77)
    md = SnapMD(x, box, np.ones(len(x), np.int32), 8, pot.coeff, [0.5], [1.0], [pot.mass], rcutfac=pot.rcutfac,
                rfac0=pot.rfac0)
    md.compute()
    st = md.state()
    s = snapcpp.initsna(oracle, [1, 8, pot.rcutfac, pot.rfac0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0], pot.coeff)
    B = np.diag(box)
    p = snapcpp.pairs_from_box(x, np.ones(len(x), np.int32), B[0], B[1], B[2], [1, 1, 1], s)
    e, f, v = snapcpp.snappotential_pairs(oracle, s, p)
    assert relerr(st["eatom"], e) < TOL
    assert relerr(st["f"], f) < TOL
    assert relerr(st["vatom"], v) < TOL
    assert -17.0 < np.mean(e) < -14.0


def test_timing_modes(oracle):
    """set_timing: every stage, or only the three SNAP compute kernels (the bench's timed region)."""
    x, box = lattice.crystal("bcc", 4, 3.316, 0.05, 13)
    md = _engine(x, box, 6, _coeff(oracle, 6, 13))
    md.compute()
    md.set_timing(True)
    md.run(3)
    t = md.timers()
    assert t["compute_yi"][1] == 3 and t["integrate"][1] == 3 and t["tally_integrate"][1] == 3
    md.set_timing(True, compute_only=True)
    md.run(4)
    t = md.timers()
    assert all(t[k][1] == 4 and t[k][0] > 0 for k in ("compute_ui", "compute_yi", "compute_dedr"))
    assert t["integrate"][1] == 0 and t["pair_list"][1] == 0 and t["tally_integrate"][1] == 0
    md.set_timing(False)
    md.run(2)
    assert md.timers()["compute_yi"][1] == 0
