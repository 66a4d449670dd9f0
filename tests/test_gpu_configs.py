"""The workloads at their stated shapes.

* 2,000-atom MD (BCC Ta 10^3 cells, twojmax=6, sigma=0.05 A): first-step energies, forces and virials of
  the whole system against the CPU oracle (1e-10 relative to the max norm), and a 100-step NVE
  trajectory (skin 0.3 A, rebuilt every 10 steps) against host velocity Verlet driven by the oracle's
  forces at every step (positions and velocities within 1e-9).
* Surrogate inference at 128,000 atoms (BCC Ta 40^3 cells, twojmax=4 descriptors): engine forces on a
  seeded sample of atoms against the float64 descriptor chain rule on the oracle's descriptors (BF16
  network: within 5e-2 of the max force), and the sum of all forces against zero.
* Audited surrogate MD: with a zero tolerance every audit falls back, and the forces, virials and
  energies after an audit step are plain SNAP's on that state; with an unreachable tolerance nothing is
  replaced and the trajectory equals the unaudited one; the RMS force difference an audit reports is
  the one between the two force fields.  Then a few steps of the 1,018,880-atom vacancy system
  (80^3 BCC cells, 0.5% vacant sites, 600 K, dt = 1 fs) with audits.
"""
import numpy as np
import pytest

import mlp_ref as R
import sampled_oracle
from conftest import relerr
from mdp_amd import lattice, snapcpp

pytestmark = pytest.mark.gpu

RCUTFAC, RFAC0, MASS = 4.67637, 0.99363, 180.95
TOL = 1e-10


def _param(J):
    return [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0]


def _coeff(oracle, J, seed):
    s = snapcpp.initsna(oracle, _param(J), [0.5], [1.0])
    return lattice.snap_coefficients(s.ncoeff, beta0=-11.7, sigma=0.05, seed=seed)


def _engine(x, box, J, coeff, v=None, dt=0.0005, skin=0.3, every=10):
    from mdp_amd.md import SnapMD
    return SnapMD(x, box, np.ones(len(x), np.int32), J, coeff, [0.5], [1.0], [MASS], v=v, dt=dt, skin=skin,
                  rebuild_every=every)


def _rand_model(seed, nin=14):
    """A seeded network with non-trivial biases and a descriptor affine map of O(1) scale."""
    rng = np.random.default_rng(seed)
    p = R.init_params(seed, nin)
    p[1024:1088] = rng.normal(0, 0.3, 64)
    p[5184:5248] = rng.normal(0, 0.3, 64)
    p[5312] = 0.1
    aff = np.zeros(34)
    aff[16:16 + nin] = 0.2
    aff[32], aff[33] = 1.3, -2.0
    return p.astype(np.float32), aff


# ---------------------------------------------------------------- 2,000-atom twojmax=6 MD


def _config0(oracle):
    J = 6
    x, box = lattice.crystal("bcc", 10, 3.316, 0.05, 20241200)
    v = lattice.maxwell_boltzmann(len(x), MASS, 300.0, seed=20241201)
    coeff = _coeff(oracle, J, 20241202)
    return J, x, box, v, coeff


def test_config0_first_step_whole_system(oracle):
    J, x, box, v, coeff = _config0(oracle)
    assert len(x) == 2000
    md = _engine(x, box, J, coeff, v=v)
    print("2000-atom twojmax=6 first step", flush=True)
    md.compute()
    st = md.state()
    s = snapcpp.initsna(oracle, _param(J), [0.5], [1.0], coeff)
    B = np.diag(box)
    p = snapcpp.pairs_from_box(x, np.ones(len(x), np.int32), B[0], B[1], B[2], [1, 1, 1], s)
    e, f, w = snapcpp.snappotential_pairs(oracle, s, p)
    assert md.stats()["pairs"] == p["ijnum"]
    assert relerr(st["eatom"], e) < TOL
    assert relerr(st["f"], f) < TOL
    assert relerr(st["vatom"], w) < TOL


def test_config0_100_steps_vs_oracle_verlet(oracle):
    """100 NVE steps, dt = 0.5 fs, skin 0.3 A and a rebuild every 10 steps on the GPU, against velocity
    Verlet on the host with exact pairs and oracle forces at every step."""
    from mdp_amd.lattice import FTM2V_METAL
    J, x, box, v, coeff = _config0(oracle)
    dt, steps = 0.0005, 100
    md = _engine(x, box, J, coeff, v=v, dt=dt)
    md.compute()
    print("2000-atom twojmax=6: 100 GPU steps", flush=True)
    md.run(steps)
    g = md.state()
    st = md.stats()
    assert st["rebuilds"] == 1 + steps // 10 and st["overflows"] == 0
    s = snapcpp.initsna(oracle, _param(J), [0.5], [1.0], coeff)
    B = np.diag(box)
    t = np.ones(len(x), np.int32)

    def force(xx):
        p = snapcpp.pairs_from_box(np.mod(xx, box), t, B[0], B[1], B[2], [1, 1, 1], s)
        return snapcpp.snappotential_pairs(oracle, s, p)

    dtf = 0.5 * dt * FTM2V_METAL / MASS
    xh, vh = x.copy(), v.copy()
    e, f, _ = force(xh)
    print("2000-atom twojmax=6: host Verlet with oracle forces", flush=True)
    for _ in range(steps):
        vh += dtf * f
        xh += dt * vh
        e, f, w = force(xh)
        vh += dtf * f
    d = np.mod(g["x"] - xh + 0.5 * box, box) - 0.5 * box
    assert np.max(np.abs(d)) < 1e-9
    assert relerr(g["v"], vh) < 1e-9
    assert relerr(g["f"], f) < 1e-9
    assert abs(g["eatom"].sum() - e.sum()) < 1e-9 * abs(e.sum())


# ---------------------------------------------------------------- surrogate inference at 128,000 atoms


def test_config3_inference_128k_sampled(oracle):
    from mdp_amd import surrogate as S
    from mdp_amd.md import SnapMD
    x, box = lattice.crystal("bcc", 40, 3.316, 0.1, 20241213)
    assert len(x) == 128000
    p, aff = _rand_model(7)
    md = SnapMD(x, box, np.ones(len(x), np.int32), 8, S.plain_coefficients(), [0.5], [1.0], [MASS])
    md.set_surrogate(S.Model(p, aff, 14), audit_every=0)
    print("surrogate inference, 128,000 atoms", flush=True)
    md.compute()
    st = md.state()
    F = st["f"]
    assert np.all(np.isfinite(F))
    assert np.max(np.abs(F.sum(axis=0))) < 1e-9 * len(x) * np.max(np.abs(F))
    sample = np.random.default_rng(13).choice(len(x), size=24, replace=False)
    sna4 = sampled_oracle.init(oracle, 4, None)
    Fs, es = sampled_oracle.sampled_surrogate_forces(oracle, sna4, x, box, sample, RCUTFAC,
                                                     lambda Bm: R.forward(p, aff, 14, Bm))
    assert relerr(F[sample], Fs) < 5e-2
    assert relerr(st["eatom"][sample], es) < 2e-2


# ---------------------------------------------------------------- audited surrogate MD


def _audit_pair(oracle, every, tol, steps, seed=41):
    from mdp_amd import surrogate as S
    x, box = lattice.crystal("bcc", 5, 3.316, 0.1, seed)
    v = lattice.maxwell_boltzmann(len(x), MASS, 600.0, seed=seed + 1)
    coeff = _coeff(oracle, 8, seed + 2)
    p, aff = _rand_model(seed + 3)
    md = _engine(x, box, 8, coeff, v=v, dt=0.001)
    md.set_surrogate(S.Model(p, aff, 14), audit_every=every, audit_tol=tol)
    md.compute()
    md.run(steps)
    return md, (x, box, v, coeff, p, aff)


def test_config4_audit_zero_tol_falls_back_to_plain(oracle):
    """Every step audited with tolerance 0: every audit falls back, and the state after the last step
    carries plain SNAP's forces, virials and energies for its positions."""
    md, (x, box, v, coeff, p, aff) = _audit_pair(oracle, 1, 0.0, 6)
    aud = md.audits()
    assert aud["audits"] == 6 and aud["fallbacks"] == 6
    st = md.state()
    plain = _engine(np.mod(st["x"], box), box, 8, coeff, v=st["v"], dt=0.001)
    plain.compute()
    ps = plain.state()
    assert relerr(st["f"], ps["f"]) < TOL
    assert relerr(st["vatom"], ps["vatom"]) < TOL
    assert relerr(st["eatom"], ps["eatom"]) < TOL
    # the reported RMS difference is the one between the surrogate's and plain SNAP's forces on that state
    from mdp_amd import surrogate as S
    sur = _engine(np.mod(st["x"], box), box, 8, coeff, v=st["v"], dt=0.001)
    sur.set_surrogate(S.Model(p, aff, 14), audit_every=0)
    sur.compute()
    rms = float(np.sqrt(np.mean((sur.state()["f"] - ps["f"]) ** 2)))
    assert abs(aud["last_rmse"] - rms) < 1e-2 * rms


def test_config4_audit_loose_tol_keeps_surrogate(oracle):
    """Audits every other step with an unreachable tolerance: audits counted, nothing replaced, and
    the trajectory is the unaudited surrogate run's."""
    a, _ = _audit_pair(oracle, 2, 1e30, 8)
    b, _ = _audit_pair(oracle, 0, 0.05, 8)
    aud = a.audits()
    assert aud["audits"] == 4 and aud["fallbacks"] == 0 and aud["max_rmse"] > 0
    assert b.audits()["audits"] == 0
    sa, sb = a.state(), b.state()
    assert relerr(sa["x"], sb["x"]) < 1e-12
    assert relerr(sa["v"], sb["v"]) < 1e-12
    assert relerr(sa["f"], sb["f"]) < 1e-12


def test_config4_vacancy_system_1m_atoms(oracle):
    """80^3 BCC cells with 0.5% vacant sites (1,018,880 atoms), 600 K, dt = 1 fs: surrogate steps with an
    audit every 2 steps (a seeded untrained network, so every audit falls back) and an audit-free run."""
    from mdp_amd import surrogate as S
    from mdp_amd.md import SnapMD
    seed = 20241212
    x, box = lattice.crystal("bcc", 80, 3.316, 0.1, seed)
    keep = np.ones(len(x), bool)
    keep[np.random.default_rng(seed + 1).choice(len(x), int(round(0.005 * len(x))), replace=False)] = False
    x = np.ascontiguousarray(x[keep])
    n = len(x)
    assert n == 1018880
    v = lattice.maxwell_boltzmann(n, MASS, 600.0, seed=seed + 2)
    p, aff = _rand_model(seed + 3)
    md = SnapMD(x, box, np.ones(n, np.int32), 8, S.plain_coefficients(), [0.5], [1.0], [MASS], v=v, dt=0.001)
    md.set_surrogate(S.Model(p, aff, 14), audit_every=2, audit_tol=0.05)
    print("1,018,880-atom surrogate MD with audits", flush=True)
    md.compute()
    p0 = MASS * v.sum(axis=0)
    md.run(4)
    aud = md.audits()
    assert aud["audits"] == 2 and aud["fallbacks"] == 2
    st = md.state()
    assert np.all(np.isfinite(st["x"])) and np.all(np.isfinite(st["v"])) and np.all(np.isfinite(st["f"]))
    # plain forces were taken on the last audit step: they sum to zero, and momentum is conserved
    assert np.max(np.abs(st["f"].sum(axis=0))) < 1e-9 * n * np.max(np.abs(st["f"]))
    assert np.max(np.abs(MASS * st["v"].sum(axis=0) - p0)) < 1e-6 * np.max(np.abs(MASS * v))
    assert md.stats()["overflows"] == 0
