"""The CPU baseline legs of bench.py, on small systems: the sliced force call adds up to the whole one,
the all-cores driver runs its worker processes over the same task, and the host MD loop follows velocity
Verlet with forces evaluated from scratch at every step."""
import os
import sys

import numpy as np
import pytest

from conftest import relerr
from mdp_amd import lattice, snapcpp

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

RCUTFAC, RFAC0, MASS = 4.67637, 0.99363, 180.95


def _system(oracle, J, ncell, sigma, seed):
    x, box = lattice.crystal("bcc", ncell, 3.316, sigma, seed)
    s0 = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0])
    coeff = lattice.snap_coefficients(s0.ncoeff, seed=seed)
    return x, box, bench._sna(oracle, J, coeff)


def test_force_chain_slices_add_up(oracle):
    x, box, s = _system(oracle, 4, 4, 0.1, 3)
    n = len(x)
    p = bench.periodic_pairs(x, box, RCUTFAC)
    e, f, v = snapcpp.snappotential_pairs(oracle, s, p)
    cuts = [0, 37, 64, n]
    parts = [bench.force_chain(oracle, s, p["rij"], p["ai"], p["aj"], lo, hi, n) for lo, hi in zip(cuts, cuts[1:])]
    assert relerr(np.concatenate([q[0] for q in parts]), e) < 1e-14
    assert relerr(sum(q[1] for q in parts), f) < 1e-12
    assert relerr(sum(q[2] for q in parts), v) < 1e-12


def test_all_cores_runs_every_worker(oracle):
    x, box, s = _system(oracle, 2, 3, 0.1, 4)
    p = bench.periodic_pairs(x, box, RCUTFAC)
    _, path = bench.cpu_library()
    rate, slowest = bench.all_cores(dict(task="force", lib=path, n=len(x), twojmax=2, coeff=s.coeffelem,
                                         rij=p["rij"], ai=p["ai"], aj=p["aj"]), 2)
    assert rate > 0 and slowest > 0


def test_cpu_md_is_velocity_verlet(oracle):
    from mdp_amd.lattice import FTM2V_METAL
    x, box, s = _system(oracle, 4, 4, 0.05, 5)
    v = lattice.maxwell_boltzmann(len(x), MASS, 300.0, seed=5)
    steps, dt = 4, 0.0005
    xc, vc, _ = bench.cpu_md(oracle, s, x, box, v, steps, dt=dt, every=2)
    B = np.diag(box)

    def force(xx):
        q = snapcpp.pairs_from_box(np.mod(xx, box), np.ones(len(x), np.int32), B[0], B[1], B[2], [1, 1, 1], s)
        return snapcpp.snappotential_pairs(oracle, s, q)[1]
    dtf = 0.5 * dt * FTM2V_METAL / MASS
    xh, vh = x.copy(), v.copy()
    f = force(xh)
    for _ in range(steps):
        vh += dtf * f
        xh += dt * vh
        f = force(xh)
        vh += dtf * f
    assert np.max(np.abs(xc - xh)) < 1e-10
    assert relerr(vc, vh) < 1e-10


@pytest.mark.parametrize("every", [3])
def test_fit_cpu_baseline_small(every):
    cfg = [(x, box) for x, box, _ in lattice.fitting_batch(nconfig=6, nmin=16, nmax=40)]
    out = bench.fit_cpu_baseline(cfg, every, cores=2)
    assert out["value"] > 0 and out["cores"] == 2


def test_tungsten_coefficients_load():
    from mdp_amd import surrogate as S
    pot = S.tungsten(os.path.join(os.path.dirname(__file__), "golden", "snapW.coeff"))
    assert pot.coeff.shape == (56,) and np.all(np.isfinite(pot.coeff))
    assert (pot.a0, pot.rcutfac, pot.mass) == (3.1803, 4.8123, 183.84)
    assert pot.zbl == (4.0, 4.8, 74.0) and S.tungsten(os.path.join(os.path.dirname(__file__), "golden",
                                                                  "snapW.coeff"), with_zbl=False).zbl is None
