"""Linear-fit consumer of the batched descriptors (mdp_fit_descriptors, mdp_fit_normal_equations,
mdp_lsqsolve).

* The per-configuration tallies bi, bd, bv against the oracle's descriptor chain (snapdescriptors on
  the same pairs, cpuSnapTallyBispectrum{,Deriv,Virial}) within 1e-10 of the max norm.
* The weighted normal equations of the energy, force and stress rows against numpy on the oracle's
  tallies (rows as in emldescriptors.jl, weights per configuration, normalizeenergy on and off).
* Coefficient recovery across routes: energies and forces of a SNAP potential with known coefficients
  come from the oracle's force chain (Ui -> Yi -> Fi, no dB/dr involved); fitting the GPU descriptors to
  them returns those coefficients.
"""
import numpy as np
import pytest

from conftest import relerr
from mdp_amd import lattice, snapcpp

pytestmark = pytest.mark.gpu

RCUTFAC, RFAC0 = 4.67637, 0.99363
TOL = 1e-10


def _batch(seed=3):
    cfg = []
    # the 250-atom cell has more pairs than one LDS tile of the derivative tally (4096)
    for k, (kind, n, a) in enumerate([("bcc", 2, 3.3), ("fcc", (2, 2, 3), 4.1), ("bcc", 3, 3.25),
                                      ("bcc", (1, 2, 2), 3.4), ("bcc", 3, 3.12), ("bcc", 5, 3.3)]):
        x, box = lattice.crystal(kind, n, a, 0.1, seed + k)
        cfg.append((x, box))
    x, box = lattice.random_packing(40, 0.0554, 2.2, seed=seed + 10)
    cfg.append((np.mod(x + np.random.default_rng(seed).normal(0, 0.1, x.shape), box), box))
    # configurations without pairs (zero force rows, isolated-atom descriptors)
    cfg.append((np.array([[3.0, 4.0, 5.0]]), np.array([20.0, 20.0, 20.0])))
    cfg.append((np.array([[1.0, 1.0, 1.0], [11.0, 11.0, 11.0]]), np.array([20.0, 20.0, 20.0])))
    return cfg


def _oracle_tallies(oracle, J, fit, c, cfg):
    """bi (K,), bd (3, N_c, K), bv (6, K) of configuration c from the oracle on the GPU batch's pairs."""
    x = cfg[c][0]
    n = len(x)
    ai, aj, rij = fit.pairs()
    _, p_off = fit.layout()
    sl = slice(int(p_off[c]), int(p_off[c + 1]))
    sna = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0])
    P = sl.stop - sl.start
    one = np.ones(P, np.int32)
    p = dict(rij=np.ascontiguousarray(rij[sl]).ravel(), ai=np.ascontiguousarray(ai[sl], np.int32),
             aj=np.ascontiguousarray(aj[sl], np.int32), ti=one, tj=one, atomtype=np.ones(n, np.int32),
             ilist=np.arange(n, dtype=np.int32), n=n, ijnum=P)
    bi, bd, bv, _ = snapcpp.snapdescriptors_pairs(oracle, sna, p)
    return bi, bd.transpose(2, 1, 0), bv.T


@pytest.mark.parametrize("J", [4, 8])
def test_tallies_vs_oracle(oracle, J):
    from mdp_amd.fit import SnapFit
    cfg = _batch()
    fit = SnapFit(cfg, J)
    fit.compute()
    got = fit.descriptors()
    print(f"tallies J={J}: {[len(c[0]) for c in cfg]} atoms", flush=True)
    for c in range(len(cfg)):
        ob, od, ov = _oracle_tallies(oracle, J, fit, c, cfg)
        gb, gd, gv = got[c]
        # bi sums B - bzero, which vanishes for atoms without pairs: scale by N_c (J + 1)
        assert np.max(np.abs(gb - ob)) < TOL * max(np.max(np.abs(ob)), len(cfg[c][0]) * (J + 1.0)), c
        assert relerr(gd, od) < TOL, c
        assert relerr(gv, ov) < TOL, c


@pytest.mark.parametrize("normalizeenergy", [0, 1])
def test_normal_equations_vs_numpy(oracle, normalizeenergy):
    from mdp_amd.fit import SnapFit
    J = 6
    cfg = _batch(5)
    C = len(cfg)
    fit = SnapFit(cfg, J)
    fit.compute()
    rng = np.random.default_rng(11)
    nat = np.array([len(c[0]) for c in cfg])
    e = rng.normal(0, 10, C)
    f = rng.normal(0, 1, (nat.sum(), 3))
    s = rng.normal(0, 1, (C, 6))
    we, wf, ws = rng.uniform(0.5, 2, C), rng.uniform(0.5, 2, C), rng.uniform(0.5, 2, C)
    L, h = fit.normal_equations(e, f, s, we, wf, ws, normalizeenergy)
    K = fit.ncoeff
    refL, refh = np.zeros((3, K, K)), np.zeros((3, K))
    refL1, refh1 = np.zeros((K, K)), np.zeros(K)   # energy rows with unit weights, no normalisation
    a0 = 0
    for c in range(C):
        bi, bd, bv = _oracle_tallies(oracle, J, fit, c, cfg)
        refL1 += np.outer(bi, bi)
        refh1 += bi * e[c]
        n = nat[c]
        scale = 1.0 / n if normalizeenergy else 1.0
        ae, be = bi * scale * we[c], e[c] * scale * we[c]
        refL[0] += np.outer(ae, ae)
        refh[0] += ae * be
        Af = bd.transpose(1, 0, 2).reshape(3 * n, K) * wf[c]    # rows comp + 3*atom
        bf = f[a0:a0 + n].reshape(-1) * wf[c]
        refL[1] += Af.T @ Af
        refh[1] += Af.T @ bf
        As, bs = bv * ws[c], s[c] * ws[c]
        refL[2] += As.T @ As
        refh[2] += As.T @ bs
        a0 += n
    for g in range(3):
        assert relerr(L[g], refL[g]) < TOL, g
        assert relerr(h[g], refh[g]) < TOL, g
    # default weights; a missing target leaves its group zero
    L2, h2 = fit.normal_equations(e=e)
    assert relerr(L2[0], refL1) < TOL and relerr(h2[0], refh1) < TOL
    assert not L2[1].any() and not h2[1].any() and not L2[2].any() and not h2[2].any()


@pytest.mark.parametrize("J", [4, 6])
def test_fit_recovers_coefficients(oracle, J):
    """Forces and energies from the oracle's force chain with known coefficients (beta0 = 0); the fit
    on the GPU descriptors returns them (cf from forces, cef from energies and forces)."""
    from mdp_amd.fit import SnapFit, lsqsolve
    cfg = _batch(7)
    fit = SnapFit(cfg, J)
    fit.compute()
    K = fit.ncoeff
    beta = lattice.snap_coefficients(K, seed=J)[1:]   # drop the constant
    coeff = np.concatenate([[0.0], beta])
    sna = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0], coeff)
    E, F = [], []
    for x, box in cfg:
        B = np.diag(np.asarray(box, float))
        p = snapcpp.pairs_from_box(x, np.ones(len(x), np.int32), B[0], B[1], B[2], [1, 1, 1], sna)
        ea, fa, _ = snapcpp.snappotential_pairs(oracle, sna, p)
        E.append(ea.sum())
        F.append(fa)
    E, F = np.array(E), np.concatenate(F)
    L, h = fit.normal_equations(e=E, f=F)
    ce, cf, cef, cefs = lsqsolve(L[0], L[1], L[2], h[0], h[1], h[2])
    assert cefs is None   # no stress rows
    print(f"J={J}: cond(Lf) = {np.linalg.cond(L[1]):.3g}", flush=True)
    assert relerr(cf, beta) < 1e-6
    assert relerr(cef, beta) < 1e-6
    # the fitted coefficients reproduce the reference forces
    fits = fit.descriptors()
    pred = np.concatenate([np.einsum("cnk,k->nc", bd, cf) for _, bd, _ in fits])
    assert relerr(pred, F) < 1e-8
