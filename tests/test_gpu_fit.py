"""Batched fitting descriptors on the GPU (mdp_fit_*).

* Pair lists: bit-exact against the oracle's brute-force periodic builder (same (j, image) order),
  r_ij bit-exact against x_j + s*L - x_i, for boxes below, at and above twice the cutoff (the GPU
  takes a single-image path above it).
* Configurations without pairs: B - bzero of an isolated atom vanishes (closed form).
* B and dB/dr of every configuration: against the oracle's per-configuration chain cpuComputeUij ->
  cpuZeroUarraytot2 -> cpuAddUarraytot -> cpuComputeZi2 -> cpuComputeBi2 -> cpuComputeDbidrj on the
  same pairs, within 1e-10 relative to the max norm of the oracle array.
* At the full 512-configuration size: a sample of configurations against the oracle, and batch
  independence (the reversed batch gives bit-identical per-configuration outputs with the same dB/dr
  variant, since nothing is accumulated across atoms or configurations).
* The dB/dr variants the autotuner chooses from agree with each other to rounding.
* Graph replays after new positions (including pair-list growth past the buffers) equal a fresh run.
"""
import ctypes

import numpy as np
import pytest

from conftest import relerr
from mdp_amd import lattice, snapcpp

pytestmark = pytest.mark.gpu

RCUTFAC, RFAC0 = 4.67637, 0.99363
RC = 2 * 0.5 * RCUTFAC
TOL = 1e-10


def _small_batch(seed=7):
    cfg = []
    for k, (kind, n, a) in enumerate([("bcc", 2, 3.3), ("fcc", (2, 2, 3), 4.1), ("bcc", 3, 3.25),
                                      ("fcc", 1, 4.2), ("bcc", (1, 2, 2), 3.4),
                                      # box side 9.36 just above twice the cutoff 9.3527: one image per
                                      # pair, with pairs near half the box and near the cutoff
                                      ("bcc", 3, 3.12)]):
        x, box = lattice.crystal(kind, n, a, 0.1, seed + k)
        cfg.append((x, box))
    for k, n in enumerate((40, 9)):
        x, box = lattice.random_packing(n, 0.0554, 2.2, seed=seed + 10 + k)
        x = np.mod(x + np.random.default_rng(seed + 20 + k).normal(0, 0.1, x.shape), box)
        cfg.append((x, box))
    # configurations without pairs: one atom, and two atoms farther apart than the cutoff
    cfg.append((np.array([[3.0, 4.0, 5.0]]), np.array([20.0, 20.0, 20.0])))
    cfg.append((np.array([[1.0, 1.0, 1.0], [11.0, 11.0, 11.0]]), np.array([20.0, 20.0, 20.0])))
    return cfg


def _bruteforce(oracle, x, box):
    n = len(x)
    off = np.zeros(n + 1, np.int32)
    maxnb = 400 * n
    nj = np.zeros(maxnb, np.int32)
    ns = np.zeros(3 * maxnb, np.int32)
    f = oracle.cdll.oracle_neighbor_bruteforce
    f.restype = ctypes.c_long
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C")
    f.argtypes = [dp, ctypes.c_int, dp, ctypes.c_double, ip, ip, ip, ctypes.c_long]
    cnt = f(np.ascontiguousarray(x), n, np.ascontiguousarray(box), RC, off, nj, ns, maxnb)
    assert cnt >= 0
    return off, nj[:cnt], ns[:3 * cnt].reshape(-1, 3)


def _oracle_descriptors(oracle, J, x, ai, aj, rij, bnorm=0, bzero=1):
    n, P = len(x), len(ai)
    sna = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, bzero, 1, 0, 0, bnorm, 0], [0.5], [1.0])
    one = np.ones(P, np.int32)
    p = dict(rij=np.ascontiguousarray(rij).ravel(), ai=np.ascontiguousarray(ai, np.int32),
             aj=np.ascontiguousarray(aj, np.int32), ti=one, tj=one, atomtype=np.ones(n, np.int32),
             ilist=np.arange(n, dtype=np.int32), n=n, ijnum=P)
    _, _, _, raw = snapcpp.snapdescriptors_pairs(oracle, sna, p)
    K = sna.idxb_max
    return raw["blist"].reshape(K, n).T, raw["dblist"].reshape(K, 3, P).transpose(2, 1, 0)


def _check_config(oracle, fit, c, cfgs, b, db, ai, aj, rij, J, bnorm=0, bzero=1):
    x, box = cfgs[c][0], cfgs[c][1]
    bc, dbc, sl = fit.config(c, b, db)
    off, nj, ns = _bruteforce(oracle, x, box)
    # the pair list: same owners and images in the same order, r_ij bit-exact
    ci = np.repeat(np.arange(len(x)), np.diff(off))
    assert np.array_equal(ai[sl], ci)
    assert np.array_equal(aj[sl], nj)
    ref_rij = (x[nj] + ns * box) - x[ci]
    assert np.array_equal(rij[sl], ref_rij)
    ob, odb = _oracle_descriptors(oracle, J, x, ai[sl], aj[sl], rij[sl], bnorm, bzero)
    # B - bzero can vanish (an atom without pairs: U = identity, Z = identity, B = j + 1 = bzero), so
    # its error is measured against the size of B before the subtraction (at least J + 1)
    scale = max(np.max(np.abs(ob)), J + 1.0)
    assert np.max(np.abs(bc - ob)) < TOL * scale, c
    if len(odb):
        assert relerr(dbc, odb) < TOL, c
    else:   # no pairs: the closed form, B - bzero = 0 to rounding (with bzero subtracted)
        if bzero:
            assert np.max(np.abs(bc)) < 1e-13 * (J + 1), c


@pytest.mark.parametrize("J,variant", [(4, "0"), (6, "0"), (8, "0"), (4, "1"), (6, "1"), (8, "1"), (8, "2"),
                                       (4, "3"), (8, "3"), (4, "4"), (6, "4"), (8, "4"), (4, "5"), (6, "5"),
                                       (8, "5"), (4, "6"), (6, "6"), (8, "6")])
def test_small_batch_vs_oracle(oracle, monkeypatch, J, variant):
    from mdp_amd.fit import SnapFit
    monkeypatch.setenv("MDP_FIT_DBDR_VARIANT", variant)   # each dB/dr kernel variant in turn
    cfg = _small_batch()
    print(f"small batch J={J} dB/dr variant {variant}: {[len(c[0]) for c in cfg]} atoms", flush=True)
    fit = SnapFit(cfg, J)
    fit.compute()
    b, db = fit.raw()
    ai, aj, rij = fit.pairs()
    for c in range(len(cfg)):
        _check_config(oracle, fit, c, cfg, b, db, ai, aj, rij, J)


@pytest.mark.parametrize("waves", ["4", "8", "12"])
@pytest.mark.parametrize("J", [6, 8])
def test_z_kernel_waves_vs_oracle(oracle, monkeypatch, J, waves):
    """The Z + B kernel pinned to 4, 8 or 12 waves per workgroup (B and, through Z, dB/dr against the oracle)."""
    from mdp_amd.fit import SnapFit
    monkeypatch.setenv("MDP_FIT_ZB_WAVES", waves)
    monkeypatch.setenv("MDP_FIT_DBDR_VARIANT", "6")
    cfg = _small_batch()
    print(f"small batch J={J}, Z kernel {waves} waves", flush=True)
    fit = SnapFit(cfg, J)
    fit.compute()
    assert fit.tuning()["chosen"]["fit_zb_waves"] == int(waves)
    b, db = fit.raw()
    ai, aj, rij = fit.pairs()
    for c in range(len(cfg)):
        _check_config(oracle, fit, c, cfg, b, db, ai, aj, rij, J)


@pytest.mark.parametrize("J", [4, 8])
def test_dbdr_variants_agree(monkeypatch, J):
    """The autotuner may pick any dB/dr variant: B is the same kernel in all of them (bit for bit), and
    the dB/dr variants agree to rounding (the compiler contracts their term sums differently)."""
    from mdp_amd.fit import SnapFit
    cfg = _small_batch(9)
    out = {}
    for v in ("0", "3", "4", "5", "6"):
        monkeypatch.setenv("MDP_FIT_DBDR_VARIANT", v)
        fit = SnapFit(cfg, J)
        fit.compute()
        out[v] = fit.raw()
    for v in ("3", "4", "5", "6"):
        assert np.array_equal(out[v][0], out["0"][0]), v
        assert relerr(out[v][1], out["0"][1]) < 1e-13, v


@pytest.mark.parametrize("variant", [None, "6"])
def test_bnorm_and_no_bzero(oracle, monkeypatch, variant):
    from mdp_amd.fit import SnapFit
    if variant:
        monkeypatch.setenv("MDP_FIT_DBDR_VARIANT", variant)
    cfg = _small_batch(11)[:3]
    fit = SnapFit(cfg, 6, bnormflag=1, bzeroflag=0)
    fit.compute()
    b, db = fit.raw()
    ai, aj, rij = fit.pairs()
    for c in range(len(cfg)):
        _check_config(oracle, fit, c, cfg, b, db, ai, aj, rij, 6, bnorm=1, bzero=0)


def test_replay_after_new_positions_and_pair_growth(oracle):
    from mdp_amd.fit import SnapFit
    cfg = _small_batch(3)
    fit = SnapFit(cfg, 6)
    fit.compute()
    rng = np.random.default_rng(5)
    # new positions: small moves everywhere, and one configuration squeezed into a corner so that
    # the pair list outgrows the buffers sized by the first compute
    new = []
    for k, (x, box) in enumerate(cfg):
        y = np.mod(x + rng.normal(0, 0.05, x.shape), box)
        if k == 2:
            y = 0.45 * y
        new.append((y, box))
    fit.set_positions(np.concatenate([c[0] for c in new]))
    fit.compute()
    b, db = fit.raw()
    ai, aj, rij = fit.pairs()
    fresh = SnapFit(new, 6)
    fresh.compute()
    b2, db2 = fresh.raw()
    assert np.array_equal(b, b2) and np.array_equal(db, db2)
    for c in range(len(new)):
        _check_config(oracle, fit, c, new, b, db, ai, aj, rij, 6)


def test_fitting_batch_512(oracle, monkeypatch):
    from mdp_amd.fit import SnapFit
    batch = lattice.fitting_batch()
    cfg = [(x, box) for x, box, _ in batch]
    print(f"512-configuration batch: {sum(len(c[0]) for c in cfg)} atoms", flush=True)
    fit = SnapFit(cfg, 8)
    fit.compute()
    b, db = fit.raw()
    ai, aj, rij = fit.pairs()
    a_off, p_off = fit.layout()
    assert a_off[-1] == sum(len(c[0]) for c in cfg)
    assert np.all(np.isfinite(b)) and np.all(np.isfinite(db))
    # a sample: the smallest configuration of each kind and two seeded random ones below 300 atoms
    sizes = np.array([len(c[0]) for c in cfg])
    sample = {int(np.argmin(np.where(np.arange(len(cfg)) % 3 == k, sizes, 10**9))) for k in range(3)}
    rng = np.random.default_rng(1)
    small = np.nonzero(sizes < 300)[0]
    sample |= set(int(v) for v in rng.choice(small, 2, replace=False))
    for c in sorted(sample):
        print(f"  config {c}: {sizes[c]} atoms, {p_off[c + 1] - p_off[c]} pairs", flush=True)
        _check_config(oracle, fit, c, cfg, b, db, ai, aj, rij, 8)
    # batch independence: the reversed batch, with the dB/dr variant the autotuner chose above,
    # reproduces every configuration bit for bit
    chosen = fit.tuning().get("chosen", {}).get("fit_dbdr_variant")
    if chosen is not None:
        monkeypatch.setenv("MDP_FIT_DBDR_VARIANT", str(chosen))
    rev = SnapFit(cfg[::-1], 8)
    rev.compute()
    rb, rdb = rev.raw()
    C = len(cfg)
    for c in range(C):
        b1, d1, _ = fit.config(c, b, db)
        b2, d2, _ = rev.config(C - 1 - c, rb, rdb)
        assert np.array_equal(b1, b2) and np.array_equal(d1, d2), c
    print("  tuning:", chosen, "profile:", fit.profile(), flush=True)
