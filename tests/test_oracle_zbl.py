"""The oracle's ZBL reference pair potential (the Ta application's empirical term beside SNAP) pinned
without a GPU: against the upstream generated pair gradient and pair tallies (golden vectors written by
tests/golden/make_golden_zbl.py from the compiled upstream sources; the compiled core itself on fresh
inputs when it is built here), against central finite differences of its own energy, and against the
properties the switching construction fixes (u and du/dr vanish at cut_outer; the inner region is the
bare screened Coulomb term shifted by a constant).
"""
import ctypes
import os

import numpy as np
import pytest

import zbl_oracle
from conftest import GOLDEN, ROOT, relerr

R2 = os.path.join(ROOT, "oracle", "_ref", "libcpucore_r2.so")


def _pair(oracle, xij, mu=zbl_oracle.TA_MU):
    g, _, _ = zbl_oracle.bind(oracle.cdll)
    xij = np.ascontiguousarray(xij, np.float64).reshape(-1, 3)
    u, ux = np.zeros(len(xij)), np.zeros(3 * len(xij))
    g(u, ux, xij.ravel(), np.ascontiguousarray(mu, np.float64), len(xij))
    return u, ux.reshape(-1, 3)


def test_zbl_vs_upstream_golden(oracle):
    gd = dict(np.load(os.path.join(GOLDEN, "zbl_ta.npz"), allow_pickle=False))
    u, ux = _pair(oracle, gd["xij"])
    assert relerr(u, gd["u"]) < 1e-13
    assert relerr(ux, gd["u_xij"]) < 1e-13
    e, f, w = zbl_oracle.zbl_system(oracle, gd["x"], gd["box"])
    assert relerr(e, gd["e"]) < 1e-13
    assert relerr(f, gd["f"]) < 1e-13
    assert relerr(w, gd["vsoa"].T) < 1e-13


def test_zbl_vs_compiled_upstream(oracle):
    if not os.path.exists(R2):
        pytest.skip("upstream C++ core not built (oracle/_ref)")
    ref = ctypes.CDLL(R2)
    g, d, v = zbl_oracle.bind(ref, prefix="r2_cpu")
    rng = np.random.default_rng(5)
    n = 4000
    dirs = rng.normal(size=(n, 3))
    xij = np.ascontiguousarray(dirs / np.linalg.norm(dirs, axis=1)[:, None] * rng.uniform(0.3, 4.8, n)[:, None])
    u0, ux0 = np.zeros(n), np.zeros(3 * n)
    g(u0, ux0, xij.ravel(), n)
    u, ux = _pair(oracle, xij)
    assert relerr(u, u0) < 1e-13 and relerr(ux, ux0.reshape(-1, 3)) < 1e-13
    # the tallies on a random pair set with repeated centres
    ai = np.sort(rng.integers(0, 50, n)).astype(np.int32)
    gO, dO, vO = zbl_oracle.bind(oracle.cdll)
    for lib_d, lib_v, tag in ((d, v, "ref"), (dO, vO, "oracle")):
        e, f, w = np.zeros(50), np.zeros(150), np.zeros(300)
        lib_d(e, f, u0, ux0, ai, n, 3)
        lib_v(w, ux0, xij.ravel(), -0.5, ai, 3, 50, n)
        if tag == "ref":
            ref_out = (e, f, w)
    for a, b in zip((e, f, w), ref_out):
        assert np.array_equal(a, b)


def test_zbl_gradient_finite_differences(oracle):
    rng = np.random.default_rng(9)
    dirs = rng.normal(size=(64, 3))
    r = np.concatenate([rng.uniform(1.0, 4.75, 60), [3.98, 4.02, 4.3, 4.7]])
    x0 = dirs / np.linalg.norm(dirs, axis=1)[:, None] * r[:, None]
    _, ux = _pair(oracle, x0)
    h = 1e-6
    fd = np.zeros_like(x0)
    for c in range(3):
        d = np.zeros_like(x0)
        d[:, c] = h
        fd[:, c] = (_pair(oracle, x0 + d)[0] - _pair(oracle, x0 - d)[0]) / (2 * h)
    assert relerr(ux, fd) < 1e-7
    # a sign or factor error in the tanh-step term shows near cut_inner, where that term peaks
    near = np.abs(r - 4.0) < 0.05
    assert np.max(np.abs(ux[near] - fd[near])) < 1e-7 * np.max(np.abs(fd))


def test_zbl_switching_properties(oracle):
    mu = zbl_oracle.TA_MU
    e = np.array([[0.0, 0.0, mu[1]]])
    u, ux = _pair(oracle, e)
    assert abs(u[0]) < 1e-12 and abs(ux[0, 2]) < 1e-10        # u and du/dr vanish at cut_outer
    # inside cut_inner the switching term is the constant sw5 (tanh step fully off): u - E(r) is the
    # same at any two radii there.  E(r) = zze/r sum c_k exp(-d_k a r) with the Ta constants.
    ainv = 2 * 73.0 ** 0.23 / 0.4685
    c = np.array([0.02817, 0.28022, 0.50986, 0.18175])
    d = np.array([0.20162, 0.40290, 0.94229, 3.19980]) * ainv
    rr = np.array([1.0, 2.0, 3.0, 3.5])
    E = 73.0 * 73.0 * 14.399645 / rr * (np.exp(-np.outer(rr, d)) @ c)
    uu, _ = _pair(oracle, np.stack([np.zeros(4), np.zeros(4), rr], axis=1))
    sw5 = uu - E
    assert np.ptp(sw5) < 1e-12 * np.max(np.abs(uu))
    assert abs(sw5[0] - (-0.0163)) < 5e-4                      # the value ZBLpotential.jl:90 quotes
    assert np.allclose(d, [2.3090, 4.6140, 10.7912, 36.6444], atol=1e-4)   # ZBLpotential.jl:35


def test_zbl_system_invariants(oracle):
    from mdp_amd import lattice
    x, box = lattice.crystal("bcc", 3, 2.9, 0.15, 3)
    e, f, w = zbl_oracle.zbl_system(oracle, x, box)
    assert np.max(np.abs(f.sum(axis=0))) < 1e-11 * np.max(np.abs(f))
    # total virial = -1/2 sum over the full list of xij (x) du/dxij
    ai, xij = zbl_oracle.full_pairs(oracle, x, box, 4.8)
    _, ux = _pair(oracle, xij)
    W = -0.5 * np.einsum("pa,pb->ab", xij, ux)
    tot = w.sum(axis=0)
    assert relerr(tot, [W[0, 0], W[1, 1], W[2, 2], W[0, 1], W[0, 2], W[1, 2]]) < 1e-12
    # energy is half the sum of pair energies of the full list
    u, _ = _pair(oracle, xij)
    assert abs(e.sum() - 0.5 * u.sum()) < 1e-12 * abs(u).sum()
