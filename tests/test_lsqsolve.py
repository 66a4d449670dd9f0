"""mdp_lsqsolve (host code of libmdpsnap, no GPU): the four systems of lsqsolve (linearfit.jl:29-75)
against numpy's LU solve of the same perturbed matrices, and which systems are solved when inputs are
zero."""
import numpy as np
import pytest

from conftest import relerr


def _spd(rng, K, cond):
    q, _ = np.linalg.qr(rng.normal(size=(K, K)))
    return (q * np.logspace(0, np.log10(cond), K)) @ q.T


def _nudge(A, guard=False):
    A = A.copy()
    A[np.diag_indices_from(A)] *= 1 + 1e-15
    if guard and abs(A[0, 0]) <= 1e-10:
        A[0, 0] = 1e-10
    return A


def test_four_systems_match_numpy():
    from mdp_amd.fit import lsqsolve
    rng = np.random.default_rng(2)
    K = 14
    Le, Lf, Ls = (_spd(rng, K, 1e6) for _ in range(3))
    he, hf, hs = (rng.normal(size=K) for _ in range(3))
    ce, cf, cef, cefs = lsqsolve(Le, Lf, Ls, he, hf, hs)
    assert relerr(ce, np.linalg.solve(_nudge(Le), he)) < 1e-9
    assert relerr(cf, np.linalg.solve(_nudge(Lf, True), hf)) < 1e-9
    assert relerr(cef, np.linalg.solve(_nudge(Le + Lf), he + hf)) < 1e-9
    assert relerr(cefs, np.linalg.solve(_nudge(Le + Lf + Ls), he + hf + hs)) < 1e-9


def test_zero_groups_are_not_solved():
    from mdp_amd.fit import lsqsolve
    rng = np.random.default_rng(3)
    K = 5
    Lf, hf = _spd(rng, K, 10.0), rng.normal(size=K)
    Z, z = np.zeros((K, K)), np.zeros(K)
    ce, cf, cef, cefs = lsqsolve(Z, Lf, Z, z, hf, z)
    assert ce is None and cef is None and cefs is None
    assert relerr(cf, np.linalg.solve(_nudge(Lf, True), hf)) < 1e-12


def test_singular_matrix_reports_error():
    from mdp_amd.fit import lsqsolve
    K = 3
    L = np.zeros((K, K))
    L[0, 0] = 1.0   # rank 1: exactly singular after the nudge
    h = np.ones(K)
    with pytest.raises(RuntimeError, match="singular"):
        lsqsolve(L, np.zeros((K, K)), np.zeros((K, K)), h, np.zeros(K), np.zeros(K))
