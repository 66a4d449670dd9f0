"""ZBL reference pair potential of a periodic system on the CPU, through the oracle (test infrastructure).

Pairs come from the oracle's brute-force full neighbour list at cut_outer (d^2 <= rc^2, every periodic
image), then the oracle's restatement of the upstream empirical pair route (cpuPotentials.cpp:1587-1649):
the pair gradient, the full-list energy/force decomposition, the pair virial with factor -1/2 and the
final force sign flip (:1837).  Returns per-atom energy (N,), force (N, 3) and virial (N, 6: xx, yy, zz,
xy, xz, yz).
"""
import ctypes

import numpy as np

TA_MU = np.array([4.0, 4.8, 73.0])    # snapTA06A.jl:62-66

_dp = np.ctypeslib.ndpointer(np.float64, flags="C")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C")


def bind(cdll, prefix="oracle_"):
    """Argument types of the ZBL pair route in the oracle (prefix "oracle_") or the compiled upstream
    core (prefix "r2_cpu", whose pair gradient has no mu argument: the generated code is for 4, 4.8, 73)."""
    g = getattr(cdll, prefix + "PairaGradient")
    g.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int] if prefix == "oracle_" else [_dp, _dp, _dp, ctypes.c_int]
    d = getattr(cdll, prefix + "FullNeighPairDecomposition")
    d.argtypes = [_dp, _dp, _dp, _dp, _ip, ctypes.c_int, ctypes.c_int]
    v = getattr(cdll, prefix + "VirialPairTally")
    v.argtypes = [_dp, _dp, _dp, ctypes.c_double, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    return g, d, v


def full_pairs(oracle, x, box, rc):
    """(ai, xij) of the full neighbour list within rc (every image), xij = x_j + shift - x_i."""
    n = len(x)
    off = np.zeros(n + 1, np.int32)
    maxnb = 160 * n
    nj = np.zeros(maxnb, np.int32)
    ns = np.zeros(3 * maxnb, np.int32)
    f = oracle.cdll.oracle_neighbor_bruteforce
    f.restype = ctypes.c_long
    f.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_double, _ip, _ip, _ip, ctypes.c_long]
    cnt = f(np.ascontiguousarray(x, np.float64), n, np.ascontiguousarray(box, np.float64), rc, off, nj, ns, maxnb)
    assert cnt >= 0
    ns = ns[:3 * cnt].reshape(-1, 3)
    ai = np.repeat(np.arange(n, dtype=np.int32), np.diff(off))
    xij = x[nj[:cnt]] + ns * np.asarray(box)[None, :] - x[ai]
    return ai, np.ascontiguousarray(xij)


def zbl_system(oracle, x, box, mu=TA_MU):
    g, d, v = bind(oracle.cdll)
    mu = np.ascontiguousarray(mu, np.float64)
    ai, xij = full_pairs(oracle, x, box, mu[1])
    P, n = len(ai), len(x)
    u, ux = np.zeros(P), np.zeros(3 * P)
    g(u, ux, xij.ravel(), mu, P)
    e, f = np.zeros(n), np.zeros(3 * n)
    d(e, f, u, ux, ai, P, 3)
    w = np.zeros(6 * n)
    v(w, ux, xij.ravel(), -0.5, ai, 3, n, P)
    return e, -f.reshape(n, 3), w.reshape(6, n).T.copy()
