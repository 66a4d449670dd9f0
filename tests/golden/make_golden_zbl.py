"""Writes tests/golden/zbl_ta.npz from the upstream C++ core compiled from source (oracle/_ref/
libcpucore_r2.so, which includes the generated ZBL pair gradient of the Ta application,
src/C++/Potentials/cpuPairaGradient.cpp, and the pair tallies of src/C++/Core).

Inputs and outputs only: (1) pair vectors at distances spread over [0.5, 4.8] A with u and du/dxij of the
generated gradient; (2) a compressed 3x3x3-cell BCC Ta box (a = 2.9 A, so that the ZBL core is sampled)
with its full neighbour list at 4.8 A, the per-pair values, and the energy/force/virial tallies of the
upstream full-list route (cpuFullNeighPairDecomposition, cpuVirialPairTally with factor -1/2, force sign
flipped as cpuPotentials.cpp:1837).

    python tests/golden/make_golden_zbl.py [outdir]
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from mdp_amd import _abi, lattice  # noqa: E402
import zbl_oracle  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "libcpucore_r2.so")
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")


def main(outdir):
    ref = ctypes.CDLL(REF)
    g, d, v = zbl_oracle.bind(ref, prefix="r2_cpu")
    rng = np.random.default_rng(73)
    ng = 512
    dirs = rng.normal(size=(ng, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    r = np.concatenate([rng.uniform(0.5, 4.8, ng - 8), [3.999, 4.0, 4.001, 4.4, 4.79, 4.7999, 4.8, 2.0]])
    xij = np.ascontiguousarray(dirs * r[:, None])
    u, ux = np.zeros(ng), np.zeros(3 * ng)
    g(u, ux, xij.ravel(), ng)
    oracle = _abi.bind(ORACLE)
    x, box = lattice.crystal("bcc", 3, 2.9, 0.12, 74)
    ai, pij = zbl_oracle.full_pairs(oracle, x, box, 4.8)
    P, n = len(ai), len(x)
    pu, pux = np.zeros(P), np.zeros(3 * P)
    g(pu, pux, pij.ravel(), P)
    e, f = np.zeros(n), np.zeros(3 * n)
    d(e, f, pu, pux, ai, P, 3)
    w = np.zeros(6 * n)
    v(w, pux, pij.ravel(), -0.5, ai, 3, n, P)
    path = os.path.join(outdir, "zbl_ta.npz")
    np.savez_compressed(path, xij=xij, u=u, u_xij=ux.reshape(ng, 3), x=x, box=box, ai=ai, pij=pij, pu=pu,
                        pux=pux.reshape(P, 3), e=e, f=-f.reshape(n, 3), vsoa=w.reshape(6, n),
                        mu=zbl_oracle.TA_MU)
    print(f"wrote {path}: {ng} pair vectors, {n} atoms / {P} pairs")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else HERE)
