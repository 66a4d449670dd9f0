"""Linear SNAP fit of a training set with a reference pair potential, as the upstream fitting example runs
it (fitting/snap/Ta/fitsnap.jl: linearfit2 + validate2 of src/Julia/Optimization/linearfit.jl:355-542 and
maespace.jl:210-306, rows of src/Julia/Potential/emldescriptors.jl:154-266).

Design-matrix columns: the "single" descriptor (one per atom, fitting/snap/Ta/single.jl) followed by the
K SNAP bispectrum components.  Rows per configuration c with N_c atoms:
  energy  [N_c, bi_c] / N_c  ->  (E_c - Eref_c) / N_c            (normalizeenergy)
  force   [0, bd_c]          ->  F_c - Fref_c        (3 N_c rows)
  stress  [0, bv_c] k_c      ->  S_c - Sref_c k_c   (6 rows, k_c = 1.6021765e6 / volume, normalizestress)
where Eref, Fref and Sref are the reference pair potential's energy, forces and virial (the ZBL term of
fitsnap.jl:84-86, computed on the GPU: mdp_hipPairaGradient and the full-list tallies, over the host
full neighbour list the upstream Julia code builds the same way, empiricalpotential.jl:1-103).

GPU: SNAP descriptors of all configurations in one batch (mdp_fit_create_cells, any periodic cell), the
reference pair potential, and the weighted normal equations of the SNAP block (mdp_fit_normal_equations).
Host: the single-descriptor column of the normal equations (a sum over the energy rows), lsqsolve
(mdp_lsqsolve) and the error tables (calerrors, linearfit.jl:101-117).
"""
import ctypes

import numpy as np

from . import library, neighbors
from .devarray import DevArray
from .fit import SnapFit, lsqsolve

NKTV2P = 1.6021765e6      # eV/A^3 -> bar (emldescriptors.jl:203)
ZBL_TA = (4.0, 4.8, 73.0)  # fitsnap.jl:84-85


def _bind_pairpot():
    c = library().cdll
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    c.mdp_hipPairaGradient.argtypes = [vp] * 12 + [i] * 7
    c.mdp_hipFullNeighPairDecomposition.argtypes = [vp] * 5 + [i, i]
    c.mdp_hipVirialPairTally.argtypes = [vp] * 3 + [d, vp, i, i, i]
    for f in ("mdp_hipPairaGradient", "mdp_hipFullNeighPairDecomposition", "mdp_hipVirialPairTally"):
        getattr(c, f).restype = None
    return c


def _check(c, what):
    msg = c.mdp_snap_last_error().decode()
    if msg:
        raise RuntimeError(f"{what}: {msg}")


def reference_pair_potential(configs, mu=ZBL_TA):
    """Energy (C,), forces (list of (N_c, 3)) and virial (C, 6; xx, yy, zz, yz, xz, xy) of the ZBL pair
    potential of every configuration (x, cell rows a, b, c), as empiricalpotential.jl evaluates it: full
    neighbour list at cut_outer, e_i = sum_j u/2, F_i = sum_j du/dxij, W = -1/2 sum xij (x) du/dxij."""
    c = _bind_pairpot()
    ai_all, xij_all, nat = [], [], []
    a0 = 0
    for x, cell in configs:
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        y, alist, nl, nn = neighbors.fullneighborlist(x, cell[0], cell[1], cell[2], [1, 1, 1], mu[1])
        ai = np.repeat(np.arange(len(x)), np.diff(nn))
        xij = y[nl] - x[ai]
        d2 = xij[:, 0] * xij[:, 0] + xij[:, 1] * xij[:, 1] + xij[:, 2] * xij[:, 2]
        keep = d2 <= mu[1] * mu[1]          # neighpairlist.jl: dsq <= rcut^2
        ai_all.append(ai[keep] + a0)
        xij_all.append(xij[keep])
        nat.append(len(x))
        a0 += len(x)
    ai = np.ascontiguousarray(np.concatenate(ai_all), np.int32)
    xij = np.ascontiguousarray(np.concatenate(xij_all), np.float64)
    P, N = len(ai), a0
    dx, dai = DevArray.from_numpy(xij.ravel(), np.float64), DevArray.from_numpy(ai, np.int32)
    dmu = DevArray.from_numpy(np.asarray(mu, np.float64), np.float64)
    du, dux = DevArray.zeros(P), DevArray.zeros(3 * P)
    de, df, dv = DevArray.zeros(N), DevArray.zeros(3 * N), DevArray.zeros(6 * N)
    c.mdp_hipPairaGradient(du.ptr, dux.ptr, dx.ptr, None, None, None, None, None, None, dmu.ptr, None, None, 3, 0, 3,
                           0, 0, P, 1)
    _check(c, "mdp_hipPairaGradient")
    c.mdp_hipFullNeighPairDecomposition(de.ptr, df.ptr, du.ptr, dux.ptr, dai.ptr, P, 3)
    _check(c, "mdp_hipFullNeighPairDecomposition")
    c.mdp_hipVirialPairTally(dv.ptr, dux.ptr, dx.ptr, -0.5, dai.ptr, 3, N, P)
    _check(c, "mdp_hipVirialPairTally")
    e, f, v = de.numpy(), -df.numpy().reshape(N, 3), dv.numpy().reshape(6, N)
    off = np.concatenate([[0], np.cumsum(nat)])
    E = np.array([e[off[k]:off[k + 1]].sum() for k in range(len(nat))])
    F = [f[off[k]:off[k + 1]] for k in range(len(nat))]
    V = np.array([v[:, off[k]:off[k + 1]].sum(axis=1)[[0, 1, 2, 5, 4, 3]] for k in range(len(nat))])
    return E, F, V


class Rows:
    """Unweighted design rows and targets of one group (emldescriptors2): Ae (C, 1+K), Af (3N, 1+K),
    As (6C, 1+K), be, bf, bs."""

    def __init__(self, Ae, Af, As, be, bf, bs):
        self.Ae, self.Af, self.As, self.be, self.bf, self.bs = Ae, Af, As, be, bf, bs


def calerrors(A, b, c):
    """linearfit.jl:101-117: (mae, rmse, rsq, ssr, avg) of b - A c; -1 for empty rows."""
    if len(b) == 0 or len(c) == 0 or A.size == 0:
        return -1.0, -1.0, -1.0, -1.0, -1.0
    res = b - A @ c
    mae = np.sum(np.abs(res)) / len(b)
    ssr = float(np.sum(res * res))
    rmse = np.sqrt(ssr / len(b))
    avg = np.sum(b) / len(b)
    rsq = 1.0 - ssr / np.sum((b - avg) ** 2)
    return mae, rmse, rsq, ssr, avg


class LinearSnapFit:
    """The fit of fitsnap.jl on a list of dataset.Frames groups (one per data folder, with its weights)."""

    def __init__(self, groups, twojmax=6, rcutfac=4.67637, rfac0=0.99363, rmin0=0.0, bzeroflag=0, switchflag=1,
                 bnormflag=0, elemradius=(0.5,), elemweight=(1.0,), pair_potential=ZBL_TA, normalizeenergy=True,
                 normalizestress=True):
        self.groups = groups
        self.normalizeenergy, self.normalizestress = normalizeenergy, normalizestress
        configs, self.group_of = [], []
        for g, grp in enumerate(groups):
            for c in range(grp.nconfigs):
                x, cell, t = grp.config(c)
                configs.append((x, cell, t))
                self.group_of.append(g)
        self.group_of = np.array(self.group_of)
        self.configs = configs
        self.natom = np.array([len(cfg[0]) for cfg in configs])
        self.vol = np.array([abs(np.linalg.det(cfg[1])) for cfg in configs])
        self.fit = SnapFit(configs, twojmax, elemradius=elemradius, elemweight=elemweight, rcutfac=rcutfac,
                           rfac0=rfac0, rmin0=rmin0, switchflag=switchflag, bzeroflag=bzeroflag, bnormflag=bnormflag)
        self.fit.compute()
        self.K = self.fit.ncoeff * len(elemradius)
        self.desc = self.fit.descriptors()
        if pair_potential is not None:
            self.eref, self.fref, self.sref = reference_pair_potential([(cfg[0], cfg[1]) for cfg in configs],
                                                                       pair_potential)
        else:
            C = len(configs)
            self.eref, self.fref, self.sref = np.zeros(C), [np.zeros((n, 3)) for n in self.natom], np.zeros((C, 6))

    def _targets(self):
        E = np.concatenate([g.e for g in self.groups])
        F = np.concatenate([g.f for g in self.groups])
        S = np.concatenate([g.stress for g in self.groups])
        return E, F, S

    def rows(self, g):
        """emldescriptors2 rows of group g."""
        E, F, S = self._targets()
        idx = np.nonzero(self.group_of == g)[0]
        a0 = np.concatenate([[0], np.cumsum(self.natom)])
        Ae, Af, As, be, bf, bs = [], [], [], [], [], []
        for c in idx:
            n = self.natom[c]
            bi, bd, bv = self.desc[c]
            k = NKTV2P / self.vol[c] if self.normalizestress else 1.0
            d = np.concatenate([[n], bi])
            Ae.append(d / n if self.normalizeenergy else d)
            Af.append(np.column_stack([np.zeros(3 * n), bd.transpose(1, 0, 2).reshape(3 * n, -1)]))
            As.append(np.column_stack([np.zeros(6), k * bv]))
            e = E[c] - self.eref[c]
            be.append(e / n if self.normalizeenergy else e)
            bf.append((F[a0[c]:a0[c + 1]] - self.fref[c]).ravel())
            bs.append(S[c] - k * self.sref[c])
        return Rows(np.array(Ae), np.concatenate(Af), np.concatenate(As), np.array(be), np.concatenate(bf),
                    np.concatenate(bs))

    def normal_equations(self):
        """Le, Lf, Ls ((1+K)^2) and he, hf, hs of linearfit2 with the per-folder weights: the SNAP block
        from the GPU, the single column from the energy rows."""
        E, F, S = self._targets()
        C = len(self.configs)
        w = np.array([self.groups[g].weight for g in self.group_of])      # (C, 3)
        k = NKTV2P / self.vol if self.normalizestress else np.ones(C)
        Fr = np.concatenate(self.fref)
        L, h = self.fit.normal_equations(e=E - self.eref, f=(F - Fr).ravel(), s=S / k[:, None] - self.sref,
                                         we=w[:, 0], wf=w[:, 1], ws=w[:, 2] * k,
                                         normalizeenergy=1 if self.normalizeenergy else 0)
        K = self.K
        Lx = np.zeros((3, K + 1, K + 1))
        hx = np.zeros((3, K + 1))
        Lx[:, 1:, 1:] = L
        hx[:, 1:] = h
        # the single descriptor: 1 per atom in energy rows (N_c, or 1 after normalisation), 0 elsewhere
        one = np.ones(C) if self.normalizeenergy else self.natom.astype(float)
        bi = np.array([self.desc[c][0] / (self.natom[c] if self.normalizeenergy else 1.0) for c in range(C)])
        be = (E - self.eref) / (self.natom if self.normalizeenergy else 1.0)
        w2 = w[:, 0] ** 2
        Lx[0, 0, 0] = np.sum(w2 * one * one)
        Lx[0, 0, 1:] = Lx[0, 1:, 0] = (w2 * one) @ bi
        hx[0, 0] = np.sum(w2 * one * be)
        return Lx, hx

    def solve(self):
        """Coefficients of the energy + force + stress fit (lossfunc "energyforcestress", method "lsq")."""
        L, h = self.normal_equations()
        ce, cf, cef, cefs = lsqsolve(L[0], L[1], L[2], h[0], h[1], h[2])
        if cefs is None:
            raise RuntimeError("linear fit: the energy + force + stress system was not solved")
        self.coeff = cefs
        return cefs

    def validate(self, coeff=None):
        """validate2 (maespace.jl:210-306): energy, force and stress error tables, (n_groups + 1, 3) arrays
        of MAE, RMSE, RSQ with the whole set first."""
        coeff = self.coeff if coeff is None else coeff
        n = len(self.groups)
        tabs = []
        stats = {k: [] for k in ("e", "f", "s")}
        for g in range(n):
            r = self.rows(g)
            for key, A, b in (("e", r.Ae, r.be), ("f", r.Af, r.bf), ("s", r.As, r.bs)):
                stats[key].append((calerrors(A, b, coeff), b))
        for key in ("e", "f", "s"):
            out = -np.ones((n + 1, 3))
            sz = np.array([len(b) for _, b in stats[key]], float)
            err = np.array([s for s, _ in stats[key]])      # (n, 5): mae rmse rsq ssr avg
            ball = np.concatenate([b for _, b in stats[key]])
            if sz.sum() > 0:
                avg = np.sum(err[:, 4] * sz) / sz.sum()
                out[0, 0] = np.sum(err[:, 0] * sz) / sz.sum()
                out[0, 1] = np.sqrt(np.sum(err[:, 1] ** 2 * sz) / sz.sum())
                out[0, 2] = 1.0 - np.sum(err[:, 3]) / np.sum((ball - avg) ** 2)
                out[1:] = err[:, :3]
            tabs.append(out)
        return tuple(tabs)
