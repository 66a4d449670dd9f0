"""Python handle on the batched fitting-descriptor engine of libmdpsnap (mdp_fit_* in include/mdpsnap.h).

It replaces the per-configuration descriptor loop of the upstream fitting drivers (snapcpp.jl:539-665
per configuration, cpuComputeUij -> cpuZeroUarraytot2 -> cpuAddUarraytot -> cpuComputeZi2 ->
cpuComputeBi2 -> cpuComputeDbidrj, cpuSnap.cpp:1620-2263): B of every atom and dB/dr of every pair of
many periodic configurations come out of one batched GPU pass.
"""
import ctypes
import json

import numpy as np

from . import library


class _Params(ctypes.Structure):
    _fields_ = [("twojmax", ctypes.c_int), ("ntypes", ctypes.c_int),
                ("rcutfac", ctypes.c_double), ("rfac0", ctypes.c_double), ("rmin0", ctypes.c_double),
                ("switchflag", ctypes.c_int), ("bzeroflag", ctypes.c_int), ("bnormflag", ctypes.c_int),
                ("elemradius", ctypes.POINTER(ctypes.c_double)), ("elemweight", ctypes.POINTER(ctypes.c_double))]


_bound = False


def _bind():
    global _bound
    c = library().cdll
    if _bound:
        return c
    vp = ctypes.c_void_p
    c.mdp_fit_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(_Params), ctypes.c_int, vp, vp, vp, vp]
    c.mdp_fit_create_cells.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(_Params), ctypes.c_int, vp, vp, vp, vp]
    for fn in ("mdp_fit_destroy", "mdp_fit_compute", "mdp_fit_sync"):
        getattr(c, fn).argtypes = [vp]
    c.mdp_fit_set_positions.argtypes = [vp, vp]
    c.mdp_fit_replay.argtypes = [vp, ctypes.c_int]
    c.mdp_fit_layout.argtypes = [vp, vp, vp]
    c.mdp_fit_get.argtypes = [vp, vp, vp]
    c.mdp_fit_pairs.argtypes = [vp, vp, vp, vp]
    c.mdp_fit_tuning.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    c.mdp_fit_profile.argtypes = [vp, vp]
    c.mdp_fit_descriptors.argtypes = [vp, vp, vp, vp]
    c.mdp_fit_normal_equations.argtypes = [vp] * 7 + [ctypes.c_int, vp, vp]
    c.mdp_lsqsolve.argtypes = [ctypes.c_int] + [vp] * 11
    _bound = True
    return c


def ncoeff_of(twojmax):
    return sum(1 for j1 in range(twojmax + 1) for j2 in range(j1 + 1)
               for j in range(j1 - j2, min(twojmax, j1 + j2) + 1, 2) if j >= j1)


class SnapFit:
    """B and dB/dr of a batch of periodic configurations.

    ``configs`` is a list of (x (N_c,3), box, [types]) with box either the three lengths of an orthogonal
    box or a (3, 3) cell whose rows are the cell vectors a = (Lx, 0, 0), b = (bx, Ly, 0), c = (cx, cy, Lz)
    (the form rotatelattice.jl produces); types are 1-based (default 1).
    """

    def __init__(self, configs, twojmax, elemradius=(0.5,), elemweight=(1.0,), rcutfac=4.67637,
                 rfac0=0.99363, rmin0=0.0, switchflag=1, bzeroflag=1, bnormflag=0):
        self._c = _bind()
        self.C = len(configs)
        self.natoms = np.array([len(cfg[0]) for cfg in configs], np.int32)
        self.x = np.ascontiguousarray(np.concatenate([np.asarray(cfg[0], np.float64).reshape(-1, 3)
                                                      for cfg in configs]))
        self.cells = None
        if any(np.asarray(cfg[1]).size == 9 for cfg in configs):
            self.cells = np.ascontiguousarray(np.stack([
                np.asarray(cfg[1], np.float64).reshape(3, 3) if np.asarray(cfg[1]).size == 9
                else np.diag(np.asarray(cfg[1], np.float64).reshape(3)) for cfg in configs]))
            self.boxes = np.ascontiguousarray(self.cells[:, [0, 1, 2], [0, 1, 2]])
        else:
            self.boxes = np.ascontiguousarray(np.stack([np.asarray(cfg[1], np.float64).reshape(3)
                                                        for cfg in configs]))
        self.types = np.ascontiguousarray(np.concatenate([
            (np.asarray(cfg[2], np.int32).ravel() if len(cfg) > 2 and not isinstance(cfg[2], str)
             else np.ones(len(cfg[0]), np.int32)) for cfg in configs]))
        self.twojmax = twojmax
        self.ncoeff = ncoeff_of(twojmax)
        self._keep = [np.ascontiguousarray(elemradius, np.float64).ravel(),
                      np.ascontiguousarray(elemweight, np.float64).ravel()]
        p = _Params(twojmax, len(self._keep[0]), rcutfac, rfac0, rmin0, switchflag, bzeroflag, bnormflag,
                    self._keep[0].ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                    self._keep[1].ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        h = ctypes.c_void_p()
        if self.cells is None:
            self._check(self._c.mdp_fit_create(ctypes.byref(h), ctypes.byref(p), self.C, self.natoms.ctypes.data,
                                               self.boxes.ctypes.data, self.x.ctypes.data, self.types.ctypes.data),
                        "mdp_fit_create")
        else:
            self._check(self._c.mdp_fit_create_cells(ctypes.byref(h), ctypes.byref(p), self.C,
                                                     self.natoms.ctypes.data, self.cells.ctypes.data,
                                                     self.x.ctypes.data, self.types.ctypes.data),
                        "mdp_fit_create_cells")
        self._h = h

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {library().cdll.mdp_snap_last_error().decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._c.mdp_fit_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def compute(self):
        self._check(self._c.mdp_fit_compute(self._h), "mdp_fit_compute")

    def set_positions(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(self.x.shape)
        self._check(self._c.mdp_fit_set_positions(self._h, x.ctypes.data), "mdp_fit_set_positions")

    def replay(self, n):
        self._check(self._c.mdp_fit_replay(self._h, int(n)), "mdp_fit_replay")

    def sync(self):
        self._check(self._c.mdp_fit_sync(self._h), "mdp_fit_sync")

    def layout(self):
        a = np.zeros(self.C + 1, np.int64)
        p = np.zeros(self.C + 1, np.int64)
        self._check(self._c.mdp_fit_layout(self._h, a.ctypes.data, p.ctypes.data), "mdp_fit_layout")
        return a, p

    def raw(self):
        """(blist, dblist) of the whole batch, configuration blocks in the upstream layouts."""
        a, p = self.layout()
        b = np.zeros(int(a[-1]) * self.ncoeff)
        db = np.zeros(int(p[-1]) * 3 * self.ncoeff)
        self._check(self._c.mdp_fit_get(self._h, b.ctypes.data, db.ctypes.data), "mdp_fit_get")
        return b, db

    def pairs(self):
        """(ai, aj, rij (P,3)) with configuration-local atom indices."""
        _, p = self.layout()
        P = int(p[-1])
        ai = np.zeros(max(P, 1), np.int32)
        aj = np.zeros(max(P, 1), np.int32)
        rij = np.zeros((max(P, 1), 3))
        self._check(self._c.mdp_fit_pairs(self._h, ai.ctypes.data, aj.ctypes.data, rij.ctypes.data), "mdp_fit_pairs")
        return ai[:P], aj[:P], rij[:P]

    def config(self, c, b=None, db=None):
        """B (N_c, ncoeff) and dB/dr (P_c, 3, ncoeff) of configuration c, plus its pair slice."""
        a, p = self.layout()
        if b is None:
            b, db = self.raw()
        K = self.ncoeff
        n = int(a[c + 1] - a[c])
        P = int(p[c + 1] - p[c])
        bc = b[int(a[c]) * K:int(a[c + 1]) * K].reshape(K, n).T
        dbc = db[int(p[c]) * 3 * K:int(p[c + 1]) * 3 * K].reshape(K, 3, P).transpose(2, 1, 0)
        return bc, dbc, slice(int(p[c]), int(p[c + 1]))

    def tuning(self):
        buf = ctypes.create_string_buffer(4096)
        self._check(self._c.mdp_fit_tuning(self._h, buf, 4096), "mdp_fit_tuning")
        return json.loads(buf.value.decode() or "{}")

    def profile(self):
        ms = np.zeros(4)
        self._check(self._c.mdp_fit_profile(self._h, ms.ctypes.data), "mdp_fit_profile")
        return dict(zip(("pairs", "utot", "zb", "dbdr"), (float(v) for v in ms)))

    def descriptors(self):
        """Per-configuration tallies of the last compute, as the upstream descriptor driver returns them
        (snapcpp.jl:645-664): a list of (bi (Kt,), bd (3, N_c, Kt), bv (6, Kt) in the row order xx, yy, zz,
        yz, xz, xy), Kt = ntypes * ncoeff."""
        kt = len(self._keep[0]) * self.ncoeff
        N = int(self.natoms.sum())
        bi = np.zeros(self.C * kt)
        bd = np.zeros(3 * N * kt)
        bv = np.zeros(6 * self.C * kt)
        self._check(self._c.mdp_fit_descriptors(self._h, bi.ctypes.data, bd.ctypes.data, bv.ctypes.data),
                    "mdp_fit_descriptors")
        out, a0 = [], 0
        for c in range(self.C):
            n = int(self.natoms[c])
            d = bd[3 * a0 * kt:3 * (a0 + n) * kt].reshape(kt, n, 3).transpose(2, 1, 0)
            v = bv[6 * c * kt:6 * (c + 1) * kt].reshape(kt, 6).T[[0, 1, 2, 5, 4, 3]]
            out.append((bi[c * kt:(c + 1) * kt].copy(), np.ascontiguousarray(d), np.ascontiguousarray(v)))
            a0 += n
        return out

    def normal_equations(self, e=None, f=None, s=None, we=None, wf=None, ws=None, normalizeenergy=0):
        """Weighted normal equations (L (3, Kt, Kt), h (3, Kt)) of the energy, force and stress rows
        (linearfit.jl:168-200): e[C] energies, f[N, 3] forces, s[C, 6] stresses in the order xx, yy, zz,
        yz, xz, xy; we, wf, ws per-configuration weights (default 1).  A missing target leaves its group
        zero."""
        kt = len(self._keep[0]) * self.ncoeff
        C, N = self.C, int(self.natoms.sum())

        def arr(a, n):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float64).ravel()
            if a.size != n:
                raise ValueError(f"expected {n} values, got {a.size}")
            return a
        e, f, s = arr(e, C), arr(f, 3 * N), arr(s, 6 * C)
        we, wf, ws = arr(we, C), arr(wf, C), arr(ws, C)
        ptr = lambda a: None if a is None else a.ctypes.data
        L = np.zeros((3, kt, kt))
        h = np.zeros((3, kt))
        self._check(self._c.mdp_fit_normal_equations(self._h, ptr(e), ptr(f), ptr(s), ptr(we), ptr(wf), ptr(ws),
                                                     int(normalizeenergy), L.ctypes.data, h.ctypes.data),
                    "mdp_fit_normal_equations")
        return L, h


def lsqsolve(Le, Lf, Ls, he, hf, hs):
    """lsqsolve of linearfit.jl:29-75: (ce, cf, cef, cefs), None for a system that is not solved."""
    c = _bind()
    K = len(he)
    arrs = [np.ascontiguousarray(a, np.float64) for a in (Le, Lf, Ls, he, hf, hs)]
    outs = [np.zeros(K) for _ in range(4)]
    solved = np.zeros(4, np.int32)
    rc = c.mdp_lsqsolve(K, *[a.ctypes.data for a in arrs], *[o.ctypes.data for o in outs], solved.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"mdp_lsqsolve: {c.mdp_snap_last_error().decode()}")
    return tuple(o if ok else None for o, ok in zip(outs, solved))
