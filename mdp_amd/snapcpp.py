"""Python mirror of the Julia SNAP driver src/Julia/Potential/snapcpp.jl (route R1).

The upstream driver ``ccall``s the 23 functions of the SNAP C library in a fixed order.  This module
does the same through ctypes, on whichever library it is given (``lib``, a :class:`_abi.SnapABI`):
by default the GPU library of this package, so that ``snappotential``/``snapdescriptors`` run the
HIP kernels behind the unchanged C entry points.  Array shapes, zero-initialisation, 0-based index
shifts, the ``bv`` row reorder and the parameter vector follow the cited upstream lines.
"""
import math

import numpy as np

from . import neighbors


def factable(n=168):
    """Factorial table handed to cpuInitSna (snapcpp.jl:64-241): n! as float64, n = 0..167."""
    return np.array([float(math.factorial(k)) for k in range(n)], dtype=np.float64)


class SnaStruct:
    """Fields of snastruct.jl:1-48 (SNAP parameters, index tables and coefficient vectors)."""

    def __repr__(self):
        return f"SnaStruct(twojmax={self.twojmax}, ntypes={self.ntypes}, ncoeff={self.ncoeff})"


def initializesna(lib, ntypes, twojmax, chemflag):
    """snastruct.jl:113-154."""
    s = SnaStruct()
    jdim = twojmax + 1
    s.nelements = ntypes if chemflag else 1
    s.idxb_max = lib.idxbcount(twojmax)
    s.idxu_max = lib.idxucount(twojmax)
    s.idxz_max = lib.idxzcount(twojmax)
    s.idxcg_max = lib.idxcgcount(twojmax)
    s.ntypes, s.twojmax = ntypes, twojmax
    s.ndoubles = s.nelements ** 2
    s.ntriples = s.nelements ** 3
    s.ncoeff = s.idxb_max * s.ntriples
    s.ncoeffall = s.ncoeff + 1
    s.map = np.zeros(ntypes + 1, np.int32)
    s.idxcg_block = np.zeros(jdim ** 3, np.int32)
    s.idxz_block = np.zeros(jdim ** 3, np.int32)
    s.idxb_block = np.zeros(jdim ** 3, np.int32)
    s.idxu_block = np.zeros(jdim, np.int32)
    s.idx_max = np.zeros(5, np.int32)
    s.idxz = np.zeros(s.idxz_max * 10, np.int32)
    s.idxb = np.zeros(s.idxb_max * 3, np.int32)
    s.rcutsq = np.zeros(ntypes * ntypes, np.float64)
    s.radelem = np.zeros(ntypes + 1, np.float64)
    s.wjelem = np.zeros(ntypes + 1, np.float64)
    s.bzero = np.zeros(jdim, np.float64)
    s.rootpqarray = np.zeros((twojmax + 2) ** 2, np.float64)
    s.cglist = np.zeros(s.idxcg_max, np.float64)
    return s


def initsna(lib, param, elemradius, elemweight, snapcoeff=None):
    """snapcpp.jl:243-314.  ``param`` = [ntypes, twojmax, rcutfac, rfac0, rmin0, bzeroflag,
    switchflag, quadraticflag, chemflag, bnormflag, wselfallflag]."""
    ntypes, twojmax = int(param[0]), int(param[1])
    rcutfac, rfac0, rmin0 = float(param[2]), float(param[3]), float(param[4])
    bzeroflag, switchflag, quadraticflag, chemflag, bnormflag, wselfallflag = (int(v) for v in param[5:11])
    s = initializesna(lib, ntypes, twojmax, chemflag)
    lib.cpuInitSna(s.rootpqarray, s.cglist, factable(), s.idx_max, s.idxz, s.idxz_block,
                   s.idxb, s.idxb_block, s.idxu_block, s.idxcg_block, twojmax)
    elemradius = np.asarray(elemradius, np.float64).ravel()
    elemweight = np.asarray(elemweight, np.float64).ravel()
    s.radelem[1:] = elemradius
    s.wjelem[1:] = elemweight
    s.map[1:] = np.arange(ntypes, dtype=np.int32)
    if snapcoeff is None:
        s.coeffelem = np.zeros(s.ncoeffall, np.float64)
    else:
        snapcoeff = np.ascontiguousarray(snapcoeff, np.float64).ravel()
        if snapcoeff.size != s.ntypes * s.ncoeffall:
            raise ValueError("number of snap coefficients is incompatible with the snap parameters: "
                             f"{snapcoeff.size} != {s.ntypes * s.ncoeffall}")
        s.coeffelem = snapcoeff
    s.rcutmax = max(2.0 * r * rcutfac for r in elemradius)
    for i in range(ntypes):
        for j in range(ntypes):
            cut = (elemradius[i] + elemradius[j]) * rcutfac
            s.rcutsq[j + i * ntypes] = cut * cut
    wself = 1.0
    www = wself * wself * wself
    for j in range(twojmax + 1):
        s.bzero[j] = www if bnormflag == 1 else www * (j + 1)
    s.nperdim = s.ncoeffall - 1
    s.bnormflag, s.chemflag, s.quadraticflag = bnormflag, chemflag, quadraticflag
    s.switchflag, s.bzeroflag, s.wselfallflag = switchflag, bzeroflag, wselfallflag
    s.wself, s.rmin0, s.rfac0, s.rcutfac = wself, rmin0, rfac0, rcutfac
    return s


def pairs_from_box(x, t, a, b, c, pbc, sna):
    """fullneighborlist -> neighpairlist -> neighpairs (snapcpp.jl:375-378), 0-based."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    n = x.shape[0]
    y, alist, neighlist, neighnum = neighbors.fullneighborlist(x, a, b, c, pbc, sna.rcutmax)
    ilist = np.arange(n, dtype=np.int32)
    atomtype = np.ascontiguousarray(np.asarray(t).ravel(), np.int32)
    pairlist, pairnum = neighbors.neighpairlist(y, ilist, alist, atomtype, neighlist, neighnum,
                                                sna.rcutsq, sna.ntypes)
    rij, ai, aj, ti, tj = neighbors.neighpairs(y, pairlist, pairnum, atomtype, ilist, alist)
    return dict(rij=rij, ai=ai, aj=aj, ti=ti, tj=tj, atomtype=atomtype, ilist=ilist, n=n,
                ijnum=int(pairnum[-1]))


def snappotential_pairs(lib, sna, p):
    """The C call chain of snapcpp.jl:386-423 on prepared pair arrays ``p`` (see pairs_from_box).
    Returns eatom (N,), fatom (N,3), vatom (N,6) -- the Julia (3,N)/(6,N) column-major arrays."""
    N, ijnum, nelem = p["n"], p["ijnum"], sna.nelements
    iu = sna.idxu_max
    Ur = np.zeros(N * iu * nelem)
    Ui = np.zeros(N * iu * nelem)
    lib.cpuSnapComputeUi(Ur, Ui, sna.rootpqarray, p["rij"], sna.wjelem, sna.radelem, sna.rmin0,
                         sna.rfac0, sna.rcutfac, sna.map, p["ai"], p["ti"], p["tj"], sna.twojmax,
                         iu, N, ijnum, sna.switchflag, sna.chemflag)
    lib.cpuAddWself2Ui(Ur, Ui, sna.wself, sna.idxu_block, p["atomtype"], sna.map, sna.wselfallflag,
                       sna.chemflag, iu, nelem, sna.twojmax, N)
    eatom = np.zeros(N)
    lib.cpuSnapComputeEi(eatom, Ur, Ui, sna.cglist, sna.bzero, sna.coeffelem, p["ilist"], sna.map,
                         p["atomtype"], sna.idxb, sna.idxcg_block, sna.idxu_block, sna.twojmax,
                         sna.idxb_max, iu, nelem, sna.ncoeffall, sna.bnormflag, sna.bzeroflag,
                         sna.wselfallflag, sna.chemflag, N)
    Yr = np.zeros(N * iu * nelem)
    Yi = np.zeros(N * iu * nelem)
    lib.cpuSnapComputeYi(Yr, Yi, Ur, Ui, sna.cglist, sna.coeffelem, sna.map, p["atomtype"], sna.idxz,
                         sna.idxb_block, sna.idxu_block, sna.idxcg_block, sna.twojmax, sna.idxb_max,
                         iu, sna.idxz_max, nelem, sna.ncoeffall, sna.bnormflag, sna.chemflag, N)
    fatom = np.zeros(3 * N)
    vatom = np.zeros(6 * N)
    lib.cpuSnapComputeFi(fatom, vatom, Yr, Yi, sna.rootpqarray, p["rij"], sna.wjelem, sna.radelem,
                         sna.rmin0, sna.rfac0, sna.rcutfac, sna.map, p["ai"], p["ai"], p["aj"], p["ti"],
                         p["tj"], sna.twojmax, iu, N, N, ijnum, sna.switchflag, sna.chemflag)
    return eatom, fatom.reshape(N, 3), vatom.reshape(N, 6)


def snappotential(lib, x, t, a, b, c, pbc, sna):
    """snapcpp.jl:323-426: per-atom energy, force and virial of the configuration."""
    return snappotential_pairs(lib, sna, pairs_from_box(x, t, a, b, c, pbc, sna))


def snapdescriptors_pairs(lib, sna, p, dim=3):
    """The C call chain of snapcpp.jl:599-664 on prepared pairs.  Returns bi (ntypes*ncoeff,),
    bd (ntypes*ncoeff, N, 3) and bv (ntypes*ncoeff, 6) with bv rows reordered [1,2,3,6,5,4]."""
    N, ijnum, nelem = p["n"], p["ijnum"], sna.nelements
    iu, iz, ib = sna.idxu_max, sna.idxz_max, sna.idxb_max
    n = iu * ijnum
    ur, ui = np.zeros(n), np.zeros(n)
    dur, dui = np.zeros(3 * n), np.zeros(3 * n)
    lib.cpuComputeUij(ur, ui, dur, dui, sna.rootpqarray, p["rij"], sna.wjelem, sna.radelem, sna.rmin0,
                      sna.rfac0, sna.rcutfac, sna.idxu_block, p["ti"], p["tj"], sna.twojmax, iu, ijnum,
                      sna.switchflag)
    Ur, Ui = np.zeros(N * iu * nelem), np.zeros(N * iu * nelem)
    lib.cpuZeroUarraytot2(Ur, Ui, sna.wself, sna.idxu_block, p["atomtype"], sna.map, p["ai"],
                          sna.wselfallflag, sna.chemflag, iu, nelem, sna.twojmax, N)
    lib.cpuAddUarraytot(Ur, Ui, ur, ui, sna.map, p["ai"], p["tj"], iu, N, ijnum, sna.chemflag)
    zr, zi = np.zeros(iz * sna.ndoubles * N), np.zeros(iz * sna.ndoubles * N)
    lib.cpuComputeZi2(zr, zi, Ur, Ui, sna.cglist, sna.idxz, sna.idxu_block, sna.idxcg_block,
                      sna.twojmax, iu, iz, nelem, sna.bnormflag, N)
    blist = np.zeros(ib * sna.ntriples * N)
    lib.cpuComputeBi2(blist, zr, zi, Ur, Ui, sna.bzero, p["ilist"], p["atomtype"], sna.map, sna.idxb,
                      sna.idxu_block, sna.idxz_block, sna.twojmax, ib, iu, iz, nelem, sna.bzeroflag,
                      sna.wselfallflag, sna.chemflag, N)
    dblist = np.zeros(ib * sna.ntriples * dim * ijnum)
    lib.cpuComputeDbidrj(dblist, zr, zi, dur, dui, sna.idxb, sna.idxu_block, sna.idxz_block, sna.map,
                         p["ai"], p["tj"], sna.twojmax, ib, iu, iz, nelem, sna.bnormflag, sna.chemflag,
                         N, ijnum)
    # Upstream sizes these tallies for ncoeff entries per type even with quadraticflag=1, where the
    # quadratic terms are written past the end (snapcpp.jl:652-658 vs cpuSnap.cpp:2281-2288); here
    # each type block holds the linear and the packed quadratic terms.
    nperdim = sna.ncoeff + (sna.ncoeff * (sna.ncoeff + 1) // 2 if sna.quadraticflag else 0)
    K = sna.ntypes * nperdim
    bi = np.zeros(K)
    lib.cpuSnapTallyBispectrum(bi, blist, p["ilist"], p["atomtype"], N, sna.ncoeff, nperdim,
                               sna.ntypes, sna.quadraticflag)
    bd = np.zeros(K * 3 * N)
    lib.cpuSnapTallyBispectrumDeriv(bd, blist, dblist, p["ai"], p["ai"], p["aj"], p["ti"], N, ijnum,
                                    sna.ncoeff, nperdim, sna.ntypes, sna.quadraticflag)
    bv = np.zeros(6 * K)
    lib.cpuSnapTallyBispectrumVirial(bv, blist, dblist, p["rij"], p["ai"], p["ti"], N, ijnum, sna.ncoeff,
                                     nperdim, sna.ntypes, sna.quadraticflag)
    bd = bd.reshape(K, N, 3)
    bv = bv.reshape(K, 6)[:, [0, 1, 2, 5, 4, 3]]
    return bi, bd, bv, dict(blist=blist, dblist=dblist, Ur=Ur, Ui=Ui, zr=zr, zi=zi)


def snapdescriptors(lib, x, t, a, b, c, pbc, sna):
    """snapcpp.jl:539-665: global bispectrum bi, its per-atom position derivative bd and virial bv."""
    bi, bd, bv, _ = snapdescriptors_pairs(lib, sna, pairs_from_box(x, t, a, b, c, pbc, sna))
    return bi, bd, bv
