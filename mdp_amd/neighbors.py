"""Host-side neighbour and pair construction of the Julia SNAP driver (route R1).

Mirrors src/Julia/Potential/fullneighborlist.jl:1-38 (ghost images + full list at rcutmax, test
``d <= rcutmax``), neighpairlist.jl:29-64 (per-type-pair filter ``dsq <= rcutsq``) and
neighpairs.jl:46-79 (pair arrays with ``aj`` = owner of the neighbour).  Indices are 0-based here;
the upstream Julia code is 1-based and shifts to 0-based before calling the C library
(snapcpp.jl:380-384).

Difference kept on purpose: upstream replicates the lattice with ``m = round(rcutmax/a)`` images per
direction when rcutmax exceeds a box length (latticecoords.jl:7-20), which misses images for boxes
between rcutmax/1.5 and rcutmax; here every periodic image within rcutmax is included, so the result
is the complete set {(i, owner j, image shift) : |x_j + shift - x_i| <= rcutmax}.

This is the reference-faithful host path used by the drop-in driver (``snapcpp.py``) on small
systems.  The MD engine builds its lists on the GPU (cell lists, ``libmdpsnap``).
"""
import numpy as np

try:
    from scipy.spatial import cKDTree
except ImportError:  # pragma: no cover - scipy is part of the image
    cKDTree = None


def _cell_matrix(a, b, c):
    H = np.column_stack([np.asarray(a, float)[:3], np.asarray(b, float)[:3], np.asarray(c, float)[:3]])
    return H


def ghost_atoms(x, a, b, c, pbc, rcut):
    """Extended coordinates: the N owned atoms followed by every periodic image (owner-major) whose
    distance to the box is at most ``rcut``.  Returns (y (M,3), alist (M,), shifts (M,3) int)."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    H = _cell_matrix(a, b, c)
    Hinv = np.linalg.inv(H)
    n = x.shape[0]
    # perpendicular widths of the cell
    vol = abs(np.linalg.det(H))
    cols = [H[:, 0], H[:, 1], H[:, 2]]
    width = np.array([vol / np.linalg.norm(np.cross(cols[1], cols[2])),
                      vol / np.linalg.norm(np.cross(cols[2], cols[0])),
                      vol / np.linalg.norm(np.cross(cols[0], cols[1]))])
    pbc = np.asarray(pbc, dtype=int).ravel()[:3]
    frac = x @ Hinv.T
    delta = rcut / width
    reach = [int(np.ceil(delta[d] + 1.0)) if pbc[d] else 0 for d in range(3)]
    shifts = np.array([(sx, sy, sz)
                       for sz in range(-reach[2], reach[2] + 1)
                       for sy in range(-reach[1], reach[1] + 1)
                       for sx in range(-reach[0], reach[0] + 1)
                       if (sx, sy, sz) != (0, 0, 0)], dtype=np.int64).reshape(-1, 3)
    ys, owners, sh = [x], [np.arange(n)], [np.zeros((n, 3), np.int64)]
    if len(shifts):
        # candidate images, owner-major
        f = frac[:, None, :] + shifts[None, :, :]
        keep = np.all((f >= -delta) & (f <= 1.0 + delta), axis=2)
        oi, si = np.nonzero(keep)
        ys.append(x[oi] + shifts[si] @ H.T)
        owners.append(oi)
        sh.append(shifts[si])
    y = np.ascontiguousarray(np.concatenate(ys), dtype=np.float64)
    return y, np.concatenate(owners).astype(np.int32), np.concatenate(sh).astype(np.int32)


def fullneighborlist(x, a, b, c, pbc, rcutmax):
    """fullneighborlist.jl:1-38.  Returns (y, alist, neighlist, neighnum): for owned atom i its
    neighbours are neighlist[neighnum[i]:neighnum[i+1]], indices into y in ascending order."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    y, alist, _ = ghost_atoms(x, a, b, c, pbc, rcutmax)
    if cKDTree is not None:
        tree = cKDTree(y)
        cand = tree.query_ball_point(x, r=rcutmax * (1.0 + 1e-9))
    else:  # brute force
        cand = [np.arange(len(y))] * n
    lists = []
    for i in range(n):
        g = np.sort(np.asarray(cand[i], dtype=np.int64))
        d = y[g] - x[i]
        dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
        g = g[(dist <= rcutmax) & (g != i)]
        lists.append(g)
    neighnum = np.zeros(n + 1, dtype=np.int32)
    neighnum[1:] = np.cumsum([len(g) for g in lists])
    neighlist = (np.concatenate(lists) if lists else np.zeros(0)).astype(np.int32)
    return y, alist, neighlist, neighnum


def neighpairlist(y, ilist, alist, atomtype, neighlist, neighnum, rcutsq, ntypes):
    """neighpairlist.jl:29-64 (0-based): keep neighbours with dsq <= rcutsq[(jtype-1) + (itype-1)*ntypes]."""
    pairs, counts = [], []
    for i in ilist:
        g = neighlist[neighnum[i]:neighnum[i + 1]]
        jt = atomtype[alist[g]]
        rsq = rcutsq[(jt - 1) + (atomtype[i] - 1) * ntypes]
        d = y[g] - y[i]
        dsq = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        sel = g[dsq <= rsq]
        pairs.append(sel)
        counts.append(len(sel))
    pairnum = np.zeros(len(ilist) + 1, dtype=np.int32)
    pairnum[1:] = np.cumsum(counts)
    pairlist = (np.concatenate(pairs) if pairs else np.zeros(0)).astype(np.int32)
    return pairlist, pairnum


def neighpairs(y, pairlist, pairnum, atomtype, ilist, alist):
    """neighpairs.jl:46-79 (0-based): rij = y_j - y_i (AoS, 3 per pair), ai, aj (owner), ti, tj."""
    n = pairnum[-1]
    ai = np.repeat(np.asarray(ilist, np.int32), np.diff(pairnum)).astype(np.int32)
    g = pairlist[:n]
    aj = alist[g].astype(np.int32)
    rij = np.ascontiguousarray((y[g] - y[ai]).reshape(-1), dtype=np.float64)
    ti = atomtype[ai].astype(np.int32)
    tj = atomtype[aj].astype(np.int32)
    return rij, ai, aj, ti, tj
