"""Failure modes of the SNAP path reported as errors, never as GPU faults or silent truncation.

* More than 96 neighbours within rcut + skin (a lattice compressed to a = 2 A): the neighbour build stops
  with an error and counts the overflow, on the cell-list path and on the small-box path.  (The upstream
  cell-list builder truncates silently, cpuNeighborList3D.cpp:399-409.)
* Non-finite positions or velocities: rejected when the engine is created and when the host sets a new
  state, leaving the engine usable; a coordinate that runs away during a run (a velocity of 1e300 A/ps,
  finite but unphysical) is reported by the next neighbour rebuild.
* Coincident atoms: the drop-in pair filter drops pairs with r^2 <= 1e-20 (cpuSnap.cpp:278) exactly as the
  oracle does, and the MD engine's pair list does the same, so forces stay finite.
"""
import numpy as np
import pytest

from conftest import relerr
from mdp_amd import lattice, snapcpp

pytestmark = pytest.mark.gpu

MASS = 180.95


def _engine(x, box, v=None, J=4, every=10, dt=0.0005):
    from mdp_amd.md import SnapMD
    return SnapMD(x, box, np.ones(len(x), np.int32), J, lattice.snap_coefficients(14, seed=1), [0.5], [1.0], [MASS],
                  v=v, dt=dt, rebuild_every=every)


@pytest.mark.parametrize("ncell", [10, 4])       # cell lists / small-box brute force
def test_neighbour_overflow_is_an_error(ncell):
    x, box = lattice.crystal("bcc", ncell, 2.0, 0.0)
    md = _engine(x, box)
    print(f"overflow: bcc a=2.0, {len(x)} atoms", flush=True)
    with pytest.raises(RuntimeError, match="96 neighbours"):
        md.compute()
    assert md.stats()["overflows"] == 1


def test_non_finite_input_rejected():
    from mdp_amd.md import SnapMD
    x, box = lattice.crystal("bcc", 4, 3.316, 0.05, 1)
    bad = x.copy()
    bad[5, 1] = np.nan
    with pytest.raises(RuntimeError, match="non-finite"):
        SnapMD(bad, box, np.ones(len(x), np.int32), 4, lattice.snap_coefficients(14, seed=1), [0.5], [1.0], [MASS])
    v = np.zeros_like(x)
    v[3, 0] = np.inf
    with pytest.raises(RuntimeError, match="non-finite"):
        SnapMD(x, box, np.ones(len(x), np.int32), 4, lattice.snap_coefficients(14, seed=1), [0.5], [1.0], [MASS], v=v)
    md = _engine(x, box)
    md.compute()
    f0 = md.state()["f"]
    with pytest.raises(RuntimeError, match="non-finite"):
        md.set_state(bad)
    md.compute()                                         # unchanged and still usable
    assert np.array_equal(md.state()["f"], f0)
    md.set_state(x + 0.01)                               # a valid new state is taken
    md.compute()
    assert not np.array_equal(md.state()["f"], f0)


def test_runaway_position_reported_by_rebuild():
    x, box = lattice.crystal("bcc", 4, 3.316, 0.05, 2)
    v = np.zeros_like(x)
    v[7] = [1e300, 0.0, 0.0]
    md = _engine(x, box, v=v, every=1)
    md.compute()
    with pytest.raises(RuntimeError, match="runaway"):
        md.run(3)


def test_coincident_atoms_dropped(oracle):
    import mdp_amd
    from mdp_amd import neighbors
    x, box = lattice.crystal("bcc", 3, 3.316, 0.1, 4)
    x = np.concatenate([x, x[[5]]])                       # atom 54 sits on atom 5
    n = len(x)
    B = np.diag(box)
    y, alist, nl, nn = neighbors.fullneighborlist(x, B[0], B[1], B[2], [1, 1, 1], 4.97637)
    # the host list keeps the distance-0 entry; the pair filter must drop it
    assert any(np.all(y[j] == x[5]) for j in nl[nn[5]:nn[6]])
    t = np.ones(len(y), np.int32)
    rcutsq = np.full(4, 4.67637 ** 2)
    ilist = np.arange(n, dtype=np.int32)
    yy = np.ascontiguousarray(y.ravel())
    res = []
    libs = [mdp_amd.library(), oracle]
    for lib in libs:
        pn, pl = np.zeros(n, np.int32), np.zeros(len(nl), np.int32)
        lib.cpuNeighPairList(pn, pl, yy, rcutsq, ilist, nl, nn, t, alist, n, 3, 1)
        res.append((pn.copy(), pl.copy()))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    ps = np.concatenate([[0], np.cumsum(res[0][0])])
    for i in (5, 54):
        js = res[0][1][ps[i]:ps[i + 1]]
        assert not any(np.all(y[j] == x[i]) for j in js)
    # the MD engine: same pair set, finite forces, equal to the oracle on the coincident system
    md = _engine(x, box)
    md.compute()
    st = md.state()
    assert np.all(np.isfinite(st["f"]))
    s = snapcpp.initsna(oracle, [1, 4, 4.67637, 0.99363, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0],
                        lattice.snap_coefficients(14, seed=1))
    p = snapcpp.pairs_from_box(x, np.ones(n, np.int32), B[0], B[1], B[2], [1, 1, 1], s)
    # the Julia pair filter keeps r = 0 (neighpairlist.jl tests only dsq <= rcut^2); apply the C
    # library's r^2 > 1e-20 (cpuSnap.cpp:278), which the engine follows
    r = p["rij"].reshape(-1, 3)
    keep = np.sum(r * r, axis=1) > 1e-20
    assert (~keep).sum() == 2
    p = dict(p, rij=np.ascontiguousarray(r[keep].ravel()), ai=p["ai"][keep], aj=p["aj"][keep], ti=p["ti"][keep],
             tj=p["tj"][keep], ijnum=int(keep.sum()))
    e, f, w = snapcpp.snappotential_pairs(oracle, s, p)
    assert md.stats()["pairs"] == p["ijnum"]
    assert relerr(st["f"], f) < 1e-10 and relerr(st["eatom"], e) < 1e-10
