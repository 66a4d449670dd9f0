"""Host side of the batched fitting path: the seeded configuration generator (counts, kinds, boxes,
minimum separation of the random packings before displacement) and its determinism."""
import numpy as np

from mdp_amd import lattice


def test_fitting_batch_shape_and_determinism():
    a = lattice.fitting_batch(nconfig=24, seed=5)
    b = lattice.fitting_batch(nconfig=24, seed=5)
    assert len(a) == 24
    for (x1, b1, k1), (x2, b2, k2) in zip(a, b):
        assert k1 == k2 and np.array_equal(x1, x2) and np.array_equal(b1, b2)
    kinds = [k for _, _, k in a]
    assert kinds[:6] == ["bcc", "fcc", "liquid", "bcc", "fcc", "liquid"]
    for x, box, kind in a:
        assert 2 <= len(x) <= 2200
        assert np.all(x >= 0) and np.all(x < box)
        if kind == "bcc":
            assert len(x) % 2 == 0 and 3.2 * 0.99 < (2 * np.prod(box) / len(x)) ** (1 / 3) < 3.45 * 1.01
        elif kind == "fcc":
            assert len(x) % 4 == 0 and 4.0 * 0.99 < (4 * np.prod(box) / len(x)) ** (1 / 3) < 4.3 * 1.01
        else:
            assert abs(len(x) / np.prod(box) - 0.0554) < 1e-9


def test_random_packing_min_separation():
    from scipy.spatial import cKDTree
    x, box = lattice.random_packing(600, 0.0554, 2.2, seed=9)
    assert len(cKDTree(x, boxsize=box).query_pairs(2.2)) == 0
    assert abs(600 / np.prod(box) - 0.0554) < 1e-12


def test_supercell_counts_track_the_target():
    for n in (16, 17, 100, 777, 2000):
        cells = lattice._supercell_for(n, 2)
        assert max(cells) <= 2 * min(cells)
        assert abs(2 * np.prod(cells) - n) <= max(2, 0.25 * n)
