"""The training-data reader (mdp_amd.dataset) on the CPU: the committed Ta fixture against the upstream
JSON files when they are present, and the properties the upstream reader's rotation fixes (cells in the
a = (ax,0,0), b = (bx,by,0) form with lengths, angles and fractional coordinates unchanged)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from mdp_amd import dataset

SRC = "/root/reference/fitting/snap/Ta/JSON"


@pytest.fixture(scope="module")
def groups():
    return dataset.load_npz(os.path.join(GOLDEN, "ta_json.npz"))


def test_fixture_shape(groups):
    assert [g.name for g in groups] == ["Displaced_A15", "Displaced_BCC", "Displaced_FCC", "Elastic_BCC",
                                        "Elastic_FCC", "GSF_110", "GSF_112", "Liquid", "Surface", "Volume_A15",
                                        "Volume_BCC", "Volume_FCC"]
    assert sum(g.nconfigs for g in groups) == 363
    assert sum(int(g.natom.sum()) for g in groups) == 4224
    for g in groups:
        assert g.cell.shape == (g.nconfigs, 3, 3)
        assert np.all(g.cell[:, 0, 1:] == 0) and np.all(g.cell[:, 1, 2] == 0)     # a along x, b in xy
        assert np.all(g.t == 1)
    assert np.allclose(groups[7].weight, [467.0, 1.0, 1e-8])


def test_reader_matches_fixture_and_rotation_properties(groups):
    if not os.path.isdir(SRC):
        pytest.skip("upstream JSON data not present")
    import json
    g = dataset.read_folder(os.path.join(SRC, "Elastic_BCC"), ("Ta",), "json", True, (1e-8, 1e-8, 1e-4))
    ref = groups[3]
    assert np.array_equal(g.x, ref.x) and np.array_equal(g.cell, ref.cell) and np.array_equal(g.e, ref.e)
    # the first file in name order, read directly
    fn = sorted(os.listdir(os.path.join(SRC, "Elastic_BCC")))[0]
    t = open(os.path.join(SRC, "Elastic_BCC", fn)).read()
    d = json.loads(t[t.index("{"):])["Dataset"]["Data"][0]
    P = np.array(d["Lattice"])
    A = P[:, 0], P[:, 1], P[:, 2]              # transposelattice: the stored columns are the cell vectors
    a = g.cell[0]
    for u in range(3):
        assert abs(np.linalg.norm(a[u]) - np.linalg.norm(A[u])) < 1e-12
        for w in range(3):
            assert abs(a[u] @ a[w] - A[u] @ A[w]) < 1e-11
    X = np.array(d["Positions"])
    fr0 = np.linalg.solve(np.column_stack(A), X.T).T
    fr1 = np.linalg.solve(a.T, g.x[:len(X)].T).T
    assert np.max(np.abs(fr0 - fr1)) < 1e-12
    assert np.array_equal(g.f[:len(X)], np.array(d["Forces"]))           # forces as stored
    S = np.array(d["Stress"])
    assert np.array_equal(g.stress[0], [S[0, 0], S[1, 1], S[2, 2], S[1, 2], S[0, 2], S[0, 1]])
