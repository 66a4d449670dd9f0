"""Real-data linear SNAP fit of the upstream Ta example (fitting/snap/Ta/fitsnap.jl) on the GPU.

Data: the 12 folders of fitting/snap/Ta/JSON (363 configurations, 4,224 atoms), read with the upstream
reader's options (tests/golden/ta_json.npz, written by tests/golden/make_ta_dataset.py).  Model: the
single descriptor + twojmax=6 SNAP (bzeroflag 0) with the ZBL reference pair potential, energies
normalised per atom, stresses per volume, the per-folder weights of fitsnap.jl:18-29.

* Triclinic cells (the 200 Elastic configurations are tilted after rotatelattice) in the batched
  descriptor engine against the oracle's descriptor chain on the host pair list: bi, bd, bv within 1e-10.
* The reference pair potential of every configuration against the oracle's ZBL over brute-force pairs.
* The normal equations against A^T A of the explicit rows (1e-12).
* The error tables of validate2 against the table published in fitsnap.jl:107-159 (the only reference
  numbers for this path): MAE and RMSE within 1e-3 relative, RSQ within 1e-4, for every folder and the
  whole set (measured: within 3e-5, except GSF_110's force MAE and stress errors at 1.6e-4).  Three entries are left out: the per-folder force errors of Volume_A15/BCC/FCC.  Those
  folders are perfect lattices whose stored and fitted forces vanish by symmetry (here 1e-14), and the
  published values (RMSE 6.25, 6.07, 2.89 eV/A over 720, 126 and 372 rows) cannot be reconciled with
  the published force RMSE of the whole set, 0.16312 eV/A over 12,672 rows, which they would push above
  1.4 eV/A; the whole-set row itself is matched.
"""
import os

import numpy as np
import pytest

import zbl_oracle
from conftest import GOLDEN, relerr
from mdp_amd import dataset, snapcpp

pytestmark = pytest.mark.gpu

RCUTFAC, RFAC0 = 4.67637, 0.99363

# fitsnap.jl:107-159: (MAE, RMSE, RSQ) per folder, the whole set first
PUBLISHED = {
    "energy": [
        (0.11265174691267, 0.37977565433503, 0.99419038677079),
        (0.00235792938939, 0.00246855977428, -1.2156363735745),
        (0.00188371436433, 0.00273064614363, 0.99587522319726),
        (0.00061741229597, 0.00075009168170, 0.97160166434659),
        (0.00599579098947, 0.00604563796454, 0.99230919870267),
        (0.00354421105172, 0.00445843637691, 0.99333159772842),
        (0.00652425814607, 0.00883023587747, 0.77951696510583),
        (0.00868251914489, 0.00961453139000, 0.87318809800944),
        (0.00566811969259, 0.00655584081006, 0.99601358415720),
        (0.01352360826034, 0.01571242406414, 0.98816795126889),
        (0.24578877837149, 0.34490495333075, 0.99781261498874),
        (0.32857920069640, 0.51271284190945, 0.99853838108547),
        (0.81209055282782, 1.18123504852115, 0.98443646179057)],
    "force": [
        (0.08359239990607, 0.16311584659554, 0.98512903562289),
        (0.09806643270309, 0.12820169392009, 0.94957955563905),
        (0.14339589323270, 0.18486360986945, 0.98517811287921),
        (0.05740637316446, 0.07166197131771, 0.98061307246834),
        (0.06439647995478, 0.07899854689009, 0.41298668776769),
        (0.05065258662499, 0.06490610225971, 0.63719969979068),
        (0.02651057125534, 0.05418215537551, 0.99889872481223),
        (0.05698006876321, 0.10025780716990, 0.99737013744671),
        (0.35379720579374, 0.48364455518089, 0.94779367837188),
        (0.04760235280733, 0.10895743153042, 0.99671811972583),
        None, None, None],
    "stress": [
        (68331.2692226880, 381756.218018856, 0.93961426214254),
        (19553.8208834055, 26783.3681374323, 0.99755940235506),
        (14933.0136738189, 20558.4972892687, 0.99870935725821),
        (35323.5274494020, 49684.5017070339, 0.99113396279263),
        (170.405533015481, 268.553751189739, 0.99999977908789),
        (48132.1295201740, 66784.2466913797, 0.98419798844541),
        (1623.00572832491, 2322.56048399292, 0.99993839179953),
        (2068.90488220449, 3116.39773656099, 0.99987644651027),
        (42920.1434033482, 58562.0314191833, 0.98628360736784),
        (2309.28255361132, 4277.50368379242, 0.99977822976790),
        (132939.655632498, 357432.429798387, 0.95261560122466),
        (271466.858168366, 770712.743079258, 0.97424571846356),
        (304211.704271920, 1079230.30094437, 0.79235217127926)],
}


@pytest.fixture(scope="module")
def groups():
    return dataset.load_npz(os.path.join(GOLDEN, "ta_json.npz"))


@pytest.fixture(scope="module")
def tafit(groups):
    from mdp_amd import linearfit
    print("Ta dataset: descriptors, reference pair potential, normal equations", flush=True)
    f = linearfit.LinearSnapFit(groups)
    f.solve()
    return f


def _oracle_descriptors(oracle, x, cell, J=6):
    s = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 0, 1, 0, 0, 0, 0], [0.5], [1.0])
    p = snapcpp.pairs_from_box(x, np.ones(len(x), np.int32), cell[0], cell[1], cell[2], [1, 1, 1], s)
    bi, bd, bv, _ = snapcpp.snapdescriptors_pairs(oracle, s, p)
    return bi, bd, bv


def test_triclinic_batched_descriptors_vs_oracle(oracle, groups):
    from mdp_amd.fit import SnapFit
    from mdp_amd import lattice
    el = groups[3]                                   # Elastic_BCC: tilted after rotation
    cfgs = [el.config(c)[:2] for c in (0, 7, 42)]
    assert all(abs(cfg[1][1, 0]) > 0 for cfg in cfgs)
    # and a strongly tilted random cell
    x, box = lattice.crystal("bcc", 3, 3.3, 0.1, 5)
    cell = np.array([[box[0], 0, 0], [1.1, box[1], 0], [-0.9, 0.7, box[2]]])
    frac = x / box
    cfgs.append((frac @ cell, cell))
    fit = SnapFit(cfgs, 6, bzeroflag=0)
    fit.compute()
    got = fit.descriptors()
    for (x, cell), (bi, bd, bv) in zip(cfgs, got):
        rbi, rbd, rbv = _oracle_descriptors(oracle, x, cell)
        assert relerr(bi, rbi) < 1e-10
        assert relerr(bd, rbd.transpose(2, 1, 0)) < 1e-10
        assert relerr(bv, rbv.T) < 1e-10


def _brute_zbl(oracle, x, cell, mu=zbl_oracle.TA_MU):
    """The oracle's pair route over every image (s1, s2, s3) in [-4, 4]^3 within cut_outer."""
    g, d, v = zbl_oracle.bind(oracle.cdll)
    n = len(x)
    s = np.array(np.meshgrid(*[np.arange(-4, 5)] * 3, indexing="ij")).reshape(3, -1).T
    shifts = s @ cell
    ai, xij = [], []
    for i in range(n):
        dd = (x[None, :, :] + shifts[:, None, :] - x[i]).reshape(-1, 3)
        r2 = np.sum(dd * dd, axis=1)
        ok = (r2 <= mu[1] ** 2) & (r2 > 0)
        ai.append(np.full(ok.sum(), i, np.int32))
        xij.append(dd[ok])
    ai, xij = np.concatenate(ai), np.ascontiguousarray(np.concatenate(xij))
    P = len(ai)
    u, ux = np.zeros(P), np.zeros(3 * P)
    g(u, ux, xij.ravel(), np.asarray(mu, np.float64), P)
    e, f, w = np.zeros(n), np.zeros(3 * n), np.zeros(6 * n)
    d(e, f, u, ux, ai, P, 3)
    v(w, ux, xij.ravel(), -0.5, ai, 3, n, P)
    return e.sum(), -f.reshape(n, 3), w.reshape(6, n).sum(axis=1)[[0, 1, 2, 5, 4, 3]]


def test_reference_pair_potential_vs_oracle(oracle, groups):
    from mdp_amd import linearfit
    picks = [(0, 0), (3, 5), (7, 1), (8, 2), (10, 0), (11, 3)]   # A15, tilted BCC, liquid, surface, volumes
    cfgs = [groups[g].config(c)[:2] for g, c in picks]
    E, F, V = linearfit.reference_pair_potential(cfgs)
    for k, (x, cell) in enumerate(cfgs):
        e0, f0, v0 = _brute_zbl(oracle, x, cell)
        assert abs(E[k] - e0) < 1e-10 * max(1.0, abs(e0))
        assert np.max(np.abs(F[k] - f0)) < 1e-10 * max(1.0, np.max(np.abs(f0)))
        assert relerr(V[k], v0) < 1e-10


def test_normal_equations_vs_explicit_rows(tafit, groups):
    L, h = tafit.normal_equations()
    for k in range(3):
        Lr = np.zeros_like(L[k])
        hr = np.zeros_like(h[k])
        for g in range(len(groups)):
            r = tafit.rows(g)
            A, b = ((r.Ae, r.be), (r.Af, r.bf), (r.As, r.bs))[k]
            w = groups[g].weight[k]
            Lr += w * w * A.T @ A
            hr += w * w * A.T @ b
        assert relerr(L[k], Lr) < 1e-12
        assert relerr(h[k], hr) < 1e-12


def test_fit_errors_match_published_table(tafit, groups):
    te, tf, ts = tafit.validate()
    print("coefficients:", np.array2string(tafit.coeff, precision=6), flush=True)
    worst = 0.0
    for key, tab in (("energy", te), ("force", tf), ("stress", ts)):
        for row, (got, pub) in enumerate(zip(tab, PUBLISHED[key])):
            if pub is None:
                continue
            name = "ALL" if row == 0 else groups[row - 1].name
            assert abs(got[0] - pub[0]) < 1e-3 * abs(pub[0]), (key, name, "MAE", got[0], pub[0])
            assert abs(got[1] - pub[1]) < 1e-3 * abs(pub[1]), (key, name, "RMSE", got[1], pub[1])
            assert abs(got[2] - pub[2]) < 1e-4, (key, name, "RSQ", got[2], pub[2])
            worst = max(worst, abs(got[1] - pub[1]) / abs(pub[1]))
    print(f"largest relative RMSE difference from the published table: {worst:.2e}", flush=True)
    # the perfect-lattice folders (rows 10-12): stored forces zero, fitted forces zero by symmetry
    assert np.all(tf[10:13, 1] < 1e-10)
