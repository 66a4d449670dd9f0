"""Learned surrogate on the GPU.

* MLP kernels (BF16 MFMA, FP32 accumulate) against the float64 reference of tests/mlp_ref.py with the
  same bfloat16 rounding points: energies and descriptor gradients within 1e-2, parameter gradients
  within 2e-2, relative to the max norm (the residual is FP32 accumulation order and double rounding
  FP32 -> BF16 of a few activations); Adam exactly as the reference formula (1e-6).
* Descriptor chain rule: GPU surrogate forces against central finite differences (h = 1e-5 A) of the
  float64 surrogate energy built from the oracle's twojmax=4 descriptors (within 5e-2 of the max force:
  the GPU path is BF16), and the float64 chain rule itself against the same differences (1e-6).
* Training: on plain-SNAP data the trained network's held-out force error is below the untrained
  network's and below the linear twojmax=4 SNAP least-squares fit (the method without the network).
"""
import numpy as np
import pytest

import mlp_ref as R
from conftest import relerr
from mdp_amd import lattice, snapcpp

pytestmark = pytest.mark.gpu

RCUTFAC, RFAC0 = 4.67637, 0.99363


def _rand_model(seed, nin=14):
    rng = np.random.default_rng(seed)
    p = R.init_params(seed, nin)
    p[1024:1088] = rng.normal(0, 0.3, 64)
    p[5184:5248] = rng.normal(0, 0.3, 64)
    p[5312] = 0.1
    aff = np.zeros(34)
    aff[:nin] = rng.normal(0, 1, nin)
    aff[16:16 + nin] = rng.uniform(0.5, 2.0, nin)
    aff[32], aff[33] = 1.3, -2.0
    return p.astype(np.float32), aff, rng


def _run_eval(p, aff, nin, X):
    from mdp_amd import surrogate as S
    from mdp_amd.devarray import DevArray
    c = S._lib()
    n = len(X)
    tp = DevArray.from_numpy(p, np.float32)
    ta = DevArray.from_numpy(aff, np.float64)
    tq = DevArray(10240, np.int16)
    tx = DevArray.from_numpy(np.ascontiguousarray(X.T), np.float64)
    te = DevArray.zeros(n)
    tg = DevArray.zeros(nin * n)
    S._chk(c.mdp_mlp_pack(tp.ptr, tq.ptr, None), "pack")
    S._chk(c.mdp_mlp_eval(tq.ptr, tp.ptr, ta.ptr, nin, tx.ptr, n, te.ptr, tg.ptr, None), "eval")
    return te.numpy(), tg.numpy().reshape(nin, n).T, (c, tp, ta, tq, tx)


@pytest.mark.parametrize("n,nin", [(1003, 14), (16, 5), (4096, 16)])
def test_mlp_eval_vs_reference(n, nin):
    p, aff, rng = _rand_model(n + nin, nin)
    X = rng.normal(0, 2.0, (n, nin))
    print(f"mlp eval n={n} nin={nin}", flush=True)
    e, g, _ = _run_eval(p, aff, nin, X)
    er, gr = R.forward(p, aff, nin, X, quant=True)
    assert relerr(e, er) < 1e-2
    assert relerr(g, gr) < 1e-2
    e64, g64 = R.forward(p, aff, nin, X, quant=False)
    assert relerr(e, e64) < 5e-2 and relerr(g, g64) < 5e-2


def test_mlp_train_grad_and_adam_vs_reference():
    from mdp_amd import surrogate as S
    from mdp_amd.devarray import DevArray
    n, nin, nwaves = 2000, 14, 37
    p, aff, rng = _rand_model(11, nin)
    X = rng.normal(0, 2.0, (n, nin))
    U = rng.normal(0, 1.0, (n, nin))
    alpha = rng.normal(0, 1.0, n)
    _, _, (c, tp, ta, tq, tx) = _run_eval(p, aff, nin, X)
    tu = DevArray.from_numpy(np.ascontiguousarray(U.T), np.float64)
    tal = DevArray.from_numpy(alpha, np.float64)
    part = DevArray.zeros(nwaves * S.NPARAM, np.float32)
    grad = DevArray.zeros(S.NPARAM, np.float32)
    print("mlp train grad", flush=True)
    S._chk(c.mdp_mlp_train(tq.ptr, tp.ptr, ta.ptr, nin, tx.ptr, tu.ptr, tal.ptr, n, part.ptr, nwaves, grad.ptr, None),
           "train")
    g = grad.numpy().astype(np.float64)
    gr = R.train_grad(p, aff, nin, X, U, alpha, quant=True)
    for name, (a, b) in dict(W1=(0, 1024), b1=(1024, 1088), W2=(1088, 5184), b2=(5184, 5248), w3=(5248, 5312),
                             b3=(5312, 5313)).items():
        assert relerr(g[a:b], gr[a:b]) < 2e-2, name
    # Adam on the GPU gradient
    m = DevArray.zeros(S.NPARAM, np.float32)
    v = DevArray.zeros(S.NPARAM, np.float32)
    p0 = tp.numpy().astype(np.float64)
    for t in (1, 2):
        S._chk(c.mdp_mlp_adam(tp.ptr, grad.ptr, m.ptr, v.ptr, 1e-3, 0.9, 0.999, 1e-8, t, 0.01, None), "adam")
    pr, mr, vr = R.adam(p0, g, np.zeros_like(g), np.zeros_like(g), 1e-3, 0.9, 0.999, 1e-8, 1, 0.01)
    pr, mr, vr = R.adam(pr, g, mr, vr, 1e-3, 0.9, 0.999, 1e-8, 2, 0.01)
    assert np.max(np.abs(tp.numpy() - pr)) < 1e-6


def _cpu_energy_and_chain_forces(oracle, x, box, p, aff, nin):
    """float64 surrogate energy of a configuration from the oracle's twojmax=4 descriptors, and the
    forces by the descriptor chain rule F_i = sum_pairs(i,j) g_i.dB_i/dr_ij - sum_pairs(a,i) g_a.dB_a/dr_ai."""
    sna = snapcpp.initsna(oracle, [1, 4, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0])
    cell = [[box[0], 0, 0], [0, box[1], 0], [0, 0, box[2]]]
    pr = snapcpp.pairs_from_box(x, np.ones(len(x)), *cell, [1, 1, 1], sna)
    _, _, _, raw = snapcpp.snapdescriptors_pairs(oracle, sna, pr)
    n, P, K = len(x), pr["ijnum"], sna.idxb_max
    B = raw["blist"].reshape(K, n).T
    db = raw["dblist"].reshape(K, 3, P)
    e, g = R.forward(p, aff, nin, B)
    s = np.einsum("kp,kcp->pc", g[pr["ai"]].T, db)
    F = np.zeros((n, 3))
    np.add.at(F, pr["ai"], s)
    np.add.at(F, pr["aj"], -s)
    return e.sum(), F


def test_chain_rule_forces_vs_finite_differences(oracle):
    from mdp_amd import surrogate as S
    x, box = lattice.crystal("bcc", 2, 3.316, 0.15, 21)
    rng = np.random.default_rng(2)
    p, aff, _ = _rand_model(5)
    aff[:14] = 0.0
    aff[16:30] = 0.2
    E0, Fc = _cpu_energy_and_chain_forces(oracle, x, box, p, aff, 14)
    h = 1e-5
    Ffd = np.zeros_like(x)
    for i in range(len(x)):
        for c in range(3):
            d = np.zeros_like(x)
            d[i, c] = h
            Ep, _ = _cpu_energy_and_chain_forces(oracle, x + d, box, p, aff, 14)
            Em, _ = _cpu_energy_and_chain_forces(oracle, x - d, box, p, aff, 14)
            Ffd[i, c] = -(Ep - Em) / (2 * h)
    assert relerr(Fc, Ffd) < 1e-6
    # the GPU path on the same configuration
    cfg = [dict(x=x, box=box, E=E0, F=Ffd)]
    ds = S.DeviceSet(cfg)
    ev = S.Evaluator(ds, S.Model(p, aff, 14))
    ev.forward()
    Fg = ev.F.numpy().reshape(-1, 3)
    # BF16 activations leave ~1% error on dE/dB; forces sum pair terms that partly cancel
    assert relerr(Fg, Ffd) < 5e-2
    assert abs(float(ev.Ec.numpy()[0]) - E0) < 2e-2 * max(1.0, abs(E0))


def test_training_beats_untrained_and_linear():
    from mdp_amd import surrogate as S
    print("generating plain-SNAP training data", flush=True)
    data, _ = S.make_training_set(nconfig=120, ncell=3, per_run=20, every=10, equil=20, tmin=300, tmax=1500, seed=5)
    perm = np.random.default_rng(3).permutation(len(data))
    train, test = [data[i] for i in perm[:96]], [data[i] for i in perm[96:]]
    ds_tr, ds_te = S.DeviceSet(train), S.DeviceSet(test)
    mu, sg, es, esh = S.affine_from(ds_tr)
    m0 = S.Model.init(mu, sg, es, esh, seed=1)
    f0, _ = S.evaluate(ds_te, m0)
    m1, hist = S.train(ds_tr, m0, steps=1500, lr=3e-3, lr_end=2e-4, log=print, every=250)
    f1, e1 = S.evaluate(ds_te, m1)
    fl, el, _ = S.linear_baseline(ds_tr, ds_te)
    print(f"held-out force RMSE: untrained {f0:.3f}, linear J4 {fl:.3f}, MLP {f1:.3f} eV/A", flush=True)
    assert np.isfinite(f1) and f1 < f0
    assert f1 < fl


@pytest.mark.parametrize("kind", ["bcc", "packing"])
def test_md_engine_surrogate_vs_chain_rule(oracle, kind):
    """The MD engine's surrogate force call (twojmax=4 descriptors through the LDS-staged B and per-atom
    Y kernels, BF16 MLP, generator dE/dr) against the float64 chain rule on the oracle's descriptors."""
    from mdp_amd import surrogate as S
    from mdp_amd.md import SnapMD
    if kind == "bcc":
        x, box = lattice.crystal("bcc", 3, 3.316, 0.15, 21)
    else:
        x, box = lattice.random_packing(40, 0.0554, 2.2, seed=9)
    p, aff, _ = _rand_model(5)
    aff[:14] = 0.0
    aff[16:30] = 0.2
    E0, Fc = _cpu_energy_and_chain_forces(oracle, x, box, p, aff, 14)
    s8 = snapcpp.initsna(oracle, [1, 8, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0])
    coeff = lattice.snap_coefficients(s8.ncoeff, seed=3)
    md = SnapMD(x, box, np.ones(len(x), np.int32), 8, coeff, [0.5], [1.0], [180.95])
    md.set_surrogate(S.Model(p, aff, 14), audit_every=0)
    print(f"engine surrogate, {kind}: {len(x)} atoms", flush=True)
    md.compute()
    st = md.state()
    assert relerr(st["f"], Fc) < 5e-2
    assert abs(st["eatom"].sum() - E0) < 2e-2 * max(1.0, abs(E0))


@pytest.mark.parametrize("pot", ["Ta", "W"])
def test_md_engine_surrogate_vs_training_path(oracle, pot):
    """The MD engine's surrogate forces (adjoint route: B, MLP, per-atom beta, Y, dE/dr) against the training
    path's forces for the same model and configuration (stored twojmax=4 dB/dr contracted with the same
    MLP's dE/dB).  Both run the same BF16 network on FP64 descriptors and differ only in the FP64 order of
    the chain rule (measured: 2e-15 of the max force)."""
    import os
    from mdp_amd import surrogate as S
    from mdp_amd.md import SnapMD
    if pot == "W":
        P = S.tungsten(os.path.join(os.path.dirname(__file__), "golden", "snapW.coeff"), with_zbl=False)
    else:
        P = S.seeded_tantalum()
    x, box = lattice.crystal("bcc", 5, P.a0, 0.12, 31)
    x = np.ascontiguousarray(np.delete(x, [7], axis=0))
    p, aff, _ = _rand_model(5)
    aff[:14] = 0.0
    aff[16:30] = 0.2
    model = S.Model(p, aff, 14)
    ds = S.DeviceSet([dict(x=x, box=box, E=0.0, F=np.zeros_like(x))], rcutfac=P.rcutfac, rfac0=P.rfac0)
    ev = S.Evaluator(ds, model)
    ev.forward()
    Ft = ev.F.numpy().reshape(-1, 3)
    md = SnapMD(x, box, np.ones(len(x), np.int32), 8, P.coeff, [0.5], [1.0], [P.mass], rcutfac=P.rcutfac,
                rfac0=P.rfac0)
    md.set_surrogate(model, audit_every=0)
    print(f"engine vs training path, {pot}: {len(x)} atoms", flush=True)
    md.compute()
    Fe = md.state()["f"]
    err = relerr(Fe, Ft)
    print(f"relative difference {err:.3e}; RMS {np.sqrt(np.mean((Fe - Ft) ** 2)):.3e} of RMS "
          f"{np.sqrt(np.mean(Ft ** 2)):.3e}", flush=True)
    assert err < 1e-10
