"""Learned surrogate of plain SNAP: a per-atom MLP over cheap twojmax=4 bispectrum descriptors,
trained on energies and forces of plain twojmax=8 SNAP.

Everything that touches the data runs on the GPU through libmdpsnap:
  * labels: plain FP64 twojmax=8 SNAP MD (mdp_md_*) sampled at 300-3000 K;
  * descriptors: B and dB/dr at twojmax=4 by the batched fitting engine (mdp_fit_*);
  * model: mdp_mlp_eval (BF16 MFMA forward + input gradient), mdp_mlp_train (parameter gradient of
    the energy + force loss, including the second-order term of the force loss), mdp_mlp_adam;
  * descriptor chain rule F_i = -dE/dx_i: mdp_sur_pair_s / mdp_sur_atom_f, and its adjoint mdp_sur_u.
Device buffers are DevArray (mdp_dev_*), allocated by the library's own HIP runtime.  There is no
upstream counterpart: the reference's
potentials are linear/quadratic SNAP fitted by least squares (src/Julia/Potential/linearfit.jl).
"""
import ctypes
import json

import numpy as np

from . import library
from . import lattice
from .devarray import DevArray

NPARAM = 5313
J_SUR = 4
NIN = 14           # bispectrum components at twojmax = 4
J_PLAIN = 8
RCUTFAC, RFAC0 = 4.67637, 0.99363
MASS = 180.95

_bound = False


def _lib():
    global _bound
    c = library().cdll
    if _bound:
        return c
    vp, i, l, d, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_float
    c.mdp_mlp_pack.argtypes = [vp, vp, vp]
    c.mdp_mlp_eval.argtypes = [vp, vp, vp, i, vp, i, vp, vp, vp]
    c.mdp_mlp_train.argtypes = [vp, vp, vp, i, vp, vp, vp, i, vp, i, vp, vp]
    c.mdp_mlp_adam.argtypes = [vp, vp, vp, vp, f, f, f, f, i, f, vp]
    c.mdp_sur_pair_s.argtypes = [vp, vp, vp, i, i, l, vp, vp]
    c.mdp_sur_atom_f.argtypes = [vp, vp, vp, vp, l, i, vp, vp, vp, vp]
    c.mdp_sur_u.argtypes = [vp, vp, vp, vp, i, i, l, d, vp, vp]
    c.mdp_sur_energy.argtypes = [vp, vp, vp, i, d, vp, vp, vp]
    c.mdp_sumsq.argtypes = [vp, l, vp, i, vp]
    _bound = True
    return c


def _p(t):
    return t.ptr if t is not None else None


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {library().cdll.mdp_snap_last_error().decode()}")


def _stream():
    return None   # the library stream


def plain_coefficients(seed=20241202 + 2):
    """The plain twojmax=8 linear coefficients of the headline MD workload."""
    return lattice.snap_coefficients(55, beta0=-11.7, sigma=0.05, seed=seed)


class Potential:
    """A single-element plain twojmax=8 SNAP potential and the BCC lattice the surrogate samples around:
    coeff = [beta0, beta_1..beta_55], lattice constant a0 (A), rcutfac, rfac0, mass (amu)."""

    def __init__(self, name, coeff, a0, rcutfac, rfac0, mass, zbl=None):
        self.name, self.coeff, self.a0 = name, np.ascontiguousarray(coeff, np.float64), float(a0)
        self.rcutfac, self.rfac0, self.mass = float(rcutfac), float(rfac0), float(mass)
        self.zbl = zbl   # (cut_inner, cut_outer, z) of the ZBL reference pair term that goes with SNAP, or None


def seeded_tantalum():
    """The headline workload's seeded coefficients on BCC Ta (a = 3.316 A)."""
    return Potential("seeded coefficients, BCC Ta", plain_coefficients(), 3.316, RCUTFAC, RFAC0, MASS)


def load_coefficients(path):
    """A SNAP coefficient file as the upstream examples ship it: whitespace-separated numbers."""
    with open(path) as fh:
        return np.array([float(v) for v in fh.read().split()])


def tungsten(path, with_zbl=True):
    """The upstream tungsten potential (simulating/snap/snapW.m:34-60: BCC a = 3.1803 A, rcutfac 4.8123,
    rfac0 0.99363, twojmax 8, mass 183.84, and the ZBL reference term with inner/outer cutoffs 4 and 4.8 A,
    z = 74) with the coefficients of snapW.coeff at ``path``."""
    c = load_coefficients(path)
    if len(c) != 56:
        raise ValueError(f"{path}: expected 56 coefficients, found {len(c)}")
    return Potential("snapW.coeff + ZBL, BCC W" if with_zbl else "snapW.coeff, BCC W", c, 3.1803, 4.8123, 0.99363,
                     183.84, (4.0, 4.8, 74.0) if with_zbl else None)


class Model:
    """Parameters (FP32, the kernels' flat layout) and the descriptor/energy affine map."""

    def __init__(self, params, aff, nin=NIN):
        self.params = np.ascontiguousarray(params, np.float32)
        self.aff = np.ascontiguousarray(aff, np.float64)
        self.nin = nin

    @staticmethod
    def init(mu, sigma, e_scale, e_shift, seed=0, nin=NIN):
        rng = np.random.default_rng(seed)
        p = np.zeros(NPARAM, np.float32)
        W1 = np.zeros((16, 64))
        W1[:nin] = rng.normal(0, 1.0 / np.sqrt(nin), (nin, 64))
        p[0:1024] = W1.ravel()
        p[1088:5184] = rng.normal(0, 1.0 / 8.0, 64 * 64)
        p[5248:5312] = rng.normal(0, 1.0 / 8.0, 64)
        aff = np.zeros(34)
        aff[:nin] = mu
        aff[16:16 + nin] = 1.0 / np.maximum(sigma, 1e-12)
        aff[32], aff[33] = e_scale, e_shift
        return Model(p, aff, nin)

    def save(self, path):
        np.savez(path, params=self.params, aff=self.aff, nin=self.nin)

    @staticmethod
    def load(path):
        z = np.load(path, allow_pickle=False)
        return Model(z["params"], z["aff"], int(z["nin"]))


def make_training_set(nconfig=4000, ncell=5, tmin=300.0, tmax=3000.0, every=10, equil=10, per_run=10, seed=777,
                      fmax=200.0, rmin=2.0, log=None, pot=None, kinds=("bcc", "liquid"), vacancies=0, sigma=0.05):
    """Configurations of 250 atoms (BCC 5^3 cells at the potential's lattice constant -- the seeded Ta
    workload by default, a = 3.316 A -- and random packings at the BCC density).  With a ZBL reference
    term the dynamics include it, and the labels are the SNAP part alone (the surrogate stands in for SNAP;
    the engine adds the exact ZBL term to either)
    sampled from short plain twojmax=8 SNAP NVE runs started at temperatures spread over [tmin, tmax];
    labels are the plain SNAP total energy and forces.  Short runs keep the samples near their
    starting structures: with untrained (random) coefficients long hot runs drift into collapsed
    states.  Samples with a pair closer than ``rmin`` A or a force above ``fmax`` eV/A are dropped and
    counted.  ``kinds`` alternates the starting structures; BCC cells start with Gaussian displacements (per component) of
    RMS ``sigma`` A, and ``vacancies`` atoms are taken out of each at random.
    Returns (list of dict(x, box, E, F, kind, T), number dropped)."""
    import os
    from scipy.spatial import cKDTree
    from .md import SnapMD
    pot = pot or seeded_tantalum()
    coeff, a0, mass = pot.coeff, pot.a0, pot.mass
    out = []
    dropped = 0
    old = os.environ.get("MDP_AUTOTUNE")
    os.environ["MDP_AUTOTUNE"] = "0"     # small systems: the default launch variants
    try:
        nruns_est = max(2, (nconfig + per_run - 1) // per_run)
        run = 0
        while len(out) < nconfig:
            nk = len(kinds)
            kind = kinds[run % nk]
            T = tmin + (tmax - tmin) * ((run // nk) % max(1, nruns_est // nk)) / max(1, nruns_est // nk - 1)
            if kind == "bcc":
                x, box = lattice.crystal("bcc", ncell, a0, sigma, seed + run)
                if vacancies:
                    gone = np.random.default_rng(seed + 200000 + run).choice(len(x), vacancies, replace=False)
                    x = np.ascontiguousarray(np.delete(x, gone, axis=0))
            else:
                x, box = lattice.random_packing(2 * ncell ** 3, 2 / a0 ** 3, 2.2, seed=seed + run)
            v = lattice.maxwell_boltzmann(len(x), mass, T, seed=seed + 100000 + run)
            md = SnapMD(x, box, np.ones(len(x), np.int32), J_PLAIN, coeff, [0.5], [1.0], [mass], rcutfac=pot.rcutfac,
                        rfac0=pot.rfac0, v=v, dt=0.0005, skin=0.3, rebuild_every=10)
            lab = None
            if pot.zbl:
                md.set_zbl(*pot.zbl)
                lab = SnapMD(x, box, np.ones(len(x), np.int32), J_PLAIN, coeff, [0.5], [1.0], [mass],
                             rcutfac=pot.rcutfac, rfac0=pot.rfac0, dt=0.0005, skin=0.3, rebuild_every=10)
            md.compute()
            md.run(equil)
            for k in range(per_run):
                if len(out) == nconfig:
                    break
                md.run(every)
                st = md.state()
                xs = np.mod(st["x"], box)
                if lab is not None:   # the SNAP part alone on the same configuration
                    lab.set_state(xs)
                    lab.compute()
                    st = lab.state()
                d, _ = cKDTree(xs, boxsize=box).query(xs, k=2)
                if d[:, 1].min() < rmin or np.abs(st["f"]).max() > fmax:
                    dropped += 1
                    continue
                out.append(dict(x=xs, box=box.copy(), E=float(st["eatom"].sum()), F=st["f"].copy(), kind=kind, T=T))
            md.close()
            if lab is not None:
                lab.close()
            run += 1
            if log and run % 50 == 0:
                log(f"  {run} runs -> {len(out)} configurations ({dropped} dropped)")
    finally:
        if old is None:
            os.environ.pop("MDP_AUTOTUNE", None)
        else:
            os.environ["MDP_AUTOTUNE"] = old
    return out, dropped


class DeviceSet:
    """Descriptors, pair chain-rule arrays and labels of a list of configurations, on the GPU."""

    def __init__(self, configs, chunk=500, log=None, rcutfac=RCUTFAC, rfac0=RFAC0):
        from .fit import SnapFit
        C = len(configs)
        nat = np.array([len(c["x"]) for c in configs])
        a_off = np.concatenate([[0], np.cumsum(nat)])
        N = int(a_off[-1])
        X = np.zeros((NIN, N))
        blocks, pis, pjs = [], [], []
        p_tot = 0
        for c0 in range(0, C, chunk):
            sub = configs[c0:c0 + chunk]
            fit = SnapFit([(c["x"], c["box"]) for c in sub], J_SUR, rcutfac=rcutfac, rfac0=rfac0)
            fit.compute()
            b, db = fit.raw()
            ai, aj, _ = fit.pairs()
            la, lp = fit.layout()
            for k, c in enumerate(sub):
                g0 = int(a_off[c0 + k])
                n = int(la[k + 1] - la[k])
                P = int(lp[k + 1] - lp[k])
                X[:, g0:g0 + n] = b[int(la[k]) * NIN:int(la[k + 1]) * NIN].reshape(NIN, n)
                blocks.append(db[int(lp[k]) * 3 * NIN:int(lp[k + 1]) * 3 * NIN].reshape(NIN, 3, P))
                pis.append(ai[int(lp[k]):int(lp[k + 1])] + g0)
                pjs.append(aj[int(lp[k]):int(lp[k + 1])] + g0)
                p_tot += P
            fit.close()
            if log:
                log(f"  descriptors of configurations {c0}..{c0 + len(sub) - 1}")
        pi = np.concatenate(pis).astype(np.int32)
        pj = np.concatenate(pjs).astype(np.int32)
        dbdr = np.concatenate(blocks, axis=2)          # (K, 3, P), pairs sorted by centre atom
        own_off = np.concatenate([[0], np.cumsum(np.bincount(pi, minlength=N))]).astype(np.int32)
        in_idx = np.argsort(pj, kind="stable").astype(np.int32)
        in_off = np.concatenate([[0], np.cumsum(np.bincount(pj, minlength=N))]).astype(np.int32)
        self.C, self.N, self.P = C, N, int(p_tot)
        self.nat = nat
        self.X = DevArray.from_numpy(X)
        self.dbdr = DevArray.from_numpy(dbdr)
        del dbdr, blocks
        self.pi = DevArray.from_numpy(pi)
        self.pj = DevArray.from_numpy(pj)
        self.own_off = DevArray.from_numpy(own_off)
        self.in_off = DevArray.from_numpy(in_off)
        self.in_idx = DevArray.from_numpy(in_idx)
        self.cfg_off = DevArray.from_numpy(a_off.astype(np.int32))
        self.Eref_h = np.array([c["E"] for c in configs], dtype=np.float64)
        self.Fref_h = np.ascontiguousarray(np.concatenate([c["F"] for c in configs]), dtype=np.float64)
        self.Eref = DevArray.from_numpy(self.Eref_h)
        self.Fref = DevArray.from_numpy(self.Fref_h)
        self.Xh = X   # host copy for statistics


class Evaluator:
    """Energies and forces of a model on a DeviceSet, and the training step."""

    def __init__(self, ds, model, nwaves=1024):
        self.c = _lib()
        self.ds = ds
        self.model = model
        self.params = DevArray.from_numpy(model.params, np.float32)
        self.aff = DevArray.from_numpy(model.aff, np.float64)
        self.packed = DevArray(10240, np.int16)
        self.e = DevArray.zeros(ds.N)
        self.g = DevArray.zeros(NIN * ds.N)
        self.U = DevArray.zeros(NIN * ds.N)
        self.alpha = DevArray.zeros(ds.N)
        self.s = DevArray.zeros(3 * max(ds.P, 1))
        self.F = DevArray.zeros(3 * ds.N)
        self.R = DevArray.zeros(3 * ds.N)
        self.Ec = DevArray.zeros(ds.C)
        self.nwaves = nwaves
        self.part = DevArray.zeros(nwaves * NPARAM, np.float32)
        self.grad = DevArray.zeros(NPARAM, np.float32)
        self.m = DevArray.zeros(NPARAM, np.float32)
        self.v = DevArray.zeros(NPARAM, np.float32)
        self.ssq = DevArray.zeros(256)
        self.t = 0

    def forward(self, wE=1.0, need_alpha=False):
        c, ds, st = self.c, self.ds, _stream()
        _chk(c.mdp_mlp_pack(_p(self.params), _p(self.packed), st), "mdp_mlp_pack")
        _chk(c.mdp_mlp_eval(_p(self.packed), _p(self.params), _p(self.aff), self.model.nin, _p(ds.X), ds.N,
                            _p(self.e), _p(self.g), st), "mdp_mlp_eval")
        _chk(c.mdp_sur_energy(_p(self.e), _p(ds.cfg_off), _p(ds.Eref), ds.C, wE / ds.C, _p(self.Ec),
                              _p(self.alpha) if need_alpha else None, st), "mdp_sur_energy")
        _chk(c.mdp_sur_pair_s(_p(self.g), _p(ds.dbdr), _p(ds.pi), NIN, ds.N, ds.P, _p(self.s), st), "mdp_sur_pair_s")
        _chk(c.mdp_sur_atom_f(_p(self.s), _p(ds.own_off), _p(ds.in_off), _p(ds.in_idx), ds.P, ds.N, _p(ds.Fref),
                              _p(self.F), _p(self.R), st), "mdp_sur_atom_f")

    def metrics(self):
        """(force RMSE eV/A, energy RMSE eV/atom) of the last forward."""
        c, ds, st = self.c, self.ds, _stream()
        _chk(c.mdp_sumsq(_p(self.R), 3 * ds.N, _p(self.ssq), 256, st), "mdp_sumsq")
        frmse = float(np.sqrt(self.ssq.numpy().sum() / (3 * ds.N)))
        de = (self.Ec.numpy() - ds.Eref_h) / ds.nat
        return frmse, float(np.sqrt(np.mean(de * de)))

    def step(self, lr, wE=1.0, wF=1.0, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        """One Adam step on L = wE mean_c ((E_c - E_c^ref)/N_c)^2 + wF mean_i,c (F - F^ref)^2."""
        c, ds, st = self.c, self.ds, _stream()
        self.forward(wE, need_alpha=True)
        _chk(c.mdp_sur_u(_p(self.R), _p(ds.dbdr), _p(ds.own_off), _p(ds.pj), NIN, ds.N, ds.P, 2.0 * wF / (3 * ds.N),
                         _p(self.U), st), "mdp_sur_u")
        _chk(c.mdp_mlp_train(_p(self.packed), _p(self.params), _p(self.aff), self.model.nin, _p(ds.X), _p(self.U),
                             _p(self.alpha), ds.N, _p(self.part), self.nwaves, _p(self.grad), st), "mdp_mlp_train")
        self.t += 1
        _chk(c.mdp_mlp_adam(_p(self.params), _p(self.grad), _p(self.m), _p(self.v), lr, b1, b2, eps, self.t, wd, st),
             "mdp_mlp_adam")

    def current_model(self):
        return Model(self.params.numpy(), self.model.aff, self.model.nin)


def affine_from(ds):
    """Descriptor standardization and per-atom energy scale of a training set."""
    mu = ds.Xh.mean(axis=1)
    sg = ds.Xh.std(axis=1)
    eat = ds.Eref_h / ds.nat
    return mu, sg, max(float(eat.std()), 0.1), float(eat.mean())


def train(ds, model, steps=2000, lr=2e-3, lr_end=1e-4, wE=1.0, wF=1.0, log=None, every=100):
    ev = Evaluator(ds, model)
    hist = []
    for t in range(steps):
        frac = t / max(1, steps - 1)
        ev.step(lr * (lr_end / lr) ** frac, wE, wF)
        if log and (t % every == 0 or t == steps - 1):
            ev.forward(wE)
            fr, er = ev.metrics()
            hist.append((t, fr, er))
            log(f"  step {t}: force RMSE {fr:.4f} eV/A, energy RMSE {er:.4f} eV/atom")
    return ev.current_model(), hist


def evaluate(ds, model):
    ev = Evaluator(ds, model)
    ev.forward()
    return ev.metrics()


def linear_baseline(train_ds, test_ds, wE=1.0, wF=1.0):
    """The method without the network: linear twojmax=4 SNAP (E_atom = b0 + beta . B) fitted by least
    squares on the same energies and forces; returns (force RMSE, energy RMSE per atom) on test_ds."""
    def columns(ds):
        c, st = _lib(), _stream()
        g = DevArray(NIN * ds.N)
        s = DevArray(3 * max(ds.P, 1))
        F = DevArray(3 * ds.N)
        cols = []
        for k in range(NIN):
            gh = np.zeros(NIN * ds.N)
            gh[k * ds.N:(k + 1) * ds.N] = 1.0
            g.upload(gh)
            _chk(c.mdp_sur_pair_s(_p(g), _p(ds.dbdr), _p(ds.pi), NIN, ds.N, ds.P, _p(s), st), "mdp_sur_pair_s")
            _chk(c.mdp_sur_atom_f(_p(s), _p(ds.own_off), _p(ds.in_off), _p(ds.in_idx), ds.P, ds.N, None, _p(F), None,
                                  st), "mdp_sur_atom_f")
            cols.append(F.numpy())
        Ff = np.stack(cols, axis=1)                      # (3N, K): force per unit beta_k
        a_off = np.concatenate([[0], np.cumsum(ds.nat)])
        Bs = np.add.reduceat(ds.Xh, a_off[:-1], axis=1).T  # (C, K): sum of B over each configuration
        return Ff, Bs

    Ff, Bs = columns(train_ds)
    nat = train_ds.nat.astype(np.float64)
    Ae = np.concatenate([nat[:, None], Bs], axis=1) / nat[:, None]
    be = train_ds.Eref_h / nat
    Af = np.concatenate([np.zeros((Ff.shape[0], 1)), Ff], axis=1)
    bf = train_ds.Fref_h.ravel()
    A = np.concatenate([np.sqrt(wE / len(be)) * Ae, np.sqrt(wF / len(bf)) * Af])
    b = np.concatenate([np.sqrt(wE / len(be)) * be, np.sqrt(wF / len(bf)) * bf])
    coef, *_ = np.linalg.lstsq(A, b, rcond=None)
    Ff2, Bs2 = columns(test_ds)
    fpred = Ff2 @ coef[1:]
    epred = (test_ds.nat * coef[0] + Bs2 @ coef[1:]) / test_ds.nat
    fr = float(np.sqrt(np.mean((fpred - test_ds.Fref_h.ravel()) ** 2)))
    er = float(np.sqrt(np.mean((epred - test_ds.Eref_h / test_ds.nat) ** 2)))
    return fr, er, coef


def summary_json(d):
    return json.dumps({k: (float(v) if isinstance(v, (np.floating,)) else v) for k, v in d.items()})
