#!/usr/bin/env python3
"""Plain FP64 SNAP molecular dynamics on one MI355X: atom-steps/s at twojmax=8, 128,000 atoms.

Workload (one "step" = one velocity-Verlet NVE step of the whole system): BCC tantalum,
40 x 40 x 40 cubic cells (128,000 atoms), a = 3.316 A, seeded Gaussian displacements sigma = 0.1 A,
twojmax = 8 (55 bispectrum components), rcutfac = 4.67637, rfac0 = 0.99363, rmin0 = 0, switching and
bzero on, linear coefficients beta_k ~ N(0, 0.05) with beta0 = -11.7 eV, 300 K Maxwell-Boltzmann
velocities, mass 180.95 amu, dt = 0.5 fs, neighbour skin 0.3 A rebuilt every 10 steps.  Every step
integrates, amortises the neighbour rebuild and evaluates SNAP energy, forces and virials in FP64.

With --workload fit it measures the batched fitting-descriptor path instead: per-atom B and per-pair
dB/dr (twojmax = 8, FP64) of 512 independent periodic configurations (atom counts log-uniform in
[16, 2000]; BCC a ~ U[3.2, 3.45], FCC a ~ U[4.0, 4.3] and random packings at 0.0554 atoms/A^3 with 2.2 A
minimum separation, all displaced by sigma = 0.1 A), one step = one replay of the captured batch graph
(pair build, Utot, Z + B, dB/dr), reported as atoms/s.

With --workload surrogate it measures the learned surrogate: a per-atom 14 -> 64 -> 64 -> 1 SiLU MLP
(BF16 matrix-core kernels) over twojmax=4 descriptors, trained on the GPU on 4000 250-atom BCC/liquid
configurations sampled from plain twojmax=8 SNAP MD at 300-3000 K (energies and forces, coefficients of
the headline workload), then used for NVE MD of 80^3 BCC cells with 0.5% of the sites vacant
(1,018,880 atoms, 600 K, dt = 1 fs) with a plain FP64 twojmax=8 audit every 50 steps that falls back
to plain forces when the RMS force difference exceeds 0.05 eV/A.  Reported: surrogate atom-steps/s with
audits (value), without audits, the plain-GPU rate on the same system, and the force errors.

Prints one JSON line on stdout; diagnostics go to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s, SNAP 2J=8 @128k atoms; surrogate vs plain at ≤0.05 eV/Å RMSE"
FP64_PEAK_TFLOPS = 78.6      # MI355X published FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0
J, NCELL, A0, SIGMA = 8, 40, 3.316, 0.1
RCUTFAC, RFAC0, MASS, DT, SKIN, EVERY, TEMP = 4.67637, 0.99363, 180.95, 0.0005, 0.3, 10, 300.0
SEED = 20241202


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(ncell=NCELL):
    from mdp_amd import lattice
    from mdp_amd.flops import snap_flops  # noqa: F401
    x, box = lattice.crystal("bcc", ncell, A0, SIGMA, SEED)
    v = lattice.maxwell_boltzmann(len(x), MASS, TEMP, seed=SEED + 1)
    ncoeff = 55
    coeff = lattice.snap_coefficients(ncoeff, beta0=-11.7, sigma=0.05, seed=SEED + 2)
    return x, box, v, coeff


def periodic_pairs(x, box, rc):
    """Minimum-image pair arrays (centre-sorted) of an orthogonal periodic box, box > 2 rc."""
    from scipy.spatial import cKDTree
    xw = np.mod(x, box)
    tree = cKDTree(xw, boxsize=box)
    ij = tree.query_pairs(rc, output_type="ndarray")
    ij = np.concatenate([ij, ij[:, ::-1]])
    d = xw[ij[:, 1]] - xw[ij[:, 0]]
    d -= box * np.round(d / box)
    d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    keep = (d2 < rc * rc) & (d2 > 1e-20)
    ij, d = ij[keep], d[keep]
    order = np.lexsort((ij[:, 1], ij[:, 0]))
    ij, d = ij[order], d[order]
    n = len(x)
    return dict(rij=np.ascontiguousarray(d.ravel()), ai=ij[:, 0].astype(np.int32), aj=ij[:, 1].astype(np.int32),
                ti=np.ones(len(ij), np.int32), tj=np.ones(len(ij), np.int32), atomtype=np.ones(n, np.int32),
                ilist=np.arange(n, dtype=np.int32), n=n, ijnum=len(ij))


def cpu_library():
    """(kind, path) of the CPU SNAP library timed as the baseline: the upstream C library compiled from
    the upstream source (oracle/_ref), or this project's CPU restatement where that is absent."""
    ref = os.path.join(ROOT, "oracle", "_ref", "libcpusnap_ref.so")
    port = os.path.join(ROOT, "oracle", "liboracle.so")
    return ("reference", ref) if os.path.exists(ref) else ("port", port)


def host_cores():
    """Host threads for the all-cores baseline: this process's CPU affinity, at most 16 (one GPU's share
    of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def force_chain(lib, s, rij, ai, aj, lo, hi, n):
    """Ui -> AddWself2Ui -> Ei -> Yi -> Fi (snapcpp.jl:386-423 order) of the CPU library for the centre
    atoms [lo, hi) of centre-sorted pair arrays (ai ascending): U, E and Y are held for those atoms only,
    forces and virials land in the full n-atom arrays (cpuSnapComputeFi's aii/ai/aj and inum/anum
    arguments, cpuSnap.cpp:990-993).  Returns (eatom of the slice, fatom (n,3), vatom (n,6))."""
    p0, p1 = np.searchsorted(ai, [lo, hi])
    rij = np.ascontiguousarray(rij[3 * p0:3 * p1])
    aig = np.ascontiguousarray(ai[p0:p1])
    ajg = np.ascontiguousarray(aj[p0:p1])
    ail = np.ascontiguousarray(aig - lo, np.int32)
    m, P = hi - lo, p1 - p0
    ones_p, ones_m = np.ones(P, np.int32), np.ones(m, np.int32)
    iu = s.idxu_max
    Ur, Ui = np.zeros(m * iu), np.zeros(m * iu)
    lib.cpuSnapComputeUi(Ur, Ui, s.rootpqarray, rij, s.wjelem, s.radelem, s.rmin0, s.rfac0, s.rcutfac, s.map,
                         ail, ones_p, ones_p, s.twojmax, iu, m, P, s.switchflag, s.chemflag)
    lib.cpuAddWself2Ui(Ur, Ui, s.wself, s.idxu_block, ones_m, s.map, s.wselfallflag, s.chemflag, iu,
                       s.nelements, s.twojmax, m)
    e = np.zeros(m)
    lib.cpuSnapComputeEi(e, Ur, Ui, s.cglist, s.bzero, s.coeffelem, np.arange(m, dtype=np.int32), s.map, ones_m,
                         s.idxb, s.idxcg_block, s.idxu_block, s.twojmax, s.idxb_max, iu, s.nelements,
                         s.ncoeffall, s.bnormflag, s.bzeroflag, s.wselfallflag, s.chemflag, m)
    Yr, Yi = np.zeros(m * iu), np.zeros(m * iu)
    lib.cpuSnapComputeYi(Yr, Yi, Ur, Ui, s.cglist, s.coeffelem, s.map, ones_m, s.idxz, s.idxb_block,
                         s.idxu_block, s.idxcg_block, s.twojmax, s.idxb_max, iu, s.idxz_max, s.nelements,
                         s.ncoeffall, s.bnormflag, s.chemflag, m)
    f, v = np.zeros(3 * n), np.zeros(6 * n)
    lib.cpuSnapComputeFi(f, v, Yr, Yi, s.rootpqarray, rij, s.wjelem, s.radelem, s.rmin0, s.rfac0, s.rcutfac,
                         s.map, ail, aig, ajg, ones_p, ones_p, s.twojmax, iu, m, n, P, s.switchflag, s.chemflag)
    return e, f.reshape(n, 3), v.reshape(n, 6)


def _sna(lib, twojmax, coeff):
    from mdp_amd import snapcpp
    return snapcpp.initsna(lib, [1, twojmax, RCUTFAC, RFAC0, 0.0, 1, 1, 0, 0, 0, 0], [0.5], [1.0], coeff)


def cpu_worker(task_path, rank, nrank):
    """One process of the all-cores CPU baseline (started by all_cores): loads the task, prepares its
    share, prints "ready", waits for "go" on stdin, times its share and prints {"elapsed", "units"}."""
    from mdp_amd import _abi, snapcpp
    rank, nrank = int(rank), int(nrank)
    d = np.load(task_path)
    lib = _abi.bind(str(d["lib"]))
    task = str(d["task"])
    if task == "force":
        # a contiguous slice of centre atoms of one force call
        n = int(d["n"])
        s = _sna(lib, int(d["twojmax"]), d["coeff"])
        lo, hi = rank * n // nrank, (rank + 1) * n // nrank
        rij, ai, aj = d["rij"], d["ai"], d["aj"]
        work = lambda: force_chain(lib, s, rij, ai, aj, lo, hi, n)   # noqa: E731
        units = hi - lo
    else:
        # every nrank-th configuration of the fitting batch: the descriptor chain of snapcpp.jl:599-664
        s = snapcpp.initsna(lib, [1, int(d["twojmax"]), RCUTFAC, RFAC0, 0.0, 1, 1, 0, 0, 0, 0], [0.5], [1.0])
        off, xs, boxes = d["off"], d["x"], d["box"]
        mine = range(rank, len(off) - 1, nrank)
        pairs = [snapcpp.pairs_from_box(xs[off[c]:off[c + 1]], np.ones(off[c + 1] - off[c], np.int32),
                                        [boxes[c, 0], 0, 0], [0, boxes[c, 1], 0], [0, 0, boxes[c, 2]], [1, 1, 1], s)
                 for c in mine]

        def work():
            for p in pairs:
                snapcpp.snapdescriptors_pairs(lib, s, p)
        units = int(sum(off[c + 1] - off[c] for c in mine))
    print("ready", flush=True)
    if sys.stdin.readline().strip() != "go":
        sys.exit(3)
    t0 = time.perf_counter()
    work()
    print(json.dumps({"elapsed": time.perf_counter() - t0, "units": units}), flush=True)


def fit_cpu_baseline(cfg, every, cores=None):
    """The CPU library's descriptor chain (snapcpp.jl:599-664: Uij, Utot, Zi, Bi, dB/dr and the three
    tallies -- the per-configuration body of the loop in Regression.cpp:68-95, neighbour lists built
    beforehand) over every `every`-th configuration of the batch, one process per host core taking
    configurations in turn."""
    kind, path = cpu_library()
    sample = cfg[::every]
    off = np.concatenate([[0], np.cumsum([len(x) for x, _ in sample])]).astype(np.int64)
    nw = cores or host_cores()
    rate, slowest = all_cores(dict(task="fit", lib=path, twojmax=J, off=off,
                                   x=np.concatenate([x for x, _ in sample]), box=np.array([b for _, b in sample])), nw)
    return dict(value=rate, unit="atoms/s", cores=nw, kind=kind,
                sample=f"{len(sample)} configurations (every {every}th of the batch, {int(off[-1])} atoms): the "
                       f"descriptor chain per configuration, {nw} processes, slowest {slowest:.2f} s",
                seconds_per_configuration=slowest * nw / len(sample))


def all_cores(arrays, nwork):
    """Runs cpu_worker in nwork child processes on the same task, starts their timed parts together and
    returns (units done per second over the slowest process, slowest process's seconds)."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "task.npz")
        np.savez(path, **arrays)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", path, str(r),
                                   str(nwork)], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True, env=env)
                 for r in range(nwork)]
        try:
            for p in procs:
                if p.stdout.readline().strip() != "ready":
                    raise RuntimeError("CPU baseline worker failed during set-up")
            for p in procs:
                p.stdin.write("go\n")
                p.stdin.flush()
            res = [json.loads(p.stdout.readline()) for p in procs]
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError("CPU baseline worker failed")
        finally:
            for p in procs:
                if p.poll() is None:
                    p.kill()
                    p.wait()
    slowest = max(r["elapsed"] for r in res)
    return sum(r["units"] for r in res) / slowest, slowest


def cpu_baseline(ncell, full_x=None, full_box=None, cores=None):
    """The CPU SNAP library (cpu_library) timed for the force call (Ui -> AddWself2Ui -> Ei -> Yi -> Fi,
    snapcpp.jl:386-423 order) of the workload's lattice: on one core for an ncell^3-cell cut, and on all
    host cores (one process per core over disjoint slices of centre atoms) for the full system
    (full_x, full_box) when given."""
    from mdp_amd import _abi
    kind, path = cpu_library()
    lib = _abi.bind(path)
    x, box, _, coeff = workload(ncell)
    s = _sna(lib, J, coeff)
    p = periodic_pairs(x, box, RCUTFAC)
    t0 = time.perf_counter()
    force_chain(lib, s, p["rij"], p["ai"], p["aj"], 0, len(x), len(x))
    dt = time.perf_counter() - t0
    one = len(x) / dt
    out = dict(value=one, unit="atom-steps/s", cores=1, kind=kind,
               sample=f"first E/F/V force call of {len(x)} atoms ({ncell}^3 BCC cells, same a/sigma/twojmax/"
                      f"coefficients, {p['ijnum']} pairs), single thread, {dt:.2f} s",
               single_core_value=one,
               build=("upstream cpuSnap.cpp" if kind == "reference" else "oracle/snap_oracle.c") +
                     " built by oracle/Makefile at -O2 -fno-fast-math -ffp-contract=off")
    if full_x is not None:
        nw = cores or host_cores()
        q = periodic_pairs(full_x, full_box, RCUTFAC)
        rate, slowest = all_cores(dict(task="force", lib=path, n=len(full_x), twojmax=J, coeff=coeff,
                                       rij=q["rij"], ai=q["ai"], aj=q["aj"]), nw)
        out.update(value=rate, cores=nw, all_cores_value=rate,
                   sample=f"first E/F/V force call of the full {len(full_x)}-atom system ({q['ijnum']} pairs) "
                          f"split over {nw} processes by centre atom, slowest {slowest:.2f} s; single thread on "
                          f"{len(x)} atoms ({ncell}^3 cells): {one:.1f} atom-steps/s in {dt:.2f} s")
    return out


def bench_fit(a):
    """Batched B + dB/dr of the 512-configuration set, timed over graph replays."""
    import mdp_amd
    from mdp_amd import lattice
    from mdp_amd.fit import SnapFit
    from mdp_amd.flops import snap_flops

    mdp_amd.library()
    t0 = time.perf_counter()
    batch = lattice.fitting_batch()
    cfg = [(x, box) for x, box, _ in batch]
    natoms = sum(len(c[0]) for c in cfg)
    log(f"512 configurations, {natoms} atoms, generated in {time.perf_counter() - t0:.1f} s")
    fit = SnapFit(cfg, J, rcutfac=RCUTFAC, rfac0=RFAC0)
    t0 = time.perf_counter()
    fit.compute()
    log(f"first compute (pair sizing, autotuning, graph capture): {1e3 * (time.perf_counter() - t0):.1f} ms")
    log(f"autotuner: {json.dumps(fit.tuning())}")
    _, p_off = fit.layout()
    npairs = int(p_off[-1])
    prof = fit.profile()
    log(f"stages (ms, one uncaptured run): {json.dumps(prof)}")
    fit.replay(a.warmup)
    fit.sync()
    t0 = time.perf_counter()
    fit.replay(a.steps)
    fit.sync()
    elapsed = time.perf_counter() - t0
    ms = 1e3 * elapsed / a.steps
    value = natoms / (elapsed / a.steps)
    fl = snap_flops(J)
    dom = max(("zb", "dbdr"), key=lambda k: prof[k])
    work = fl["dbdr"] * npairs if dom == "dbdr" else fl["zb"] * natoms
    achieved = work / (prof[dom] * 1e-3) / 1e12
    out_bytes = 8 * 3 * 55 * npairs + 8 * 55 * natoms
    log(f"{a.steps} replays: {ms:.3f} ms per batch, {value:.4e} atoms/s, {npairs} pairs "
        f"({npairs / natoms:.2f}/atom); output {out_bytes / 1e9:.2f} GB per batch; {dom}: {achieved:.2f} TFLOP/s")
    # the linear-fit consumer of the same descriptors (not part of the timed replays): tallies and
    # normal equations of energy, force and stress rows on the GPU, then the host solve
    from mdp_amd.fit import lsqsolve
    rng = np.random.default_rng(0)
    e, f, s = rng.normal(size=fit.C), rng.normal(size=(natoms, 3)), rng.normal(size=(fit.C, 6))
    fit.normal_equations(e, f, s)
    t0 = time.perf_counter()
    L, h = fit.normal_equations(e, f, s)
    t_ne = 1e3 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    lsqsolve(L[0], L[1], L[2], h[0], h[1], h[2])
    t_sol = 1e3 * (time.perf_counter() - t0)
    log(f"linear fit on the batch: tallies + normal equations (3 groups, {3 * natoms + 7 * fit.C} rows) "
        f"{t_ne:.2f} ms, lsqsolve (4 systems of {L.shape[1]}) {t_sol:.2f} ms")
    out = {
        "metric": "atoms/s, batched SNAP B + dB/dr, 2J=8, 512 configurations", "value": value, "unit": "atoms/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded BCC/FCC/random-packing configurations)",
        "config": {"workload": "batched fitting descriptors: per-atom B and per-pair dB/dr of 512 periodic "
                               "configurations (16-2000 atoms), twojmax=8, one captured graph per batch",
                   "atoms": natoms, "pairs": npairs, "twojmax": J, "parallelism": "1 GPU",
                   "autotuned": fit.tuning().get("chosen", {})},
        "roofline": {"bound": "mfma", "kernel": "fit_" + dom, "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "pipe": ("FP64 matrix (v_mfma_f64_16x16x4f64)"
                              if dom == "dbdr" and str(fit.tuning().get("chosen", {}).get(
                                  "fit_dbdr_variant", os.environ.get("MDP_FIT_DBDR_VARIANT"))) == "6"
                              else "FP64 vector"),
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": work, "ms_per_launch": prof[dom]},
        "cpu_baseline": None,
    }
    fit.close()
    if not a.no_cpu_baseline:
        t0 = time.perf_counter()
        out["cpu_baseline"] = fit_cpu_baseline(cfg, 4, a.cpu_cores or None)
        log(f"cpu baseline: {out['cpu_baseline']['value']:.1f} atoms/s ({time.perf_counter() - t0:.1f} s)")
    print(json.dumps(out), flush=True)


def bench_surrogate(a):
    """Train the surrogate on plain-SNAP data, then run surrogate MD on 80^3 BCC cells with vacancies."""
    import mdp_amd
    from mdp_amd import lattice
    from mdp_amd import surrogate as S
    from mdp_amd.md import SnapMD

    mdp_amd.library()
    pot = S.tungsten(os.path.join(ROOT, "tests", "golden", "snapW.coeff")) if a.potential == "W" else S.seeded_tantalum()
    log(f"potential: {pot.name} (a0 {pot.a0} A, rcutfac {pot.rcutfac}, mass {pot.mass})")
    t0 = time.perf_counter()
    kinds = ("bcc", "liquid") if a.train_kinds == "both" else ("bcc",)
    log(f"training runs: {'/'.join(kinds)} at {a.train_tmin:.0f}-{a.train_tmax:.0f} K, "
        f"{a.train_vacancies} vacancies per BCC cell, {a.train_equil} steps before the first sample, "
        f"{a.train_every} between samples, BCC displacements {a.train_sigma} A")
    data, dropped = S.make_training_set(nconfig=a.train_configs, log=log, pot=pot, tmin=a.train_tmin,
                                        tmax=a.train_tmax, kinds=kinds, vacancies=a.train_vacancies,
                                        equil=a.train_equil, every=a.train_every, sigma=a.train_sigma)
    log(f"plain-SNAP training data: {len(data)} configurations in {time.perf_counter() - t0:.1f} s "
        f"({dropped} samples dropped: a pair under 2 A or a force above 200 eV/A)")
    ntest = max(1, len(data) // 10)
    rng = np.random.default_rng(SEED + 7)
    perm = rng.permutation(len(data))
    train = [data[i] for i in perm[ntest:]]
    test = [data[i] for i in perm[:ntest]]
    t0 = time.perf_counter()
    ds_tr = S.DeviceSet(train, log=log, rcutfac=pot.rcutfac, rfac0=pot.rfac0)
    ds_te = S.DeviceSet(test, rcutfac=pot.rcutfac, rfac0=pot.rfac0)
    log(f"twojmax=4 descriptors: {ds_tr.N + ds_te.N} atoms, {ds_tr.P + ds_te.P} pairs in {time.perf_counter() - t0:.1f} s")
    mu, sg, es, esh = S.affine_from(ds_tr)
    m0 = S.Model.init(mu, sg, es, esh, seed=1)
    t0 = time.perf_counter()
    model, hist = S.train(ds_tr, m0, steps=a.train_steps, lr=3e-3, lr_end=1e-4, log=log, every=250)
    t_train = time.perf_counter() - t0
    f_te, e_te = S.evaluate(ds_te, model)
    f_lin, e_lin, _ = S.linear_baseline(ds_tr, ds_te)
    log(f"training {a.train_steps} steps in {t_train:.1f} s; held-out force RMSE {f_te:.4f} eV/A "
        f"(linear twojmax=4 SNAP {f_lin:.4f}), energy RMSE {e_te:.4f} eV/atom (linear {e_lin:.4f})")

    # production system: 80^3 BCC cells, 0.5% vacancies, 600 K, dt = 1 fs
    x, box = lattice.crystal("bcc", a.sur_ncell, pot.a0, SIGMA, SEED + 10)
    keep = np.ones(len(x), bool)
    keep[np.random.default_rng(SEED + 11).choice(len(x), int(round(0.005 * len(x))), replace=False)] = False
    x = np.ascontiguousarray(x[keep])
    n = len(x)
    v = lattice.maxwell_boltzmann(n, pot.mass, 600.0, seed=SEED + 12)

    def engine():
        e = SnapMD(x, box, np.ones(n, np.int32), J, pot.coeff, [0.5], [1.0], [pot.mass], rcutfac=pot.rcutfac,
                   rfac0=pot.rfac0, v=v, dt=0.001, skin=SKIN, rebuild_every=EVERY)
        if pot.zbl:   # the exact reference pair term, with plain SNAP and with the surrogate alike
            e.set_zbl(*pot.zbl)
        return e

    def run_chunks(md, steps, chunk=50):
        """Steps in chunks; a run that fails (e.g. the neighbour build reports a collapsed system) stops
        there.  Returns (steps done, seconds, error or None)."""
        done, t0 = 0, time.perf_counter()
        try:
            while done < steps:
                k = min(chunk, steps - done)
                md.run(k)
                done += k
        except RuntimeError as e:
            return done, time.perf_counter() - t0, str(e)
        return done, time.perf_counter() - t0, None

    # each phase starts from the same state
    md = engine()
    md.compute()
    md.run(a.warmup)
    md.set_timing(True)
    t0 = time.perf_counter()
    md.run(a.plain_steps)
    t_plain = (time.perf_counter() - t0) / a.plain_steps
    temp_plain = md.thermo()["temp"]
    log(f"plain MD after {a.warmup + a.plain_steps} steps: {temp_plain:.0f} K")
    md.close()
    md = engine()
    md.set_surrogate(model, audit_every=0)
    md.compute()
    w_done, _, na_err = run_chunks(md, a.warmup)
    na_done, na_s = 0, 0.0
    if not na_err:
        na_done, na_s, na_err = run_chunks(md, a.noaudit_steps)
    t_noaudit = na_s / max(na_done, 1)
    md.close()
    if na_err:
        log(f"surrogate without audits stopped after {w_done + na_done} steps: {na_err}")
    md = engine()
    md.set_surrogate(model, audit_every=50, audit_tol=0.05)
    md.compute()
    md.set_timing(True)
    au_done, elapsed, au_err = run_chunks(md, a.steps)
    if au_err:
        log(f"surrogate with audits stopped after {au_done} steps: {au_err}")
    timers = md.timers()
    aud = md.audits()
    st = md.stats()
    temp_sur = md.thermo()["temp"]
    log(f"surrogate MD with audits after {au_done} steps: {temp_sur:.0f} K")
    md.close()
    steps_timed = max(au_done, 1)
    elapsed = elapsed * a.steps / steps_timed if au_done else elapsed
    value = n * a.steps / elapsed
    log(f"{n} atoms: plain {1e3 * t_plain:.2f} ms/step, surrogate without audits {1e3 * t_noaudit:.2f} ms/step, "
        f"with audits {1e3 * elapsed / a.steps:.2f} ms/step; audits {aud}")
    for k, (ms, cnt) in timers.items():
        if cnt:
            log(f"  {k:16s} {ms:9.2f} ms total, {cnt:6d} launches, {ms / max(cnt, 1):8.4f} ms avg")
    out = {
        "metric": "atom-steps/s, surrogate MD (twojmax=4 MLP) with plain twojmax=8 audits every 50 steps",
        "value": value, "unit": "atom-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 MLP / f64 descriptors, forces and integration",
        "data": "synthetic (plain-SNAP MD samples for training; seeded BCC lattice with vacancies for MD)",
        "config": {"workload": f"surrogate NVE MD, BCC {a.sur_ncell}^3 cells with 0.5% vacancies ({n} atoms), "
                               "600 K, dt=1 fs, audit every 50 steps", "potential": pot.name, "atoms": n,
                   "parallelism": "1 GPU",
                   "train_configs": len(train), "train_steps": a.train_steps, "samples_dropped": dropped,
                   "train_runs": {"kinds": list(kinds), "tmin": a.train_tmin, "tmax": a.train_tmax,
                                  "vacancies": a.train_vacancies, "equil": a.train_equil,
                                  "every": a.train_every, "sigma": a.train_sigma}},
        "surrogate": {"atom_steps_per_s_without_audits": n / t_noaudit if na_done else None,
                      "plain_gpu_atom_steps_per_s": n / t_plain,
                      "steps_without_audits": w_done + na_done, "error_without_audits": na_err,
                      "steps_with_audits": au_done, "error_with_audits": au_err,
                      "speedup_vs_plain_with_audits": (n * a.steps / elapsed) / (n / t_plain),
                      "speedup_vs_plain_without_audits": t_plain / t_noaudit if na_done else None,
                      "heldout_force_rmse": f_te, "heldout_energy_rmse_per_atom": e_te,
                      "linear_j4_force_rmse": f_lin, "linear_j4_energy_rmse_per_atom": e_lin,
                      "audit_force_rmse_mean": aud["mean_rmse"], "audit_force_rmse_max": aud["max_rmse"],
                      "md_temperature_plain_K": temp_plain, "md_temperature_surrogate_K": temp_sur,
                      "audits": aud["audits"], "fallbacks": aud["fallbacks"], "target_force_rmse": 0.05,
                      "pairs_per_atom": st["pairs"] / n},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def cpu_md(lib, s, x, box, v, steps, dt=DT, skin=SKIN, every=EVERY):
    """Velocity-Verlet NVE on the host with the CPU library's force call (force_chain) at every step and a
    minimum-image pair list of radius rcut + skin rebuilt every `every` steps (the upstream step loop,
    Integration.cpp:17-156, on one core).  Returns (final x, final v, seconds)."""
    from scipy.spatial import cKDTree
    from mdp_amd.lattice import FTM2V_METAL
    n = len(x)
    rc = s.rcutmax
    x, v = x.copy(), v.copy()
    dtf = 0.5 * dt * FTM2V_METAL / MASS
    ij = None

    def forces():
        d = x[ij[:, 1]] - x[ij[:, 0]]
        d -= box * np.round(d / box)
        keep = np.einsum("ij,ij->i", d, d) < rc * rc
        return force_chain(lib, s, np.ascontiguousarray(d[keep].ravel()), ij[keep, 0], ij[keep, 1], 0, n, n)[1]

    t0 = time.perf_counter()
    for step in range(steps + 1):
        if step % every == 0:
            xw = np.mod(x, box)
            q = cKDTree(xw, boxsize=box).query_pairs(rc + skin, output_type="ndarray")
            q = np.concatenate([q, q[:, ::-1]]).astype(np.int32)
            ij = q[np.lexsort((q[:, 1], q[:, 0]))]
        f = forces()
        if step:
            v += dtf * f
        if step < steps:
            v += dtf * f
            x += dt * v
    return x, v, time.perf_counter() - t0


def bench_config0(a):
    """The 2,000-atom plain-SNAP MD case (BCC Ta 10^3 cells, sigma = 0.05 A, twojmax = 6, 100 NVE steps)
    on the GPU engine, beside the same 100 steps driven by the CPU library on one host core."""
    import mdp_amd
    from mdp_amd import _abi, lattice
    from mdp_amd.md import SnapMD

    mdp_amd.library()
    J0, steps_cpu = 6, 100
    x, box = lattice.crystal("bcc", 10, A0, 0.05, 20241200)
    v = lattice.maxwell_boltzmann(len(x), MASS, TEMP, seed=20241201)
    coeff = lattice.snap_coefficients(30, beta0=-11.7, sigma=0.05, seed=20241202)
    n = len(x)

    def engine():
        return SnapMD(x, box, np.ones(n, np.int32), J0, coeff, [0.5], [1.0], [MASS], rcutfac=RCUTFAC,
                      rfac0=RFAC0, v=v, dt=DT, skin=SKIN, rebuild_every=EVERY)
    md = engine()
    md.compute()
    md.run(a.warmup)
    md.close()
    md = engine()
    t0 = time.perf_counter()
    md.compute()
    md.run(a.steps)
    elapsed = time.perf_counter() - t0
    st = md.state()
    md.close()
    log(f"GPU: {n} atoms, {a.steps} steps (incl. the first force call) in {1e3 * elapsed:.2f} ms")
    out = {
        "metric": "atom-steps/s, plain SNAP MD, 2J=6, 2,000 atoms", "value": n * a.steps / elapsed,
        "unit": "atom-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded BCC lattice, displacements, velocities and SNAP coefficients)",
        "config": {"workload": "plain FP64 SNAP NVE MD, BCC Ta 10^3 cells (2000 atoms), sigma 0.05 A, "
                               "twojmax=6, dt=0.5 fs, skin 0.3 A rebuilt every 10 steps", "atoms": n,
                   "twojmax": J0, "parallelism": "1 GPU"},
        "roofline": None, "cpu_baseline": None,
    }
    if not a.no_cpu_baseline:
        kind, path = cpu_library()
        lib = _abi.bind(path)
        s = _sna(lib, J0, coeff)
        xc, vc, t_cpu = cpu_md(lib, s, x, box, v, steps_cpu)
        log(f"CPU ({kind}, 1 core): {steps_cpu} steps in {t_cpu:.2f} s")
        if a.steps == steps_cpu:
            d = np.mod(st["x"] - xc + 0.5 * box, box) - 0.5 * box
            log(f"GPU vs CPU after {steps_cpu} steps: max |dx| {np.max(np.abs(d)):.3e} A, "
                f"max |dv| {np.max(np.abs(st['v'] - vc)):.3e} A/ps")
        out["cpu_baseline"] = dict(value=n * steps_cpu / t_cpu, unit="atom-steps/s", cores=1, kind=kind,
                                   sample=f"the whole {steps_cpu}-step trajectory ({n} atoms), host velocity "
                                          f"Verlet with the library's force call at every step, single thread, "
                                          f"{t_cpu:.2f} s", seconds_for_100_steps=t_cpu * 100 / steps_cpu)
    print(json.dumps(out), flush=True)


def main():
    if len(sys.argv) == 5 and sys.argv[1] == "--cpu-worker":
        cpu_worker(*sys.argv[2:])
        return
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", choices=("md", "fit", "surrogate", "config0"), default="md")
    ap.add_argument("--train-configs", type=int, default=4000)
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--sur-ncell", type=int, default=80)
    ap.add_argument("--train-tmin", type=float, default=300.0, help="surrogate: lowest training-run temperature (K)")
    ap.add_argument("--train-tmax", type=float, default=3000.0, help="surrogate: highest training-run temperature (K)")
    ap.add_argument("--train-kinds", choices=("both", "bcc"), default="both",
                    help="surrogate: training runs from BCC cells and random packings, or BCC cells only")
    ap.add_argument("--train-vacancies", type=int, default=0, help="surrogate: atoms taken out of each BCC training cell")
    ap.add_argument("--train-equil", type=int, default=10, help="surrogate: MD steps of a training run before its first sample")
    ap.add_argument("--train-every", type=int, default=10, help="surrogate: MD steps between a training run's samples")
    ap.add_argument("--train-sigma", type=float, default=0.05,
                    help="surrogate: RMS displacement (A) of the BCC training cells' starting positions")
    ap.add_argument("--potential", choices=("seeded", "W"), default="seeded",
                    help="surrogate workload: the seeded Ta coefficients or the upstream tungsten potential")
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--noaudit-steps", type=int, default=500)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--ncell", type=int, default=NCELL)
    ap.add_argument("--cpu-ncell", type=int, default=24, help="cells per side of the CPU baseline sample")
    ap.add_argument("--cpu-cores", type=int, default=0, help="processes of the all-cores CPU baseline (0: host_cores())")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py runs on exactly one GPU: the SNAP MD path is measured per GPU (multi-GPU spatial "
              "decomposition is not part of this build); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if a.workload == "fit":
        bench_fit(a)
        return
    if a.workload == "surrogate":
        bench_surrogate(a)
        return
    if a.workload == "config0":
        bench_config0(a)
        return

    import mdp_amd
    from mdp_amd.flops import dedr_key, snap_flops
    from mdp_amd.md import SnapMD

    mdp_amd.library()   # fails loudly without the HIP library
    x, box, v, coeff = workload(a.ncell)
    n = len(x)
    md = SnapMD(x, box, np.ones(n, np.int32), J, coeff, [0.5], [1.0], [MASS], rcutfac=RCUTFAC, rfac0=RFAC0,
                v=v, dt=DT, skin=SKIN, rebuild_every=EVERY)
    t0 = time.perf_counter()
    md.compute()
    log(f"first force call (incl. neighbour build and autotuning): {1e3 * (time.perf_counter() - t0):.1f} ms")
    log(f"autotuner: {json.dumps(md.tuning())}")
    md.run(a.warmup)
    # per-stage breakdown for the log, from a short run with every stage timed (outside the timed region)
    md.set_timing(True)
    md.run(min(a.steps, 50))
    stage_timers = md.timers()
    e0 = sum(md.energy())
    # the timed region: HIP events around the three SNAP compute kernels only (the roofline's launch time)
    md.set_timing(True, compute_only=True)
    t0 = time.perf_counter()
    md.run(a.steps)                       # returns after the stream is synchronised
    elapsed = time.perf_counter() - t0
    timers = md.timers()
    stats = md.stats()
    e1 = sum(md.energy())
    value = n * a.steps / elapsed
    log(f"{a.steps} steps in {elapsed:.3f} s -> {value:.4e} atom-steps/s; pairs/atom {stats['pairs'] / n:.2f}; "
        f"rebuilds {stats['rebuilds']} (dangerous {stats['dangerous_rebuilds']}); "
        f"energy drift {(e1 - e0) / n:.3e} eV/atom")
    log(f"per stage, {min(a.steps, 50)}-step run with every stage timed (before the timed region):")
    for k, (ms, cnt) in stage_timers.items():
        if cnt:
            log(f"  {k:16s} {ms:9.2f} ms total, {cnt:6d} launches, {ms / max(cnt, 1):8.4f} ms avg")
    log("timed region (events around the compute kernels only):")
    for k in ("compute_ui", "compute_yi", "compute_dedr"):
        ms, cnt = timers[k]
        log(f"  {k:16s} {ms:9.2f} ms total, {cnt:6d} launches, {ms / max(cnt, 1):8.4f} ms avg")

    fl = snap_flops(J)
    pairs_mean = md.mean_pairs()          # pairs per step averaged over the timed steps
    ppa = pairs_mean / n
    dkey = dedr_key(int(md.tuning().get("chosen", {}).get("compute_dedr", 9)))
    per_launch = {"compute_yi": fl["yi_md"] * n, "compute_dedr": fl[dkey] * pairs_mean,
                  "compute_ui": fl["ui"] * pairs_mean}
    dom = max(per_launch, key=lambda k: timers[k][0])
    ms_avg = timers[dom][0] / max(timers[dom][1], 1)
    achieved = per_launch[dom] / (ms_avg * 1e-3) / 1e12
    # HBM bytes per launch of the dominant kernel, measured on this workload with rocprofv3's L2 read and
    # write request counters by size, which count exact bytes (profiles/traffic_<kernel>_bcc40_j8.*)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_bcc40_j8.json")
    if os.path.exists(tfile) and a.ncell == NCELL:
        traffic = json.load(open(tfile)).get("bytes_per_launch")
    step_flops = n * (ppa * (fl["ui"] + fl[dkey]) + fl["yi_md"])
    log(f"dominant kernel {dom}: {ms_avg:.3f} ms/launch, {achieved:.2f} TFLOP/s "
        f"({100 * achieved / FP64_PEAK_TFLOPS:.1f}% of FP64 peak); whole step "
        f"{step_flops / (elapsed / a.steps) / 1e12:.2f} TFLOP/s")
    # The algorithmic count is the upstream compute_yi's.  compute_yi variants 115 and 116 skip the terms that
    # the swap symmetry of the (j1, j1) couplings makes redundant, so the FP64 work they execute is smaller.
    executed = None
    if dom == "compute_yi" and int(md.tuning().get("chosen", {}).get("compute_yi_waves", 0)) in (115, 116):
        executed = fl["yi_md_sym"] * n
        log(f"  compute_yi executes {executed / per_launch[dom]:.3f} of the algorithmic flops: "
            f"{executed / (ms_avg * 1e-3) / 1e12:.2f} TFLOP/s executed")

    out = {
        "metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * elapsed / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded BCC lattice, displacements, velocities and SNAP coefficients)",
        "config": {"workload": f"plain FP64 SNAP NVE MD, BCC Ta {a.ncell}^3 cells ({n} atoms), twojmax=8, "
                               f"dt=0.5 fs, skin 0.3 A rebuilt every {EVERY} steps",
                   "atoms": n, "twojmax": J, "pairs_per_atom": round(ppa, 3), "parallelism": "1 GPU",
                   "autotuned": md.tuning().get("chosen", {})},
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "pipe": "FP64 vector (FMA; the FP64 vector and matrix peaks of MI355X are the same)",
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "flops_per_launch": per_launch[dom], "ms_per_launch": ms_avg},
        "cpu_baseline": None,
    }
    if executed is not None:
        out["roofline"]["executed_flops_per_launch"] = executed
        out["roofline"]["executed_frac"] = executed / (ms_avg * 1e-3) / 1e12 / FP64_PEAK_TFLOPS
    if not a.no_cpu_baseline:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(a.cpu_ncell, x, box, a.cpu_cores or None)
        log(f"cpu baseline: {out['cpu_baseline']['value']:.1f} atom-steps/s ({time.perf_counter() - t0:.1f} s)")
    md.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
