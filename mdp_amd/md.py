"""Python handle on the device-resident SNAP MD engine of libmdpsnap (mdp_md_* in include/mdpsnap.h).

The engine plays the role of the upstream velocity-Verlet drivers (src/C++/Integration/
Integration.cpp:17-156, src/Julia/Integration/velocityverlet.jl:3-113) around the SNAP force call:
positions, velocities, neighbour lists and all SNAP intermediates stay on the GPU.
"""
import ctypes

import numpy as np

from . import library


class _Params(ctypes.Structure):
    _fields_ = [("twojmax", ctypes.c_int), ("ntypes", ctypes.c_int),
                ("rcutfac", ctypes.c_double), ("rfac0", ctypes.c_double), ("rmin0", ctypes.c_double),
                ("switchflag", ctypes.c_int), ("bzeroflag", ctypes.c_int), ("bnormflag", ctypes.c_int),
                ("wselfallflag", ctypes.c_int),
                ("elemradius", ctypes.POINTER(ctypes.c_double)), ("elemweight", ctypes.POINTER(ctypes.c_double)),
                ("coeff", ctypes.POINTER(ctypes.c_double)), ("mass", ctypes.POINTER(ctypes.c_double)),
                ("dt", ctypes.c_double), ("skin", ctypes.c_double), ("rebuild_every", ctypes.c_int)]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


_bound = False


def _bind():
    global _bound
    c = library().cdll
    if _bound:
        return c
    vp = ctypes.c_void_p
    c.mdp_md_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(_Params), ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    for fn in ("mdp_md_destroy", "mdp_md_compute"):
        getattr(c, fn).argtypes = [vp]
    c.mdp_md_run.argtypes = [vp, ctypes.c_int]
    c.mdp_md_get.argtypes = [vp] + [ctypes.c_void_p] * 5
    c.mdp_md_neighbors.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_long)]
    c.mdp_md_stats.argtypes = [vp, ctypes.c_void_p]
    c.mdp_md_set_timing.argtypes = [vp, ctypes.c_int]
    c.mdp_md_timers.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    c.mdp_md_timer_name.restype = ctypes.c_char_p
    c.mdp_md_timer_name.argtypes = [ctypes.c_int]
    c.mdp_md_num_timers.argtypes = []
    c.mdp_md_energy.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    c.mdp_md_tuning.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    c.mdp_md_set_surrogate.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    c.mdp_md_clear_surrogate.argtypes = [vp]
    c.mdp_md_audits.argtypes = [vp, vp]
    c.mdp_md_mean_pairs.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    c.mdp_md_set_nvt.argtypes = [vp] + [ctypes.c_double] * 5 + [ctypes.c_int] * 2 + [ctypes.c_long] * 3
    c.mdp_md_clear_nvt.argtypes = [vp]
    c.mdp_md_set_zbl.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    c.mdp_md_set_state.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p]
    c.mdp_md_thermo.argtypes = [vp, ctypes.c_void_p]
    c.mdp_md_nvt_state.argtypes = [vp, ctypes.c_void_p]
    c.mdp_velocity_create.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_double, vp]
    c.mdp_velocity_create_loop.argtypes = [vp] * 8 + [ctypes.c_int] * 8
    _bound = True
    return c


class SnapMD:
    """One periodic orthogonal system integrated NVE with plain FP64 SNAP on the GPU.

    ``coeff`` holds [beta0, beta_1..beta_ncoeff] (first block used for every atom, as upstream does
    with chemflag = 0); ``dt`` in ps, positions in Angstrom, velocities in Angstrom/ps, masses in amu.
    """

    def __init__(self, x, box, types, twojmax, coeff, elemradius, elemweight, mass, rcutfac=4.67637,
                 rfac0=0.99363, rmin0=0.0, switchflag=1, bzeroflag=1, bnormflag=0, wselfallflag=0, v=None,
                 dt=0.0005, skin=0.3, rebuild_every=10):
        self._c = _bind()
        self.x0 = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        self.n = len(self.x0)
        self.box = np.ascontiguousarray(box, np.float64).reshape(3)
        self.types = np.ascontiguousarray(types, np.int32).reshape(self.n)
        v = np.zeros_like(self.x0) if v is None else np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        self._keep = [np.ascontiguousarray(elemradius, np.float64).ravel(),
                      np.ascontiguousarray(elemweight, np.float64).ravel(),
                      np.ascontiguousarray(coeff, np.float64).ravel(),
                      np.ascontiguousarray(mass, np.float64).ravel()]
        ntypes = len(self._keep[0])
        p = _Params(twojmax, ntypes, rcutfac, rfac0, rmin0, switchflag, bzeroflag, bnormflag, wselfallflag,
                    _dp(self._keep[0]), _dp(self._keep[1]), _dp(self._keep[2]), _dp(self._keep[3]),
                    dt, skin, rebuild_every)
        h = ctypes.c_void_p()
        rc = self._c.mdp_md_create(ctypes.byref(h), ctypes.byref(p), self.n, self.box.ctypes.data,
                                   self.x0.ctypes.data, v.ctypes.data, self.types.ctypes.data)
        self._check(rc, "mdp_md_create")
        self._h = h

    @classmethod
    def _from_handle(cls, h, x, box, types):
        """Wrap an engine handle made by another entry point (mdp_md_create_app)."""
        self = cls.__new__(cls)
        self._c = _bind()
        self.x0 = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        self.n = len(self.x0)
        self.box = np.ascontiguousarray(box, np.float64).reshape(3)
        self.types = np.ascontiguousarray(types, np.int32).reshape(self.n)
        self._keep = []
        self._h = h
        return self

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {library().cdll.mdp_snap_last_error().decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._c.mdp_md_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def compute(self):
        """Rebuild neighbours and evaluate energy/forces/virials at the current positions."""
        self._check(self._c.mdp_md_compute(self._h), "mdp_md_compute")

    def run(self, nsteps):
        self._check(self._c.mdp_md_run(self._h, int(nsteps)), "mdp_md_run")

    def state(self):
        n = self.n
        x, v, f = (np.zeros((n, 3)) for _ in range(3))
        e, w = np.zeros(n), np.zeros((n, 6))
        self._check(self._c.mdp_md_get(self._h, x.ctypes.data, v.ctypes.data, f.ctypes.data, e.ctypes.data,
                                       w.ctypes.data), "mdp_md_get")
        return dict(x=x, v=v, f=f, eatom=e, vatom=w)

    def neighbors(self):
        """(offsets[n+1], j[E], shift[E,3]) of the last neighbour build."""
        tot = ctypes.c_long()
        self._check(self._c.mdp_md_neighbors(self._h, None, None, None, ctypes.byref(tot)), "mdp_md_neighbors")
        off = np.zeros(self.n + 1, np.int32)
        nj = np.zeros(max(tot.value, 1), np.int32)
        ns = np.zeros((max(tot.value, 1), 3), np.int32)
        self._check(self._c.mdp_md_neighbors(self._h, off.ctypes.data, nj.ctypes.data, ns.ctypes.data,
                                             ctypes.byref(tot)), "mdp_md_neighbors")
        return off, nj[:tot.value], ns[:tot.value]

    def stats(self):
        out = np.zeros(6, np.int64)
        self._check(self._c.mdp_md_stats(self._h, out.ctypes.data), "mdp_md_stats")
        keys = ("pairs", "neighbor_entries", "rebuilds", "dangerous_rebuilds", "steps", "overflows")
        return dict(zip(keys, (int(v) for v in out)))

    def set_timing(self, on=True, compute_only=False):
        """Per-kernel HIP-event timers from now on (every kernel class, or with compute_only the three SNAP
        compute kernels alone, which puts fewer events into a timed run); on=False switches them off."""
        mode = 0 if not on else (2 if compute_only else 1)
        self._check(self._c.mdp_md_set_timing(self._h, mode), "mdp_md_set_timing")

    def timers(self):
        k = self._c.mdp_md_num_timers()
        ms = np.zeros(k)
        n = np.zeros(k, np.int64)
        self._check(self._c.mdp_md_timers(self._h, ms.ctypes.data, n.ctypes.data, k), "mdp_md_timers")
        return {self._c.mdp_md_timer_name(i).decode(): (float(ms[i]), int(n[i])) for i in range(k)}

    def tuning(self):
        """Launch variants the autotuner chose (and their measured times), as a dict."""
        import json
        buf = ctypes.create_string_buffer(4096)
        self._check(self._c.mdp_md_tuning(self._h, buf, 4096), "mdp_md_tuning")
        return json.loads(buf.value.decode() or "{}")

    def set_surrogate(self, model, audit_every=50, audit_tol=0.05, twojmax=4):
        """Use the learned surrogate ``model`` (mdp_amd.surrogate.Model) for forces from now on, with a
        plain-SNAP audit every ``audit_every`` steps (plain forces replace the surrogate's on an audit
        step whose RMS force difference exceeds ``audit_tol`` eV/A)."""
        self._sur_keep = (np.ascontiguousarray(model.params, np.float32), np.ascontiguousarray(model.aff, np.float64))
        self._check(self._c.mdp_md_set_surrogate(self._h, self._sur_keep[0].ctypes.data, self._sur_keep[1].ctypes.data,
                                                 int(model.nin), int(twojmax), int(audit_every), float(audit_tol)),
                    "mdp_md_set_surrogate")

    def clear_surrogate(self):
        self._check(self._c.mdp_md_clear_surrogate(self._h), "mdp_md_clear_surrogate")

    def mean_pairs(self):
        """Mean pair count per force evaluation since set_timing() was last called."""
        m = ctypes.c_double()
        self._check(self._c.mdp_md_mean_pairs(self._h, ctypes.byref(m)), "mdp_md_mean_pairs")
        return m.value

    def audits(self):
        out = np.zeros(5)
        self._check(self._c.mdp_md_audits(self._h, out.ctypes.data), "mdp_md_audits")
        return dict(audits=int(out[0]), fallbacks=int(out[1]), last_rmse=float(out[2]), max_rmse=float(out[3]),
                    mean_rmse=float(out[4]))

    def energy(self):
        ep, ek = ctypes.c_double(), ctypes.c_double()
        self._check(self._c.mdp_md_energy(self._h, ctypes.byref(ep), ctypes.byref(ek)), "mdp_md_energy")
        return ep.value, ek.value

    def set_nvt(self, t_start, t_stop=None, t_period=0.1, drag=0.0, tdof=0.0, mtchain=3, nc_tchain=1,
                beginstep=0, endstep=1, currentstep=0):
        """Nose-Hoover chain NVT for the following steps (cpuVelocityIntegration.cpp:505-648)."""
        t_stop = t_start if t_stop is None else t_stop
        self._check(self._c.mdp_md_set_nvt(self._h, t_start, t_stop, t_period, drag, tdof, mtchain, nc_tchain,
                                           beginstep, endstep, currentstep), "mdp_md_set_nvt")

    def thermo(self):
        """ThermoOutput of the upstream app: dict pe (eV/atom), temp (K), ke (eV/atom), press (bar)."""
        o = np.zeros(4)
        self._check(self._c.mdp_md_thermo(self._h, o.ctypes.data), "mdp_md_thermo")
        return dict(pe=o[0], temp=o[1], ke=o[2], press=o[3])

    def set_state(self, x, v=None):
        """New positions (and velocities) from the host; forces are recomputed at the next step."""
        x = np.ascontiguousarray(x, np.float64).reshape(self.n, 3)
        v = None if v is None else np.ascontiguousarray(v, np.float64).reshape(self.n, 3)
        self._check(self._c.mdp_md_set_state(self._h, x.ctypes.data, None if v is None else v.ctypes.data),
                    "mdp_md_set_state")

    def set_zbl(self, cut_inner=4.0, cut_outer=4.8, z=73.0):
        """Add the ZBL reference pair potential (simulating/snap/ZBLpotential.jl; the Ta application's
        parameters by default, snapTA06A.jl:62-66) to the forces from the next evaluation on."""
        self._check(self._c.mdp_md_set_zbl(self._h, float(cut_inner), float(cut_outer), float(z)), "mdp_md_set_zbl")

    def clear_nvt(self):
        self._check(self._c.mdp_md_clear_nvt(self._h), "mdp_md_clear_nvt")

    def nvt_state(self):
        """dict of the chain: eta_mass, eta, eta_dot, eta_dotdot (10 each), t_current, t_target, step."""
        o = np.zeros(43)
        self._check(self._c.mdp_md_nvt_state(self._h, o.ctypes.data), "mdp_md_nvt_state")
        return dict(eta_mass=o[0:10], eta=o[10:20], eta_dot=o[20:30], eta_dotdot=o[30:40], t_current=o[40],
                    t_target=o[41], step=int(o[42]))


def velocity_create(x, types, mass, t_desired, seed0, dist_flag=0, loop_flag=2, mvv2e=1.0364269e-4,
                    boltz=8.617343e-5):
    """Initial velocities (N, 3) as the upstream app creates them (readdata.cpp:609-640, loop mode "geom");
    ``mass`` per type, 1-based (mass[0] unused)."""
    c = _bind()
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    types = np.ascontiguousarray(types, np.int32).reshape(len(x))
    mass = np.ascontiguousarray(mass, np.float64).ravel()
    v = np.zeros_like(x)
    rc = c.mdp_velocity_create(x.ctypes.data, types.ctypes.data, mass.ctypes.data, len(x), float(t_desired),
                               int(seed0), int(dist_flag), int(loop_flag), mvv2e, boltz, v.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"mdp_velocity_create: {library().cdll.mdp_snap_last_error().decode()}")
    return v


class RandomState:
    """The upstream app's generator state (commonstruct seed/save/second): Park-Miller state and a
    waiting normal deviate."""

    def __init__(self, seed, save=0, second=0.0):
        self.seed = np.array([seed], np.int32)
        self.save = np.array([save], np.int32)
        self.second = np.array([second], np.float64)


def velocity_draws(x, v, types, mass, state, seed0=0, sum_flag=0, dist_flag=0, loop_flag=0, mpi_rank=0, natoms=None,
                   idmap=None):
    """cpuVelocityCreate (cpuVelocityIntegration.cpp:54-135) in any loop mode, on ``v`` (N, 3) in place:
    loop_flag 0 ("all": every atom of the system draws in global-ID order; idmap[k] for k = 1..natoms
    names the local index, default k - 1), 1 ("local": seed + mpi_rank, 100 discarded draws, then local
    order) or 2 ("geom": each atom's stream seeded from seed0 and its coordinates).  ``state`` (a
    RandomState) is advanced as the upstream generator would be."""
    c = _bind()
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    if not (isinstance(v, np.ndarray) and v.dtype == np.float64 and v.flags.c_contiguous and v.shape == x.shape):
        raise ValueError("v must be a C-contiguous float64 array shaped like x")
    n = len(x)
    natoms = n if natoms is None else int(natoms)
    types = np.ascontiguousarray(types, np.int32).reshape(n)
    mass = np.ascontiguousarray(mass, np.float64).ravel()
    if idmap is None:
        idmap = np.concatenate([[-1], np.arange(natoms)])
    idmap = np.ascontiguousarray(idmap, np.int32)
    if len(idmap) < natoms + 1:
        raise ValueError("idmap needs natoms + 1 entries")
    rc = c.mdp_velocity_create_loop(x.ctypes.data, v.ctypes.data, mass.ctypes.data, state.second.ctypes.data,
                                    state.seed.ctypes.data, state.save.ctypes.data, idmap.ctypes.data,
                                    types.ctypes.data, int(seed0), int(sum_flag), int(dist_flag), int(loop_flag), 3,
                                    int(mpi_rank), n, natoms)
    if rc != 0:
        raise RuntimeError(f"mdp_velocity_create_loop: {library().cdll.mdp_snap_last_error().decode()}")
    return v
