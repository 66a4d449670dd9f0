"""The upstream application's binary inputs: writing them as the Julia preprocessing does, reading them
through libmdpsnap (include/mdpsnap.h part 7), and running the application's MD on the GPU.

Writers mirror src/Julia/Preprocessing/writeapp.jl:8-137, writeconfig.jl:8-48, writelattice.jl,
writeregion.jl and writedomain.jl: a float64 stream of the size-table length, the size table, then every
field's values (column-major, as Julia's ``[:]``).  ``snap_app`` fills the application fields the way
Preprocessing.jl:60-321 derives them from the user's settings (flag, ndims, simparam, solparam, snaparam,
squared cutoffs, the 0 placed before the masses), for a SNAP application with an optional ZBL pair
potential such as simulating/snap/snapTA06A.jl.
"""
import ctypes

import numpy as np

from . import library

APP_FIELDS = ["ndims", "flag", "bcs", "pbc", "boxoffset", "atomnumbers", "atommasses", "atomcharges", "simparam",
              "solparam", "eta", "kappa", "muml", "mu1a", "mu1b", "mu2a", "mu2b", "mu2c", "mu3a", "mu3b", "mu3c",
              "mu4a", "mu4b", "pot1a", "pot1b", "pot2a", "pot2b", "pot2c", "pot3a", "pot3b", "pot3c", "pot4a",
              "pot4b", "rcutsqml", "rcutsq2a", "rcutsq2b", "rcutsq2c", "rcutsq3a", "rcutsq3b", "rcutsq3c",
              "rcutsq4a", "rcutsq4b", "rcutsqmax", "atom1b", "atom2b", "atom2c", "atom3b", "atom3c", "atom4b", None,
              "traininglist", "validatelist", "nveparam", "nvtparam", "snaparam", "snapelemradius",
              "snapelemweight", "snapcoeff", "createvelocity", "setatomtypefraction"]
LATTICE_STYLES = {"none": 0, "sc": 1, "bcc": 2, "fcc": 3, "hcp": 4, "diamond": 5, "sq": 6, "sq2": 7, "hex": 8,
                  "custom": 9}


def _flat(a):
    """Julia's a[:] of a row vector or matrix (column-major)."""
    return np.asarray(a, np.float64).ravel(order="F")


def _write(path, nsize, arrays, ntable):
    table = np.zeros(ntable)
    table[:len(nsize)] = nsize
    with open(path, "wb") as fh:
        fh.write(np.float64(ntable).tobytes())
        fh.write(table.astype(np.float64).tobytes())
        for a in arrays:
            fh.write(_flat(a).tobytes())


def write_app(fields, path):
    """writeapp.jl: ``fields`` maps the APP_FIELDS names to arrays (missing ones are empty)."""
    nsize, arrays = [], []
    for name in APP_FIELDS:
        if name is None:
            nsize.append(0)
            continue
        a = _flat(fields.get(name, []))
        nsize.append(a.size)
        arrays.append(a)
    _write(path, nsize, arrays, 100)


def write_config(natom, a, b, c, t, x, path, e=(), q=(), v=(), f=(), we=(), wf=(), ncq=0):
    """writeconfig.jl: one or more configurations (a, b, c: (3, nconfig); x: (3, N); t: (1, N))."""
    a, b, c = (np.asarray(u, np.float64).reshape(3, -1, order="F") for u in (a, b, c))
    x = np.asarray(x, np.float64).reshape(3, -1, order="F")
    v = np.asarray(v, np.float64)
    f = np.asarray(f, np.float64)
    tmp = [len(np.atleast_1d(natom)), 3, ncq, 3 if v.size else 0, 1 if len(e) else 0, 3 if f.size else 0]
    arrays = [tmp, np.atleast_1d(natom), a, b, c, np.asarray(e, np.float64), np.asarray(t, np.float64), x,
              np.asarray(q, np.float64), v, f, np.asarray(we, np.float64), np.asarray(wf, np.float64)]
    _write(path, [_flat(u).size for u in arrays], arrays, 20)


def write_lattice(path, style, scale, orientx=(1, 0, 0), orienty=(0, 1, 0), orientz=(0, 0, 1), spacing=(1, 1, 1),
                  origin=(0, 0, 0), a1=(1, 0, 0), a2=(0, 1, 0), a3=(0, 0, 1), basis=None, types=None):
    """setlattice.jl + writelattice.jl (spaceflag 0, as setlattice sets it)."""
    st = LATTICE_STYLES[style.lower()]
    if basis is None:
        basis = {1: [[0, 0, 0]], 2: [[0, 0, 0], [0.5, 0.5, 0.5]],
                 3: [[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]],
                 4: [[0, 0, 0], [0.5, 0.5, 0], [0.5, 5 / 6, 0.5], [0, 1 / 3, 0.5]],
                 5: [[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5], [0.25, 0.25, 0.25], [0.25, 0.75, 0.75],
                     [0.75, 0.25, 0.75], [0.75, 0.75, 0.25]],
                 6: [[0, 0, 0]], 7: [[0, 0, 0], [0.5, 0.5, 0]], 8: [[0, 0, 0], [0.5, 0.5, 0]]}[st]
    basis = np.asarray(basis, np.float64).reshape(-1, 3)
    types = np.ones(len(basis)) if types is None else np.asarray(types, np.float64)
    tmp = [st, len(basis), 0, scale]
    arrays = [tmp, origin, orientx, orienty, orientz, spacing, a1, a2, a3, basis.ravel(), types]
    _write(path, [_flat(u).size for u in arrays], arrays, 11)


def write_region(path, boxhi, boxtilt=(0, 0, 0)):
    """setregion.jl + writeregion.jl (boxlo = 0)."""
    tri = 0 if all(t == 0 for t in boxtilt) else 1
    arrays = [[tri], [0, 0, 0], boxhi, boxtilt]
    _write(path, [_flat(u).size for u in arrays], arrays, 4)


def snap_app(snapcoeff, twojmax, rcutfac, atommasses, ntimesteps, dt, pot2a=None, mu2a=(), rcut2a=(),
             rfac0=0.99363, rmin0=0.0, bzeroflag=0, switchflag=1, quadraticflag=0, chemflag=0, bnormflag=0,
             wselfallflag=0, elemradius=(0.5,), elemweight=(1.0,), neighskin=1.0, neighmax=100, neighevery=1,
             neighdelay=0, neighcheck=1, globalfreq=10, ensemble="nve", nvtparam=(), createvelocity=(),
             pbc=(1, 1, 1), time=0.0, neighcell=0, neighpair=0, natomtype=None):
    """The app.bin fields of a SNAP application (Preprocessing.jl:60-321 with the initializeapp.jl
    defaults): descriptor snap, metal units, ML potential (rcutml = rcutfac), full neighbour list."""
    snapcoeff = np.asarray(snapcoeff, np.float64).ravel()
    nelem = len(elemradius)
    masses = np.atleast_1d(np.asarray(atommasses, np.float64))
    natomtype = natomtype or len(masses)
    ens = {"nve": 0, "nvelimit": 1, "nvt": 2}[ensemble]
    rcut = [rcutfac] + list(np.atleast_1d(rcut2a)) if len(np.atleast_1d(rcut2a)) else [rcutfac]
    rcutmax = max(rcut) + neighskin
    ndims = np.zeros(20)
    ndims[[0, 3, 6, 8, 10, 11, 12, 13, 14]] = [3, ntimesteps, 0, natomtype, neighmax, neighevery, neighdelay,
                                               globalfreq, 0]
    # descriptor, spectrum, training, runMD, potentialform, neighpair, energycal, forcecal, stresscal,
    # neighcell, decomposition, chemtype, dftdata, unitstyle, ensemblemode, neighcheck
    flag = [1, 1, 0, 1, 1 if pot2a is None else 2, neighpair, 0, 0, 0, neighcell, 0, 0, 0, 2, ens, neighcheck]
    f = dict(ndims=ndims, flag=flag, bcs=np.zeros(6), pbc=pbc, boxoffset=[rcutmax] * 3,
             atomnumbers=[0], atommasses=np.concatenate([[0.0], masses]), atomcharges=[0.0],
             simparam=[time, dt], solparam=[rcutmax, neighskin],
             rcutsqml=[rcutfac ** 2], rcutsqmax=[rcutmax ** 2],
             snaparam=[nelem, len(snapcoeff) // nelem, twojmax, rcutfac, rfac0, rmin0, bzeroflag, switchflag,
                       quadraticflag, chemflag, bnormflag, wselfallflag],
             snapelemradius=elemradius, snapelemweight=elemweight, snapcoeff=snapcoeff,
             nvtparam=nvtparam, createvelocity=createvelocity)
    if pot2a is not None:
        f.update(pot2a=pot2a, mu2a=mu2a, rcutsq2a=np.asarray(rcut2a, np.float64) ** 2)
    return f


_bound = False


def _lib():
    global _bound
    c = library().cdll
    if not _bound:
        vp = ctypes.c_void_p
        c.mdp_app_read.argtypes = [ctypes.POINTER(vp), ctypes.c_char_p]
        c.mdp_app_destroy.argtypes = [vp]
        c.mdp_app_field.argtypes = [vp, ctypes.c_char_p, vp, ctypes.c_long]
        c.mdp_app_field.restype = ctypes.c_long
        c.mdp_app_has.argtypes = [vp, vp]
        c.mdp_app_atoms.argtypes = [vp, vp, vp, vp, ctypes.c_long]
        c.mdp_app_atoms.restype = ctypes.c_long
        c.mdp_md_create_app.argtypes = [ctypes.POINTER(vp), vp]
        _bound = True
    return c


class App:
    """An application's inputs read by libmdpsnap from ``prefix`` (prefix + "app.bin", ...)."""

    def __init__(self, prefix):
        self._c = _lib()
        h = ctypes.c_void_p()
        if self._c.mdp_app_read(ctypes.byref(h), str(prefix).encode()) != 0:
            raise RuntimeError(f"mdp_app_read: {self._c.mdp_snap_last_error().decode()}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._c.mdp_app_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, name):
        n = self._c.mdp_app_field(self._h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(f"{name}: {self._c.mdp_snap_last_error().decode()}")
        out = np.zeros(max(n, 1))
        self._c.mdp_app_field(self._h, name.encode(), out.ctypes.data, n)
        return out[:n]

    def has(self):
        fl = np.zeros(4, np.int32)
        self._c.mdp_app_has(self._h, fl.ctypes.data)
        return dict(zip(("config", "lattice", "region", "domain"), map(bool, fl)))

    def atoms(self):
        """(x (N, 3), types (N,), box (3,)) the application starts from."""
        box = np.zeros(3)
        n = self._c.mdp_app_atoms(self._h, None, None, box.ctypes.data, 0)
        if n < 0:
            raise RuntimeError(f"mdp_app_atoms: {self._c.mdp_snap_last_error().decode()}")
        x, t = np.zeros((n, 3)), np.zeros(n, np.int32)
        if self._c.mdp_app_atoms(self._h, x.ctypes.data, t.ctypes.data, box.ctypes.data, n) != n:
            raise RuntimeError(f"mdp_app_atoms: {self._c.mdp_snap_last_error().decode()}")
        return x, t, box

    def engine(self):
        """The MD engine the application describes (mdp_amd.md.SnapMD)."""
        from .md import SnapMD
        h = ctypes.c_void_p()
        if self._c.mdp_md_create_app(ctypes.byref(h), self._h) != 0:
            raise RuntimeError(f"mdp_md_create_app: {self._c.mdp_snap_last_error().decode()}")
        x, t, box = self.atoms()
        return SnapMD._from_handle(h, x, box, t)

    def run(self, steps=None, log=None):
        """The application's MD: ntimesteps (or ``steps``) steps with thermo output (step, temperature,
        potential and kinetic energy per atom, pressure) every globalfreq steps, as Integration.cpp:537
        prints it.  Returns (engine, list of thermo dicts)."""
        md = self.engine()
        nd = self.field("ndims")
        total = int(nd[3]) if steps is None else int(steps)
        every = max(1, int(nd[13]))
        md.compute()
        rows = [dict(step=0, **md.thermo())]
        done = 0
        while done < total:
            k = min(every, total - done)
            md.run(k)
            done += k
            rows.append(dict(step=done, **md.thermo()))
            if log:
                r = rows[-1]
                log(f"step {done}: T {r['temp']:.3f} K, pe {r['pe']:.6f} eV/atom, ke {r['ke']:.6f} eV/atom, "
                    f"press {r['press']:.3f} bar")
        return md, rows
