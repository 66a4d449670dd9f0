"""The upstream application's binary inputs (include/mdpsnap.h part 7) on the CPU: files written as the
Julia preprocessing writes them (mdp_amd.appfile, mirroring writeapp.jl, writeconfig.jl, writelattice.jl,
writeregion.jl) read back field by field through libmdpsnap, and the atoms the application starts from:
the first configuration of config.bin, or the lattice points inside the region, bit for bit against the
upstream C++ core's own lattice functions compiled from source (cpuLattice, cpuLatticeBoundingBox,
cpuAtomLattice, cpuAtomAdd; oracle/_ref/libcpucore_r2.so) for several lattice styles and orientations."""
import ctypes
import os

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from mdp_amd import appfile

R2 = os.path.join(ROOT, "oracle", "_ref", "libcpucore_r2.so")
TA_COEFF = np.loadtxt(os.path.join(GOLDEN, "snapTA06A.coeff"))


def _ta_app(tmp_path, nlattice=4, style="bcc", a=3.316, **kw):
    """simulating/snap/snapTA06A.jl: bcc Ta, twojmax 6, ZBL [4, 4.8, 73], 300 K geom velocities."""
    prefix = str(tmp_path / "snapTA06A")
    f = appfile.snap_app(TA_COEFF, 6, 4.67637, [180.88], 100, 0.5e-3, pot2a=[1], mu2a=[4.0, 4.8, 73.0],
                         rcut2a=[4.8], createvelocity=[300, 4928459, 0, 0, 2, 1, 0], **kw)
    appfile.write_app(f, prefix + "app.bin")
    appfile.write_lattice(prefix + "lattice.bin", style, a)
    appfile.write_region(prefix + "region.bin", [nlattice] * 3)
    return prefix, f


def test_app_fields_round_trip(tmp_path):
    prefix, f = _ta_app(tmp_path)
    app = appfile.App(prefix)
    assert app.has() == dict(config=False, lattice=True, region=True, domain=False)
    for name, want in f.items():
        got = app.field(name)
        assert np.array_equal(got, np.asarray(want, np.float64).ravel(order="F")), name
    assert app.field("pot1a").size == 0
    assert app.field("snapcoeff").size == 31 and app.field("snapcoeff")[0] == TA_COEFF[0]
    with pytest.raises(KeyError):
        app.field("no_such_field")


def test_app_lattice_atoms_bcc(tmp_path):
    prefix, _ = _ta_app(tmp_path, nlattice=4)
    x, t, box = appfile.App(prefix).atoms()
    assert len(x) == 128 and np.all(t == 1)
    assert np.allclose(box, 4 * 3.316, rtol=0, atol=1e-12)
    # the 2 basis points of every cubic cell, k-major, as upstream's loops visit them
    cells = np.array([[i, j, k] for k in range(4) for j in range(4) for i in range(4)], float)
    want = (cells[:, None, :] + np.array([[0, 0, 0], [0.5, 0.5, 0.5]])[None]).reshape(-1, 3) * 3.316
    assert np.max(np.abs(x - want)) < 1e-12


@pytest.mark.parametrize("style,scale,orient,region", [
    ("bcc", 3.316, [[1, 0, 0], [0, 1, 0], [0, 0, 1]], [4, 4, 4]),
    ("fcc", 4.05, [[1, 1, 0], [-1, 1, 0], [0, 0, 1]], [3, 2.5, 2]),
    ("hcp", 2.95, [[1, 0, 0], [0, 1, 0], [0, 0, 1]], [3, 2, 2]),
    ("diamond", 5.43, [[1, 1, 1], [1, -1, 0], [1, 1, -2]], [2, 2, 1]),
    ("sc", 2.0, [[1, 0, 0], [0, 1, 0], [0, 0, 1]], [3.7, 1, 2.2]),
])
def test_app_lattice_atoms_vs_upstream_core(tmp_path, style, scale, orient, region):
    if not os.path.exists(R2):
        pytest.skip("upstream C++ core not built (oracle/_ref)")
    prefix = str(tmp_path / "lat")
    appfile.write_app(appfile.snap_app(TA_COEFF, 6, 4.67637, [180.88], 10, 0.5e-3), prefix + "app.bin")
    appfile.write_lattice(prefix + "lattice.bin", style, scale, orientx=orient[0], orienty=orient[1],
                          orientz=orient[2])
    appfile.write_region(prefix + "region.bin", region)
    x, t, box = appfile.App(prefix).atoms()
    r2 = ctypes.CDLL(R2).r2_app_lattice_atoms
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C")
    r2.argtypes = [dp, ip, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ip, dp, dp, dp, dp, ip,
                   ctypes.c_int, dp, dp, dp]
    cap = 100000
    xr, tr, sp = np.zeros(3 * cap), np.zeros(cap, np.int32), np.zeros(3)
    n = r2(xr, tr, cap, appfile.LATTICE_STYLES[style], scale, 0, np.array(orient, np.int32).ravel(), np.zeros(3),
           np.ones(3), np.eye(3).ravel(), np.zeros(3), np.ones(8, np.int32), 0, np.zeros(3),
           np.array(region, float), sp)
    assert n == len(x) > 0
    assert np.array_equal(x, xr[:3 * n].reshape(-1, 3))         # bit for bit (box origin is 0)
    assert np.array_equal(t, tr[:n])
    assert np.array_equal(box, sp * np.array(region, float))


def test_app_config_atoms(tmp_path):
    prefix = str(tmp_path / "cfg")
    appfile.write_app(appfile.snap_app(TA_COEFF, 6, 4.67637, [180.88], 10, 0.5e-3), prefix + "app.bin")
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 10, (3, 20))
    appfile.write_config([20], [10.0, 0, 0], [0, 11.0, 0], [0, 0, 12.0], np.ones((1, 20)), x, prefix + "config.bin")
    app = appfile.App(prefix)
    assert app.has()["config"]
    xa, ta, box = app.atoms()
    assert np.array_equal(xa, x.T) and np.all(ta == 1) and np.array_equal(box, [10.0, 11.0, 12.0])
    # a tilted cell is refused (the MD engine's box is orthogonal)
    appfile.write_config([20], [10.0, 0, 0], [0.5, 11.0, 0], [0, 0, 12.0], np.ones((1, 20)), x, prefix + "config.bin")
    with pytest.raises(RuntimeError, match="orthogonal"):
        appfile.App(prefix).atoms()
