"""The upstream Ta application (simulating/snap/snapTA06A.jl: bcc Ta 4^3 cells, twojmax=6 SNAP with the
coefficients of snapTA06A.coeff, the ZBL pair potential [4, 4.8, 73], 300 K geom velocities, NVE,
dt = 0.5 fs, 100 steps, thermo every 10) driven by its binary inputs through mdp_md_create_app:

* the engine starts from the application's atoms and initial velocities (mdp_velocity_create's sequence,
  itself bit-identical with the upstream app's);
* its first energies, forces and virials equal the oracle's SNAP plus the oracle's ZBL (1e-10);
* the 100-step run conserves the total energy, and its thermo rows come every globalfreq steps;
* an NVT application (nvtparam) drives the thermostat toward its target.
"""
import os

import numpy as np
import pytest

import zbl_oracle
from conftest import GOLDEN, relerr
from mdp_amd import appfile, lattice, snapcpp

pytestmark = pytest.mark.gpu

TA_COEFF = np.loadtxt(os.path.join(GOLDEN, "snapTA06A.coeff"))


def _ta(tmp_path, **kw):
    prefix = str(tmp_path / "snapTA06A4")
    f = appfile.snap_app(TA_COEFF, 6, 4.67637, [180.88], 100, 0.5e-3, pot2a=[1], mu2a=[4.0, 4.8, 73.0],
                         rcut2a=[4.8], createvelocity=[300, 4928459, 0, 0, 2, 1, 0], **kw)
    appfile.write_app(f, prefix + "app.bin")
    appfile.write_lattice(prefix + "lattice.bin", "bcc", 3.316)
    appfile.write_region(prefix + "region.bin", [4, 4, 4])
    return appfile.App(prefix)


def test_ta_app_first_step_vs_oracle(oracle, tmp_path):
    from mdp_amd.md import velocity_create
    app = _ta(tmp_path)
    x, t, box = app.atoms()
    md = app.engine()
    print(f"Ta application: {len(x)} atoms", flush=True)
    md.compute()
    st = md.state()
    assert np.array_equal(st["x"], np.mod(x, box))
    v0 = velocity_create(x, t, [0.0, 180.88], 300.0, 4928459, 0, 2)
    assert np.array_equal(st["v"], v0)
    s = snapcpp.initsna(oracle, [1, 6, 4.67637, 0.99363, 0, 0, 1, 0, 0, 0, 0], [0.5], [1.0], TA_COEFF)
    B = np.diag(box)
    p = snapcpp.pairs_from_box(x, t, B[0], B[1], B[2], [1, 1, 1], s)
    e, f, w = snapcpp.snappotential_pairs(oracle, s, p)
    ez, fz, wz = zbl_oracle.zbl_system(oracle, x, box)
    assert relerr(st["eatom"], e + ez) < 1e-10
    assert relerr(st["f"], f + fz) < 1e-10 or np.max(np.abs(st["f"])) < 1e-9     # perfect lattice: F ~ 0
    assert relerr(st["vatom"], w.reshape(-1, 6) + wz) < 1e-10


def test_ta_app_run(tmp_path):
    app = _ta(tmp_path)
    md, rows = app.run()
    assert [r["step"] for r in rows] == list(range(0, 101, 10))
    etot = np.array([r["pe"] + r["ke"] for r in rows])
    assert np.max(np.abs(etot - etot[0])) < 1e-4          # eV/atom over 50 fs
    assert abs(rows[0]["temp"] - 300.0) < 1e-9             # createvelocity's target on the perfect lattice
    assert 0.0 < rows[-1]["temp"] < 300.0                   # kinetic energy flows into the potential
    assert md.stats()["steps"] == 100 and md.stats()["overflows"] == 0


def test_ta_app_nvt(tmp_path):
    app = _ta(tmp_path, ensemble="nvt", nvtparam=[600.0, 600.0, 0.05])
    md, rows = app.run(steps=400)
    # from the perfect lattice at 300 K NVE can only cool (the lattice is the potential's minimum, so the
    # kinetic energy never exceeds its start); the thermostat heats it past 300 K toward its 600 K target
    temps = [r["temp"] for r in rows]
    assert np.mean(temps[-10:]) > 330.0 and temps[-1] > temps[-10]
