"""Nose-Hoover NVT of the GPU MD engine (mdp_md_set_nvt) against the upstream integrator itself:
cpuNVTInitialIntegrate / cpuNVTFinalIntegrate of the upstream C++ core (compiled from source into
oracle/_ref/libcpucore_r2.so), driven step by step with the oracle's forces.  Positions, velocities and
the chain variables agree within 1e-9 after ten steps, for a constant target and for a ramp with a
three-link chain and two sub-steps.  A longer run conserves the extended (Nose-Hoover) energy while the
thermostat changes the kinetic energy.  The thermo output (per-atom PE and KE, temperature, pressure)
matches the upstream formulas on the oracle's per-atom energies and virials.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, relerr
from mdp_amd import lattice, snapcpp

pytestmark = pytest.mark.gpu

R2LIB = os.path.join(ROOT, "oracle", "_ref", "libcpucore_r2.so")
RCUTFAC, RFAC0 = 4.67637, 0.99363
MASS = 180.95
FTM2V, MVV2E, BOLTZ = 1.0 / 1.0364269e-4, 1.0364269e-4, 8.617343e-5


def _upstream(lib):
    P, I = ctypes.c_void_p, ctypes.c_int
    out = []
    for name in ("r2_cpuNVTInitialIntegrate", "r2_cpuNVTFinalIntegrate"):
        f = getattr(lib, name)
        f.argtypes = [P] * 12 + [I] * 6
        f.restype = None
        out.append(f)
    return out


@pytest.mark.parametrize("t0,t1,mtchain,nc,period", [(600.0, 600.0, 3, 1, 0.05), (300.0, 900.0, 3, 2, 0.02)])
def test_nvt_vs_upstream_integrator(oracle, t0, t1, mtchain, nc, period):
    if not os.path.exists(R2LIB):
        pytest.skip("upstream C++ core checker not built (oracle/_ref)")
    from mdp_amd.md import SnapMD
    init, final = _upstream(ctypes.CDLL(R2LIB))
    J, dt, steps, endstep = 4, 0.0005, 10, 40
    x, box = lattice.crystal("bcc", 4, 3.316, 0.05, 21)
    n = len(x)
    v = lattice.maxwell_boltzmann(n, MASS, 300.0, seed=21)
    s0 = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0])
    coeff = lattice.snap_coefficients(s0.ncoeff, seed=21)
    md = SnapMD(x, box, np.ones(n, np.int32), J, coeff, [0.5], [1.0], [MASS], v=v, dt=dt, rebuild_every=1)
    md.set_nvt(t0, t1, period, mtchain=mtchain, nc_tchain=nc, beginstep=0, endstep=endstep, currentstep=0)
    md.compute()
    md.run(steps)
    g = md.state()
    st = md.nvt_state()
    print(f"NVT {t0}->{t1} K, chain {mtchain}x{nc}: T = {st['t_current']:.2f} K (target {st['t_target']:.2f})",
          flush=True)

    s = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0], coeff)
    B = np.diag(box)
    t = np.ones(n, np.int32)

    def force(xx):
        p = snapcpp.pairs_from_box(np.mod(xx, box), t, B[0], B[1], B[2], [1, 1, 1], s)
        return np.ascontiguousarray(snapcpp.snappotential_pairs(oracle, s, p)[1])

    xh, vh = np.ascontiguousarray(x.copy()), np.ascontiguousarray(v.copy())
    mass = np.array([0.0, MASS])
    tdof = 3.0 * (n - 1)
    dtarray = np.array([dt, 0.5 * dt * FTM2V, dt, 0.0, float(endstep), 0.0])
    tarray = np.array([t0, t1, 1.0 / period, 0.0, 0.0, tdof, BOLTZ, 0.0, MVV2E])
    eta_mass, eta, eta_dot, eta_dotdot = (np.zeros(10) for _ in range(4))
    ilist = np.arange(n, dtype=np.int32)
    types = np.ones(n, np.int32)
    f = force(xh)
    args = lambda: [a.ctypes.data for a in (xh, vh, f, mass, dtarray, tarray, eta_mass, eta, eta_dot, eta_dotdot,
                                            types, ilist)] + [1, 0, mtchain, nc, 3, n]
    for step in range(1, steps + 1):
        dtarray[5] = step
        init(*args())
        f = force(xh)
        final(*args())
    d = np.mod(g["x"] - xh + 0.5 * box, box) - 0.5 * box
    assert np.max(np.abs(d)) < 1e-9
    assert relerr(g["v"], vh) < 1e-9
    for k, ref in (("eta", eta), ("eta_dot", eta_dot), ("eta_dotdot", eta_dotdot), ("eta_mass", eta_mass)):
        assert relerr(st[k][:mtchain], ref[:mtchain]) < 1e-9, k
    assert abs(st["t_current"] - tarray[4]) < 1e-9 * tarray[4]
    assert abs(st["t_target"] - tarray[3]) < 1e-12 * tarray[3]


def test_nvt_extended_energy_conserved(oracle):
    """300 NVT steps towards 600 K from 300 K on a weak potential: the extended energy PE + KE + the
    chain energy (cpuComputeNVTEnergy, cpuComputes.cpp:261-272) stays constant while the kinetic energy
    changes by tens of percent, and the total momentum stays zero (all velocities scale alike)."""
    from mdp_amd.md import SnapMD
    J, mt = 4, 3
    x, box = lattice.crystal("bcc", 5, 3.316, 0.05, 5)
    n = len(x)
    v = lattice.maxwell_boltzmann(n, MASS, 300.0, seed=5)
    s0 = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0])
    coeff = lattice.snap_coefficients(s0.ncoeff, seed=5, sigma=0.001)
    md = SnapMD(x, box, np.ones(n, np.int32), J, coeff, [0.5], [1.0], [MASS], v=v, dt=0.0005)
    md.set_nvt(600.0, 600.0, 0.02, mtchain=mt)
    md.compute()
    ep0, ek0 = md.energy()

    def extended():
        ep, ek = md.energy()
        st = md.nvt_state()
        kt = BOLTZ * st["t_target"]
        e_nh = 3.0 * (n - 1) * kt * st["eta"][0] + 0.5 * st["eta_mass"][0] * st["eta_dot"][0] ** 2
        for k in range(1, mt):
            e_nh += kt * st["eta"][k] + 0.5 * st["eta_mass"][k] * st["eta_dot"][k] ** 2
        return ep + ek + e_nh, ek

    md.run(1)
    e1, _ = extended()
    drift, ek_min, ek_max = 0.0, ek0, ek0
    for _ in range(30):
        md.run(10)
        e, ek = extended()
        drift = max(drift, abs(e - e1))
        ek_min, ek_max = min(ek_min, ek), max(ek_max, ek)
    g = md.state()
    print(f"KE {ek0:.3f} eV -> range [{ek_min:.3f}, {ek_max:.3f}]; extended-energy drift {drift:.2e} eV", flush=True)
    assert ek_max > 1.3 * ek0             # the thermostat heats the system towards 600 K
    assert drift < 1e-3 * ek0             # ... while the extended energy is conserved
    assert np.max(np.abs(g["v"].sum(axis=0))) < 1e-9 * np.max(np.abs(g["v"])) * n


def test_thermo_output_vs_oracle(oracle):
    """ThermoOutput (Calculation.cpp:436-465) of the engine against the same formulas on the oracle's
    per-atom energies and virials at the same positions and velocities."""
    from mdp_amd.md import SnapMD
    J = 6
    x, box = lattice.crystal("bcc", 4, 3.316, 0.08, 9)
    n = len(x)
    v = lattice.maxwell_boltzmann(n, MASS, 500.0, seed=9)
    s0 = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0])
    coeff = lattice.snap_coefficients(s0.ncoeff, seed=9)
    md = SnapMD(x, box, np.ones(n, np.int32), J, coeff, [0.5], [1.0], [MASS], v=v)
    md.compute()
    th = md.thermo()
    s = snapcpp.initsna(oracle, [1, J, RCUTFAC, RFAC0, 0, 1, 1, 0, 0, 0, 0], [0.5], [1.0], coeff)
    B = np.diag(box)
    p = snapcpp.pairs_from_box(x, np.ones(n, np.int32), B[0], B[1], B[2], [1, 1, 1], s)
    e, _, w = snapcpp.snappotential_pairs(oracle, s, p)
    tdof = 3.0 * (n - 1)
    temp = MVV2E / (tdof * BOLTZ) * (MASS * (v * v).sum())
    press = (tdof * BOLTZ * temp + w[:, 0].sum() + w[:, 1].sum() + w[:, 2].sum()) / 3.0 / np.prod(box) * 1.6021765e6
    print(f"thermo: {th}", flush=True)
    assert abs(th["pe"] - e.sum() / n) < 1e-10 * abs(e.sum() / n)
    assert abs(th["temp"] - temp) < 1e-10 * temp
    assert abs(th["ke"] - 0.5 * tdof * BOLTZ * temp / n) < 1e-10 * th["ke"]
    assert abs(th["press"] - press) < 1e-9 * max(abs(press), 1.0)
