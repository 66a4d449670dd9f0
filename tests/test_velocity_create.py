"""mdp_velocity_create (host code of libmdpsnap, no GPU) against the upstream app's own sequence
(readdata.cpp:609-640: cpuVelocityCreate in loop mode "geom", cpuComputeMass, cpuComputeVCM,
cpuVelocityZeroMomentum, cpuComputeTempScalar, cpuArrayMultiplyScalar) compiled from source into
oracle/_ref/libcpucore_r2.so: the velocities agree bit for bit, for uniform and normal draws, with the
seed and lattice of the upstream Ta example (snapTA06A.jl:77).  Properties (zero momentum, exact
temperature, dependence on positions only) are checked without the compiled upstream too."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT
from mdp_amd import lattice

R2LIB = os.path.join(ROOT, "oracle", "_ref", "libcpucore_r2.so")
MASS = 180.88   # Ta, the upstream example's mass
MVV2E, BOLTZ = 1.0364269e-4, 8.617343e-5


def _system(seed=3, two_types=False):
    x, _ = lattice.crystal("bcc", 6, 3.316, 0.05, seed)
    t = np.ones(len(x), np.int32)
    if two_types:
        t[::3] = 2
    return np.ascontiguousarray(x), t, np.array([0.0, MASS, 92.9])


@pytest.mark.parametrize("dist,two_types", [(0, False), (1, False), (0, True)])
def test_bit_exact_with_upstream_sequence(dist, two_types):
    if not os.path.exists(R2LIB):
        pytest.skip("upstream C++ core checker not built (oracle/_ref)")
    from mdp_amd.md import velocity_create
    lib = ctypes.CDLL(R2LIB)
    f = lib.r2_velocity_create
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    f.argtypes = [P, P, P, P, I, D, I, I, I, D, D]
    f.restype = None
    x, t, m = _system(5, two_types)
    ref = np.zeros_like(x)
    f(x.ctypes.data, ref.ctypes.data, m.ctypes.data, t.ctypes.data, len(x), 300.0, 4928459, dist, 2, MVV2E, BOLTZ)
    got = velocity_create(x, t, m, 300.0, 4928459, dist_flag=dist)
    assert np.array_equal(got, ref)


def test_properties():
    from mdp_amd.md import velocity_create
    x, t, m = _system(7, True)
    v = velocity_create(x, t, m, 450.0, 12345)
    mm = m[t][:, None]
    assert np.max(np.abs((mm * v).sum(axis=0))) < 1e-10 * np.max(np.abs(mm * v))
    temp = (mm * v * v).sum() * MVV2E / (3 * (len(x) - 1) * BOLTZ)
    assert abs(temp - 450.0) < 1e-9
    # each atom's draw depends on the seed and its own coordinates only
    perm = np.random.default_rng(1).permutation(len(x))
    vp = velocity_create(x[perm], t[perm], m, 450.0, 12345)
    assert np.allclose(vp, v[perm], rtol=0, atol=1e-12 * np.max(np.abs(v)))
    assert not np.allclose(velocity_create(x, t, m, 450.0, 12346), v)
    with pytest.raises(RuntimeError, match="geom"):
        velocity_create(x, t, m, 450.0, 12345, loop_flag=0)
