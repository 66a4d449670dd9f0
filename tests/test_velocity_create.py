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


def _draw_args(natoms, nlocal, seed, loop_flag, map_kind):
    x, t, m = _system(11, True)
    x, t = x[:nlocal], t[:nlocal]
    if map_kind == "tags":        # global ID k -> local index k - 1
        mp = np.concatenate([[-1], np.arange(natoms)]).astype(np.int32)
    else:                         # the app's ilist padded by one entry that names no local atom
        mp = np.concatenate([np.arange(nlocal), [-1] * (natoms + 1 - nlocal)]).astype(np.int32)
    return np.ascontiguousarray(x), t, m, mp


@pytest.mark.parametrize("loop_flag,dist,sum_flag,map_kind,rank", [
    (0, 0, 0, "tags", 0), (0, 1, 0, "tags", 0), (0, 1, 1, "ilist", 0),
    (1, 0, 0, "tags", 0), (1, 1, 0, "tags", 3), (1, 1, 1, "tags", 1),
    (2, 0, 0, "tags", 0), (2, 1, 1, "tags", 0)])
def test_every_loop_mode_bit_exact_with_upstream(loop_flag, dist, sum_flag, map_kind, rank):
    """cpuVelocityCreate's three loop modes (cpuVelocityIntegration.cpp:54-135) with a caller-held
    generator state, against the compiled upstream function: velocities and the final state agree bit for
    bit.  "all" runs with the LAMMPS tag map and with the app's index list."""
    if not os.path.exists(R2LIB):
        pytest.skip("upstream C++ core checker not built (oracle/_ref)")
    import mdp_amd
    from mdp_amd import _abi  # noqa: F401
    lib = ctypes.CDLL(R2LIB)
    ours = mdp_amd.library().cdll.mdp_velocity_create_loop
    P, I = ctypes.c_void_p, ctypes.c_int
    args_t = [P] * 8 + [I] * 8
    lib.r2_cpuVelocityCreate.argtypes = args_t
    lib.r2_cpuVelocityCreate.restype = None
    ours.argtypes = args_t
    ours.restype = I
    natoms = 432
    x, t, m, mp = _draw_args(natoms, natoms, 0, loop_flag, map_kind)
    v0 = np.random.default_rng(2).normal(size=x.shape)
    res = []
    for f in (lib.r2_cpuVelocityCreate, ours):
        v = v0.copy()
        second, seed, save = np.array([0.25]), np.array([987654], np.int32), np.array([0], np.int32)
        rc = f(x.ctypes.data, v.ctypes.data, m.ctypes.data, second.ctypes.data, seed.ctypes.data, save.ctypes.data,
               mp.ctypes.data, t.ctypes.data, 4928459, sum_flag, dist, loop_flag, 3, rank, natoms, natoms)
        assert rc in (None, 0)
        res.append((v, second.copy(), seed.copy(), save.copy()))
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)
    if map_kind == "ilist" and loop_flag == 0:
        assert np.array_equal(res[1][0][0], v0[0])   # atom 0 takes no draw through the app's index list


def test_loop_modes_python_wrapper_and_state():
    """mdp_amd.md.velocity_draws: "all" over the whole system equals "local" on rank 0 only when the local
    stream's 100 discarded draws are taken out, the generator state advances by three draws per atom
    (uniform), and a bad loop mode is an error."""
    from mdp_amd.md import RandomState, velocity_draws
    x, t, m = _system(13)
    n = len(x)
    st_all = RandomState(12345)
    v_all = velocity_draws(x, np.zeros_like(x), t, m, st_all, loop_flag=0)
    st_ref = RandomState(12345)
    # "local" discards 100 draws first: start it 100 draws earlier by hand
    st_loc = RandomState(12345)
    velocity_draws(x[:1], np.zeros((1, 3)), t[:1], m, st_loc, loop_flag=1)   # 100 + 3 draws
    assert st_loc.seed[0] != st_ref.seed[0]
    assert np.all(np.abs(v_all * np.sqrt(m[t][:, None])) <= 0.5)
    # three uniform draws per atom: running "all" on 2n atoms continues the same stream
    st2 = RandomState(12345)
    x2 = np.concatenate([x, x + 100.0])
    v2 = velocity_draws(x2, np.zeros_like(x2), np.concatenate([t, t]), m, st2, loop_flag=0)
    assert np.array_equal(v2[:n], v_all)
    with pytest.raises(RuntimeError, match="loop_flag"):
        velocity_draws(x, np.zeros_like(x), t, m, RandomState(1), loop_flag=3)
