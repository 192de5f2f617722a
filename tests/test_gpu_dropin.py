"""GPU parity of the drop-in functions against the oracle (FP64: bit for bit).

Each test calls the reference-named function of
lattice_boltzmann_parallel_solver_amd, which runs a HIP kernel of
liblbm_hip.so, on the same seeded inputs as the oracle.
"""

import numpy as np
import pytest

from oracle import lbm_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (1, 7), (5, 1), (17, 13), (64, 48), (300, 300)]


def rand_f(shape, seed=0):
    rng = np.random.default_rng(seed)
    return rng.random(tuple(shape) + (9,)) * 10.0 ** rng.integers(-5, 1, tuple(shape) + (9,))


@pytest.fixture(scope="module")
def M():
    from lattice_boltzmann_parallel_solver_amd import lattice_boltzmann_method as M

    return M


@pytest.mark.parametrize("shape", SHAPES)
def test_streaming_pure_shift_bit_exact(M, shape):
    f = rand_f(shape, seed=shape[0] * 31 + shape[1])
    got = M.streaming(f)
    assert got.dtype == np.float64 and got.shape == f.shape
    assert np.array_equal(got, O.streaming(f))
    # and against numpy's own roll, channel by channel (lattice_boltzmann_method.py:156-157)
    for i in range(9):
        assert np.array_equal(got[..., i], np.roll(f[..., i], tuple(O.C[i]), axis=(0, 1)))


@pytest.mark.parametrize("shape", SHAPES)
def test_equilibrium_bit_exact(M, shape):
    rng = np.random.default_rng(11)
    rho = rng.random(shape) + 0.5
    u = rng.standard_normal(tuple(shape) + (2,)) * 0.1
    assert np.array_equal(M.equilibrium_distr_func(rho, u), O.equilibrium(rho, u))


def test_equilibrium_any_leading_shape(M):
    # the pressure boundary calls it with (ly,) and (ly, 2) (boundary_conditions.py:338)
    rng = np.random.default_rng(12)
    rho = rng.random(37) + 0.5
    u = rng.standard_normal((37, 2)) * 0.1
    assert np.array_equal(M.equilibrium_distr_func(rho, u), O.equilibrium(rho, u))


@pytest.mark.parametrize("shape", SHAPES)
def test_density_and_velocity_bit_exact(M, shape):
    f = rand_f(shape, seed=3)
    rho = M.compute_density(f)
    assert np.array_equal(rho, O.density(f))
    rho_given = rho.copy()
    rho_given.flat[0] = 0.0  # zero density -> zero velocity (lattice_boltzmann_method.py:125-126)
    u = M.compute_velocity_field(rho_given, f)
    assert np.array_equal(u, O.velocity(rho_given, f))
    assert u.flat[0] == 0.0 and u.flat[1] == 0.0


def test_reference_unit_tests(M):
    # tests/test_density_computation.py, test_velocity_computation.py, test_streaming_func.py
    f = np.ones((10, 10, 9)) / 9
    assert abs(np.sum(M.compute_density(f)) - 100) < 1e-12
    u = M.compute_velocity_field(np.ones((10, 10)), f)
    assert u.shape == (10, 10, 2) and abs(u.sum()) < 1e-15
    g = np.zeros((10, 10, 9))
    g[..., 1] = 1.0
    s = M.streaming(g)
    assert s.shape == g.shape and s.sum() == g.sum()


def test_torch_device_tensors_stay_on_device(M):
    import torch

    f = rand_f((33, 21), seed=5)
    t = torch.from_numpy(f).cuda()
    out = M.streaming(t)
    assert isinstance(out, torch.Tensor) and out.is_cuda
    assert np.array_equal(out.cpu().numpy(), O.streaming(f))
    rho = M.compute_density(t)
    assert np.array_equal(rho.cpu().numpy(), O.density(f))
    # float32 tensors run the float32 kernels: compare with a float32-rounded oracle bound
    t32 = t.float()
    rho32 = M.compute_density(t32)
    assert rho32.dtype == torch.float32
    np.testing.assert_allclose(rho32.cpu().numpy(), O.density(f), rtol=2e-6)


def test_collide_entry_point_bit_exact():
    import ctypes

    import torch

    from lattice_boltzmann_parallel_solver_amd import _lib

    rng = np.random.default_rng(9)
    f = rand_f((40, 30), seed=9)
    rho = rng.random((40, 30)) + 0.5
    u = rng.standard_normal((40, 30, 2)) * 0.05
    tf, tr, tu = (torch.from_numpy(a).cuda() for a in (f, rho, u))
    out = torch.empty_like(tf)
    _lib.check(_lib.load().lbm_collide(tf.data_ptr(), tr.data_ptr(), tu.data_ptr(), ctypes.c_double(0.7),
                                       out.data_ptr(), 1200, _lib.DT_F64,
                                       torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), O.collide(f, O.equilibrium(rho, u), 0.7))


# --------------------------------------------------------------- boundary primitives
@pytest.fixture(scope="module")
def B():
    from lattice_boltzmann_parallel_solver_amd import boundary_conditions as B

    return B


def test_boundary_reference_kats(B, M):
    """The reference's tests/test_boundary_conditions.py:12-121, through the GPU primitives."""
    shape = (10, 10)
    vel = np.zeros(shape + (2,))
    vel[..., 0] = 0.1
    feq = M.equilibrium_distr_func(np.ones(shape), vel)
    f = np.zeros(shape + (9,))
    out = B.inlet(shape, 1, 0.1)(f.copy())
    assert np.allclose(out[0], feq[0]) and np.allclose(out[1:], 0)

    out = B.outlet()(np.ones(shape + (9,)), np.zeros(shape + (9,)))
    for i in range(9):
        assert np.allclose(out[-1, :, i], 1.0 if i in (3, 6, 7) else 0.0)

    m = np.zeros(shape, dtype=bool)
    m[:, -1] = True
    out = B.rigid_wall(m)(np.ones(shape + (9,)), np.zeros(shape + (9,)))
    for i in range(9):
        assert np.allclose(out[m, i], 1.0 if i in (4, 7, 8) else 0.0)

    out = B.moving_wall(m, np.array([2, 0]), 1)(np.ones(shape + (9,)), np.zeros(shape + (9,)))
    assert np.allclose(out[m, 4], 1) and np.allclose(out[m, 7], 1 - 1 / 3) and np.allclose(out[m, 8], 1 + 1 / 3)

    p = np.zeros(shape, dtype=bool)
    p[0, :] = p[-1, :] = True
    out = B.periodic_with_pressure_variations(p, 1, 0.1)(np.ones(shape + (9,)), np.ones(shape), vel)
    assert np.allclose(out[0, :, 1], 1.29555) and np.allclose(out[0, :, 5], 1.073888)
    assert np.allclose(out[0, :, 8], 1.073888) and np.allclose(out[-1, :, 3], 0.943222)
    assert np.allclose(out[-1, :, 6], 0.9858) and np.allclose(out[-1, :, 7], 0.9858)


def test_boundary_primitives_bit_exact_on_random_arrays(B):
    lx, ly = 23, 17
    rng = np.random.default_rng(21)
    pre, post, prev = (rand_f((lx, ly), seed=s) for s in (1, 2, 3))
    top = np.zeros((lx, ly), bool)
    top[:, -1] = True
    left = np.zeros((lx, ly), bool)
    left[0, :] = True
    obj = np.zeros((lx, ly), bool)
    obj[6, 4:12] = True
    pairs = [
        (B.rigid_wall(top), O.rigid_wall(top)),
        (B.rigid_wall(left), O.rigid_wall(left)),
        (B.moving_wall(top, np.array([0.05, 0.01]), 1.02), O.moving_wall(top, np.array([0.05, 0.01]), 1.02)),
        (B.rigid_object(obj.copy()), O.rigid_object(obj.copy())),
    ]
    for ours, ref in pairs:
        a = ours(pre, post.copy())
        b = ref(pre, post.copy())
        assert np.array_equal(a, b)
    a = B.inlet((lx, ly), 1.01, 0.08)(post.copy())
    assert np.array_equal(a, O.inlet((lx, ly), 1.01, 0.08)(post.copy()))
    a = B.outlet()(prev, post.copy())
    assert np.array_equal(a, O.outlet()(prev, post.copy()))
    p = np.zeros((lx, ly), bool)
    p[0, :] = p[-1, :] = True
    rho = rng.random((lx, ly)) + 0.5
    u = rng.standard_normal((lx, ly, 2)) * 0.05
    a = B.periodic_with_pressure_variations(p, 0.34, 0.33)(pre.copy(), rho, u)
    b = O.periodic_with_pressure_variations(p, 0.34, 0.33)(pre.copy(), rho, u)
    assert np.array_equal(a, b)


def test_primitives_mutate_in_place_like_the_reference(B):
    top = np.zeros((8, 6), bool)
    top[:, -1] = True
    post = np.zeros((8, 6, 9))
    out = B.rigid_wall(top)(np.ones((8, 6, 9)), post)
    assert out is post and post[0, -1, 4] == 1.0
    obj = np.zeros((8, 6), bool)
    obj[2, 1:5] = True
    B.rigid_object(obj)
    assert not obj[2, 1] and not obj[2, 4] and obj[2, 2:4].all()
