"""GPU parity of the full step: lattice_boltzmann_step and the device-resident
Lattice (fused stream+collide kernel) against the oracle, FP64 bit for bit,
for every boundary composite of the reference."""

import numpy as np
import pytest

from oracle import lbm_oracle as O
from oracle import lbm_oracle_native as N

pytestmark = pytest.mark.gpu


def scenarios(lx, ly):
    """(name, oracle composite, package composite factory, f, rho, u, omega)."""
    from lattice_boltzmann_parallel_solver_amd import boundary_utils as BU

    out = []
    rho, u = O.sinusoidal_velocity_x((lx, ly), 0.05)
    out.append(("periodic", None, lambda: None, O.equilibrium(rho, u), rho, u, 0.8))
    rho, u = O.density_1_velocity_0_initial((lx, ly))
    out.append(("couette", O.couette_flow_boundary_conditions(lx, ly, 0.05, 1.0),
                lambda: BU.couette_flow_boundary_conditions(lx, ly, 0.05, 1.0), O.equilibrium(rho, u), rho, u, 1.0))
    out.append(("poiseuille", O.poiseuille_flow_boundary_conditions(lx, ly, 1.0015 / 3, 0.9985 / 3),
                lambda: BU.poiseuille_flow_boundary_conditions(lx, ly, 1.0015 / 3, 0.9985 / 3),
                O.equilibrium(rho, u), rho, u, 1.5))
    rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
    d = max(2, ly // 4)
    om = O.omega_from_viscosity(0.04)
    out.append(("von_karman", O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d),
                lambda: BU.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d), O.equilibrium(rho, u), rho, u, om))
    out.append(("channel", O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1),
                lambda: BU.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1),
                O.equilibrium(rho, u), rho, u, om))
    return out


NAMES = ["periodic", "couette", "poiseuille", "von_karman", "channel"]


def pick(name, lx, ly):
    return next(s for s in scenarios(lx, ly) if s[0] == name)


@pytest.mark.parametrize("name", NAMES)
def test_dropin_step_bit_exact(name):
    from lattice_boltzmann_parallel_solver_amd.lattice_boltzmann_method import lattice_boltzmann_step

    _, bo, mk, f, rho, u, om = pick(name, 37, 23)
    bc = mk()
    fo, ro, uo = f, rho, u
    fg, rg, ug = f, rho, u
    for _ in range(6):
        fo, ro, uo = O.step(fo, ro, uo, om, bo)
        fg, rg, ug = lattice_boltzmann_step(fg, rg, ug, om, bc)
        assert np.array_equal(fg, fo) and np.array_equal(rg, ro) and np.array_equal(ug, uo)


@pytest.mark.parametrize("storage_cfg", [(0, 0, 0), (0, 1, 1), (1, 0, 0), (2, 1, 0), (4, 1, 1), (2, 0, 1)])
@pytest.mark.parametrize("name", NAMES)
def test_lattice_run_bit_exact(name, storage_cfg):
    """n fused steps on the device equal n oracle steps, for every launch shape
    (0 = 16-byte groups of 2 FP64 nodes, else nodes per thread)."""
    from lattice_boltzmann_parallel_solver_amd import Lattice

    lx, ly = 130, 72
    _, bo, mk, f, rho, u, om = pick(name, lx, ly)
    lat = Lattice(lx, ly, boundary=mk(), storage="f64")
    v, xcd, nt = storage_cfg
    lat.set_launch(v, xcd, nt)
    lat.set_state(f, rho, u)
    lat.run(om, 150)
    got = lat.state()
    want = N.run(f, rho, u, om, 150, bo)
    for a, b in zip(got, want[:3]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", [(57, 33), (16, 2), (9, 1), (3, 6)])
def test_ragged_and_thin_lattices_bit_exact(shape):
    """Rows that do not split into 16-byte groups use the per-node kernel."""
    from lattice_boltzmann_parallel_solver_amd import Lattice

    lx, ly = shape
    rho, u = O.sinusoidal_velocity_x((lx, ly), 0.05)
    f = O.equilibrium(rho, u)
    lat = Lattice(lx, ly)
    lat.set_state(f, rho, u)
    lat.run(0.9, 25)
    got = lat.state()
    want = N.run(f, rho, u, 0.9, 25)
    for a, b in zip(got, want[:3]):
        assert np.array_equal(a, b)


def test_run_in_chunks_equals_one_run_and_single_steps():
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd.lattice_boltzmann_method import lattice_boltzmann_step

    lx, ly = 48, 30
    _, bo, mk, f, rho, u, om = pick("channel", lx, ly)
    bc = mk()
    a = Lattice(lx, ly, boundary=bc)
    a.set_state(f, rho, u)
    a.run(om, 40)
    one = a.state()
    b = Lattice(lx, ly, boundary=bc)
    b.set_state(f, rho, u)
    for k in (1, 7, 12, 20):
        b.run(om, k)
        b.state()  # materialising the state must not disturb the run
    chunks = b.state()
    fs, rs, us = f, rho, u
    for _ in range(40):
        fs, rs, us = lattice_boltzmann_step(fs, rs, us, om, bc)
    for x, y, z in zip(one, chunks, (fs, rs, us)):
        assert np.array_equal(x, y) and np.array_equal(x, z)


def test_passed_density_velocity_are_used_for_the_first_collision():
    """The reference collides with the density/velocity it is given, not the
    moments of f (lattice_boltzmann_method.py:213)."""
    from lattice_boltzmann_parallel_solver_amd.lattice_boltzmann_method import lattice_boltzmann_step

    rng = np.random.default_rng(4)
    f = np.abs(rng.standard_normal((20, 16, 9))) * 0.1
    rho = rng.random((20, 16)) + 0.5
    u = rng.standard_normal((20, 16, 2)) * 0.05
    got = lattice_boltzmann_step(f, rho, u, 0.9)
    want = O.step(f, rho, u, 0.9)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_probe_trace_and_moment_sums():
    from lattice_boltzmann_parallel_solver_amd import Lattice

    lx, ly = 90, 40
    _, bo, mk, f, rho, u, om = pick("von_karman", lx, ly)
    lat = Lattice(lx, ly, boundary=mk())
    lat.set_state(f, rho, u)
    lat.set_probe(3 * lx // 4, ly // 2, 1000)
    lat.run(om, 37)
    lat.run(om, 26)
    tr = lat.probe()
    _, ro, uo, want = N.run(f, rho, u, om, 63, bo, probe=(3 * lx // 4, ly // 2))
    assert tr.shape == (63, 3)
    assert np.array_equal(tr, want)
    m, jx, jy = lat.moments_sum()
    assert m == pytest.approx(ro.sum(), rel=1e-14)
    assert jx == pytest.approx((ro * uo[..., 0]).sum(), rel=1e-12)
    # probe ring smaller than the run keeps the newest samples
    lat2 = Lattice(lx, ly, boundary=mk())
    lat2.set_state(f, rho, u)
    lat2.set_probe(10, 5, 16)
    lat2.run(om, 63)
    _, _, _, want2 = N.run(f, rho, u, om, 63, bo, probe=(10, 5))
    assert np.array_equal(lat2.probe(), want2[-16:])


def test_mass_conserved_periodic_fp64():
    """tests/test_mass_preservation.py:22-34 at 1e-13 relative."""
    from lattice_boltzmann_parallel_solver_amd import Lattice

    rho = np.ones((50, 50)) * 0.5
    rho[24, 24] = 0.6
    u = np.zeros((50, 50, 2))
    lat = Lattice(50, 50)
    lat.set_equilibrium(rho, u)
    m0 = rho.sum()
    for _ in range(10):
        lat.run(0.5, 1000)
        m, _, _ = lat.moments_sum()
        assert abs(m - m0) <= 1e-13 * m0


def test_errors_are_raised():
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd._lib import LbmError
    from lattice_boltzmann_parallel_solver_amd.lattice_boltzmann_method import lattice_boltzmann_step

    lat = Lattice(8, 8)
    with pytest.raises(LbmError, match="no state"):
        lat.run(1.0, 1)
    lat.set_equilibrium(np.ones((8, 8)), np.zeros((8, 8, 2)))
    with pytest.raises(LbmError, match="omega"):
        lat.run(2.0, 1)
    with pytest.raises(AssertionError):
        lattice_boltzmann_step(np.ones((8, 8, 9)), np.ones((8, 8)), np.zeros((8, 8, 2)), 0.0)
    with pytest.raises(AssertionError):
        lattice_boltzmann_step(np.ones((8, 8, 9)), np.ones((8, 7)), np.zeros((8, 8, 2)), 1.0)
    with pytest.raises(TypeError):
        lattice_boltzmann_step(np.ones((8, 8, 9)), np.ones((8, 8)), np.zeros((8, 8, 2)), 1.0,
                               boundary=lambda *a: a[1])
    with pytest.raises(NotImplementedError):
        lattice_boltzmann_step(np.ones((8, 8, 9)), np.ones((8, 8)), np.zeros((8, 8, 2)), 1.0,
                               parallel_communication=lambda f: f)
