"""Pins of the CPU oracle (oracle/) against the reference's own fixtures,
known-answer tests and published results, and against closed forms.

Everything here runs on the CPU.  The oracle is the checker for the GPU
tests, so each of its functions is tied to something the reference produced
or to mathematics that a wrong term, sign or index would violate.
"""

import json
import os

import numpy as np
import pytest

from oracle import lbm_oracle as O
from oracle import lbm_oracle_native as N

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rand_f(shape, seed=0):
    rng = np.random.default_rng(seed)
    return rng.random(shape + (9,)) * 10.0 ** rng.integers(-6, 1, shape + (9,))


# ---------------------------------------------------------------- rounding order
def test_density_matches_numpy_sum():
    f = rand_f((31, 17))
    assert np.array_equal(O.density(f), np.sum(f, axis=-1))


def test_velocity_matches_numpy_fancy_index_sums():
    f = rand_f((23, 19), seed=1)
    rho = np.sum(f, axis=-1)
    rho[3, 4] = 0.0
    jx = np.sum(f[:, :, [1, 5, 8]], axis=2) - np.sum(f[:, :, [3, 6, 7]], axis=2)
    jy = np.sum(f[:, :, [2, 5, 6]], axis=2) - np.sum(f[:, :, [4, 7, 8]], axis=2)
    ux = np.divide(jx, rho, out=np.zeros_like(rho), where=rho != 0)
    uy = np.divide(jy, rho, out=np.zeros_like(rho), where=rho != 0)
    u = O.velocity(rho, f)
    assert np.array_equal(u, np.dstack((ux, uy)))
    assert u[3, 4, 0] == 0.0 and u[3, 4, 1] == 0.0


def test_equilibrium_matches_numpy_matmul_norm_power():
    rng = np.random.default_rng(2)
    rho = rng.random((29, 13)) + 0.5
    u = rng.standard_normal((29, 13, 2)) * 0.1
    cu = u @ O.C.T
    usq = np.power(np.linalg.norm(u, axis=-1), 2)[..., None]
    ref = O.W[None, None, :] * rho[..., None] * (1 + 3 * cu + (9 / 2) * np.power(cu, 2) - (3 / 2) * usq)
    assert np.array_equal(O.equilibrium(rho, u), ref)


def test_equilibrium_moments_invariants():
    # sum_i feq = rho and sum_i c_i feq = rho u (tests/test_navier_stokes_eq.py:12-43)
    rng = np.random.default_rng(3)
    rho = rng.random((40, 30)) + 0.5
    u = rng.standard_normal((40, 30, 2)) * 0.05
    feq = O.equilibrium(rho, u)
    np.testing.assert_allclose(feq.sum(-1), rho, rtol=1e-14)
    np.testing.assert_allclose(feq @ O.C, rho[..., None] * u, rtol=0, atol=1e-15)
    assert O.C.sum() == 0


def test_equilibrium_zero_velocity_is_weights():
    rho = np.array([[1.0, 2.0]])
    feq = O.equilibrium(rho, np.zeros((1, 2, 2)))
    assert np.array_equal(feq[0, 0], O.W)
    assert np.array_equal(feq[0, 1], 2.0 * O.W)


# ---------------------------------------------------------------- streaming
def test_streaming_is_the_periodic_shift():
    f = rand_f((11, 7), seed=4)
    s = O.streaming(f)
    for i in range(9):
        assert np.array_equal(s[..., i], np.roll(f[..., i], (O.C[i, 0], O.C[i, 1]), axis=(0, 1)))


def test_streaming_is_a_permutation():
    f = np.arange(13 * 9 * 9, dtype=np.float64).reshape(13, 9, 9)
    s = O.streaming(f)
    assert np.array_equal(np.sort(s.ravel()), f.ravel())
    # a single population at (2,3) moves one node along c_i
    for i in range(9):
        g = np.zeros((13, 9, 9))
        g[2, 3, i] = 1.0
        h = O.streaming(g)
        assert h[(2 + O.C[i, 0]) % 13, (3 + O.C[i, 1]) % 9, i] == 1.0 and h.sum() == 1.0


# ---------------------------------------------------------------- boundary KATs
# tests/test_boundary_conditions.py:12-121 of the reference
def test_inlet_kat():
    shape = (10, 10)
    vel = np.zeros(shape + (2,))
    vel[..., 0] = 0.1
    feq = O.equilibrium(np.ones(shape), vel)
    f = O.inlet(shape, 1, 0.1)(np.zeros(shape + (9,)))
    assert np.allclose(f[0], feq[0]) and np.allclose(f[1:], 0)


def test_outlet_kat():
    f = O.outlet()(np.ones((10, 10, 9)), np.zeros((10, 10, 9)))
    for i in range(9):
        assert np.allclose(f[-1, :, i], 1.0 if i in (3, 6, 7) else 0.0)


def test_rigid_wall_kat():
    m = np.zeros((10, 10), dtype=bool)
    m[:, -1] = True
    f = O.rigid_wall(m)(np.ones((10, 10, 9)), np.zeros((10, 10, 9)))
    for i in range(9):
        assert np.allclose(f[m, i], 1.0 if i in (4, 7, 8) else 0.0)


def test_moving_wall_kat():
    m = np.zeros((10, 10), dtype=bool)
    m[:, -1] = True
    f = O.moving_wall(m, np.array([2, 0]), 1)(np.ones((10, 10, 9)), np.zeros((10, 10, 9)))
    # f7 = 1 - 2*(1/36)*2/(1/3) = 1 - 1/3, f8 = 1 + 1/3, f4 = 1
    assert np.allclose(f[m, 4], 1.0)
    assert np.allclose(f[m, 7], 1 - 1 / 3)
    assert np.allclose(f[m, 8], 1 + 1 / 3)
    for i in (0, 1, 2, 3, 5, 6):
        assert np.allclose(f[m, i], 0.0)


def test_pressure_periodic_kat():
    shape = (10, 10)
    vel = np.zeros(shape + (2,))
    vel[..., 0] = 0.1
    m = np.zeros(shape, dtype=bool)
    m[0, :] = True
    m[-1, :] = True
    f = O.periodic_with_pressure_variations(m, 1, 0.1)(np.ones(shape + (9,)), np.ones(shape), vel)
    # reference constants, and the same values derived by hand:
    # rho_in = 3: f1 = (3/9)(1.33) + 1 - 1.33/9, f5 = f8 = (3/36)(1.33) + 1 - 1.33/36
    # rho_out = 0.3: f3 = (0.3/9)(0.73) + 1 - 0.73/9, f6 = f7 = (0.3/36)(0.73) + 1 - 0.73/36
    exact = {1: 3 / 9 * 1.33 + 1 - 1.33 / 9, 5: 3 / 36 * 1.33 + 1 - 1.33 / 36,
             3: 0.3 / 9 * 0.73 + 1 - 0.73 / 9, 6: 0.3 / 36 * 0.73 + 1 - 0.73 / 36}
    assert np.allclose(f[0, :, 1], 1.29555) and np.allclose(f[0, :, 1], exact[1], rtol=1e-12)
    for i in (5, 8):
        assert np.allclose(f[0, :, i], 1.073888) and np.allclose(f[0, :, i], exact[5], rtol=1e-12)
    assert np.allclose(f[-1, :, 3], 0.943222) and np.allclose(f[-1, :, 3], exact[3], rtol=1e-12)
    for i in (6, 7):
        assert np.allclose(f[-1, :, i], 0.9858) and np.allclose(f[-1, :, i], exact[6], rtol=1e-12)
    untouched = np.ones(shape + (9,), dtype=bool)
    untouched[0, :, [1, 5, 8]] = False
    untouched[-1, :, [3, 6, 7]] = False
    assert np.all(f[untouched] == 1.0)


def test_rigid_object_clears_corners_in_callers_mask():
    m = np.zeros((12, 10), dtype=bool)
    m[3, 2:8] = True
    O.rigid_object(m)
    assert not m[3, 2] and not m[3, 7] and m[3, 3:7].all()


# ---------------------------------------------------------------- initial values KATs
def test_initial_values_kat():
    # tests/test_initial_values.py:12-63 of the reference
    rho, u = O.sinusoidal_density_x((10, 10), 0.5, 0.3)
    np.testing.assert_allclose(rho[:, 0], [0.5, 0.676, 0.785, 0.785, 0.676, 0.5, 0.324, 0.215, 0.215, 0.324],
                               atol=5e-4)
    assert np.all(rho == rho[:, :1]) and np.all(u == 0)
    rho, u = O.sinusoidal_velocity_x((10, 10), 0.05)
    np.testing.assert_allclose(u[0, :, 0], [0.0, 0.029, 0.048, 0.048, 0.029, 0.0, -0.029, -0.048, -0.048, -0.029],
                               atol=5e-4)
    assert np.all(rho == 1.0) and np.all(u[..., 0] == u[:1, :, 0]) and np.all(u[..., 1] == 0)


# ---------------------------------------------------------------- reference results
def _vk_setup(lx=420, ly=180, d=40):
    rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
    return O.equilibrium(rho, u), rho, u, O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d)


def test_von_karman_probe_fixture_exact():
    """tests/test_parallelization_von_karman.py:53-57: |u| at (315, 90) after
    each of 11 steps equals tests/von_karman_vortex_shedding/vel_at_p.npy."""
    fix = np.load(os.path.join(GOLDEN, "vel_at_p.npy"))
    f, rho, u, bc = _vk_setup()
    omega = float(np.reciprocal(3 * 0.04 + 0.5))
    trace = [np.linalg.norm(u[315, 90])]
    for _ in range(11):
        f, rho, u = O.step(f, rho, u, omega, bc)
        trace.append(np.linalg.norm(u[315, 90]))
    assert np.array_equal(np.array(trace), fix)


def test_c_oracle_equals_numpy_oracle_on_all_composites():
    lx, ly = 40, 24
    om = O.omega_from_viscosity(0.04)
    cases = []
    rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
    cases.append((O.equilibrium(rho, u), rho, u, om, O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, 8)))
    cases.append((O.equilibrium(rho, u), rho, u, om,
                  O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, 8, walls=True, y_offset=1)))
    rho, u = O.density_1_velocity_0_initial((lx, ly))
    cases.append((O.equilibrium(rho, u), rho, u, 1.0, O.couette_flow_boundary_conditions(lx, ly, 0.05, 1.0)))
    cases.append((O.equilibrium(rho, u), rho, u, 1.5,
                  O.poiseuille_flow_boundary_conditions(lx, ly, 1.0015 / 3, 0.9985 / 3)))
    rho, u = O.sinusoidal_velocity_x((lx, ly), 0.05)
    cases.append((O.equilibrium(rho, u), rho, u, 0.7, None))
    for f, rho, u, om_, bc in cases:
        a = O.run(f, rho, u, om_, 40, bc)
        b = N.run(f, rho, u, om_, 40, bc)
        for x, y in zip(a[:3], b[:3]):
            assert np.array_equal(x, y)


def test_couette_published_fit():
    """figures/couette_flow/linregress.csv from experiments.py:304-363."""
    from scipy.stats import linregress

    ref = json.load(open(os.path.join(GOLDEN, "reference_csv.json")))["couette"]
    lx, ly = ref["lattice"]
    rho, u = O.density_1_velocity_0_initial((lx, ly))
    bc = O.couette_flow_boundary_conditions(lx, ly, ref["U"], np.mean(rho))
    _, _, u, _ = N.run(O.equilibrium(rho, u), rho, u, ref["omega"], ref["steps"], bc)
    r = linregress(np.arange(ly), u[int(lx / 2), :, 0])
    assert r.slope == pytest.approx(ref["slope"], rel=1e-12)
    assert r.intercept == pytest.approx(ref["intercept"], rel=1e-12)
    assert r.rvalue == pytest.approx(ref["rvalue"], rel=1e-14)
    # and the analytic line through walls at y = -1/2 and y = ly - 1/2
    assert r.slope == pytest.approx(-ref["U"] / ly, rel=1e-9)


@pytest.mark.slow
def test_poiseuille_published_areas_and_fit():
    """figures/poiseuille_flow/areas.csv and curve_fit.csv from experiments.py:377-492."""
    from scipy.optimize import curve_fit

    ref = json.load(open(os.path.join(GOLDEN, "reference_csv.json")))["poiseuille"]
    lx, ly = ref["lattice"]
    drho = ref["delta_p"] * 3
    bc = O.poiseuille_flow_boundary_conditions(lx, ly, (1 + drho / 2) / 3, (1 - drho / 2) / 3)
    rho, u = O.density_1_velocity_0_initial((lx, ly))
    _, rho, u, _ = N.run(O.equilibrium(rho, u), rho, u, ref["omega"], ref["steps"], bc)
    vx = u[..., 0]
    trapz = getattr(np, "trapezoid", None) or np.trapz
    a_in = trapz(vx[1, :], np.arange(ly))
    a_mid = trapz(vx[lx // 2, :], np.arange(ly))
    assert a_in == pytest.approx(ref["area_inlet"], rel=1e-13)
    assert a_mid == pytest.approx(ref["area_middle"], rel=1e-13)
    popt, _ = curve_fit(lambda y, a, b, c: a * y ** 2 + b * y + c, np.arange(ly), vx[lx // 2, :])
    np.testing.assert_allclose(popt, ref["popt"], rtol=1e-8)


@pytest.mark.slow
def test_von_karman_reference_traces_head_exact():
    """6000 steps of the 420x180 Re=100 channel: the |u| probe equals the
    reference's serial trace (plain norm) and its MPI trace (BLAS-dot norm,
    sqrt(fma(uy, uy, ux*ux))) bit for bit."""
    g = np.load(os.path.join(GOLDEN, "von_karman_420x180_reference.npz"))
    f, rho, u, bc = _vk_setup()
    omega = float(np.reciprocal(3 * 0.04 + 0.5))
    steps = 6000
    _, _, _, tr = N.run(f, rho, u, omega, steps, bc, probe=(315, 90))
    ux, uy = tr[:, 1], tr[:, 2]
    plain = np.concatenate([[0.1], N.norm_plain(ux, uy)])
    assert np.array_equal(plain, g["serial_trace_head"][: steps + 1])
    # the MPI run's norm rounds ux*ux + uy*uy once (a fused multiply-add)
    fused = np.concatenate([[0.1], N.norm_fma(ux, uy)])
    assert np.array_equal(fused, g["parallel_trace_head"][: steps + 1])


def test_norm_fma_is_numpys_1d_norm_here():
    """On this host numpy's 1-D norm of a 2-vector is sqrt(fma(uy, uy, ux*ux))."""
    rng = np.random.default_rng(5)
    v = rng.standard_normal((4000, 2)) * 0.05
    ref = np.array([np.linalg.norm(r) for r in v])
    assert np.array_equal(N.norm_fma(v[:, 0], v[:, 1]), ref)


# ---------------------------------------------------------------- invariants
def test_mass_preservation_periodic():
    """tests/test_mass_preservation.py:22-34 (50x50, omega 0.5, density bump),
    with the bar tightened to 1e-13 relative."""
    rho = np.ones((50, 50)) * 0.5
    rho[24, 24] = 0.6
    u = np.zeros((50, 50, 2))
    f = O.equilibrium(rho, u)
    m0 = rho.sum()
    for _ in range(10):
        f, rho, u, _ = N.run(f, rho, u, 0.5, 100)
        assert abs(rho.sum() - m0) <= 1e-13 * m0


def test_shear_wave_viscosity_closed_form():
    """300x300 shear wave (eps 0.05, omega 1): amplitude decays as
    exp(-nu k^2 t) with nu = (1/omega - 1/2)/3."""
    lx = ly = 300
    rho, u = O.sinusoidal_velocity_x((lx, ly), 0.05)
    f = O.equilibrium(rho, u)
    amps = []
    steps = 3000
    for _ in range(6):
        f, rho, u, _ = N.run(f, rho, u, 1.0, steps // 6)
        amps.append(np.abs(u[..., 0]).max())
    t = np.arange(1, 7) * (steps // 6)
    k = 2 * np.pi / ly
    nu_fit = -np.polyfit(t, np.log(amps), 1)[0] / (k * k)
    assert nu_fit == pytest.approx(O.viscosity_from_omega(1.0), rel=2e-3)
    assert amps[-1] == pytest.approx(O.shear_wave_amplitude_decay(0.05, 1 / 6, k, steps), rel=5e-3)
