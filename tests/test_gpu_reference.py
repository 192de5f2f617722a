"""The reference's own von Karman results reproduced on the GPU (FP64).

420x180 channel, plate d=40 at x=105, inlet u=0.1, nu=0.04 (Re=100), probe at
(315, 90) (milestoneQuickFunctionCalls.py:293-341, experiments.py:650-720).
The reference committed the probe |u| traces of a 100k-step serial run and of
a 200k-step MPI run, and the serial run's final density and velocity fields;
tests/golden/ holds their SHA-256 digests and subsamples.  Matching a digest
after shedding has set in (about step 40k) requires every step to agree bit
for bit, since the flow is chaotic.
"""

import hashlib
import os

import numpy as np
import pytest

from oracle import lbm_oracle_native as N

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
LX, LY, D = 420, 180, 40
PROBE = (315, 90)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "von_karman_420x180_reference.npz"))


def new_lattice():
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import von_karman_boundary_conditions
    from lattice_boltzmann_parallel_solver_amd.initial_values import density_1_velocity_x_u0_velocity_y_0_initial

    rho, u = density_1_velocity_x_u0_velocity_y_0_initial((LX, LY), 0.1)
    lat = Lattice(LX, LY, boundary=von_karman_boundary_conditions(LX, LY, 1.0, 0.1, D))
    lat.set_equilibrium(rho, u)
    return lat


OMEGA = float(np.reciprocal(3 * 0.04 + 0.5))


def test_probe_fixture_exact():
    """tests/test_parallelization_von_karman.py:56-57: exact equality."""
    fix = np.load(os.path.join(GOLDEN, "vel_at_p.npy"))
    lat = new_lattice()
    lat.set_probe(*PROBE, 16)
    lat.run(OMEGA, 11)
    tr = lat.probe()
    speed = np.concatenate([[0.1], [np.linalg.norm(r[1:]) for r in tr]])
    assert np.array_equal(speed, fix)


def test_serial_run_100k_steps_bitwise(golden):
    """The reference's serial 100k-step run: final density, final velocity and
    the full |u| probe trace, compared by SHA-256."""
    steps = int(golden["serial_steps"])
    lat = new_lattice()
    lat.set_probe(*PROBE, steps)
    lat.run(OMEGA, steps)
    _, rho, u = lat.state()
    tr = lat.probe()
    trace = np.concatenate([[0.1], N.norm_plain(tr[:, 1], tr[:, 2])])
    head = golden["serial_trace_head"]
    first_bad = np.flatnonzero(trace[: len(head)] != head)
    assert first_bad.size == 0, f"trace departs from the reference at step {first_bad[0]}"
    np.testing.assert_array_equal(trace[::100], golden["serial_trace_every100"])
    assert sha(trace) == str(golden["serial_trace_sha"])
    pts = golden["serial_points"]
    np.testing.assert_array_equal(rho[pts[:, 0], pts[:, 1]], golden["serial_density_at"])
    np.testing.assert_array_equal(u[pts[:, 0], pts[:, 1]], golden["serial_velocity_at"])
    assert sha(rho) == str(golden["serial_density_sha"])
    assert sha(u) == str(golden["serial_velocity_sha"])


def test_mpi_run_200k_steps_trace_bitwise(golden):
    """The reference's 200k-step MPI run (reynold_strouhal/vel_at_p_100.npy),
    whose |u| went through numpy's FMA-based 1-D norm."""
    steps = int(golden["parallel_steps"])
    lat = new_lattice()
    lat.set_probe(*PROBE, steps)
    lat.run(OMEGA, steps)
    tr = lat.probe()
    trace = np.concatenate([[0.1], N.norm_fma(tr[:, 1], tr[:, 2])])
    np.testing.assert_array_equal(trace[::100], golden["parallel_trace_every100"])
    assert sha(trace) == str(golden["parallel_trace_sha"])
