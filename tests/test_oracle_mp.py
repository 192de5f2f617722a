"""The multi-process oracle (the reference's MPI block decomposition restated
with processes and shared memory) equals the single-process oracle bit for
bit, and its process grid follows get_xy_size (parallelization_utils.py:55-92)."""

import numpy as np
import pytest

from oracle import lbm_oracle as O
from oracle import lbm_oracle_mp as MP


@pytest.mark.parametrize("p,grid", [(1, (1, 1)), (2, (1, 2)), (4, (2, 2)), (6, (2, 3)), (8, (2, 4)),
                                    (12, (3, 4)), (16, (4, 4)), (9, (3, 3)), (10, (2, 5))])
def test_process_grid(p, grid):
    assert MP.get_xy_size(p) == grid


@pytest.mark.parametrize("p", [3, 5, 7, 11, 13])
def test_prime_process_counts_rejected(p):
    with pytest.raises(ValueError):
        MP.get_xy_size(p)


def test_blocks_tile_the_lattice_with_remainders_last():
    lx, ly, grid = 37, 23, (3, 4)
    seen = np.zeros((lx, ly), dtype=int)
    for r in range(12):
        x0, nx, y0, ny = MP.block_of(r, lx, ly, *grid)
        seen[x0:x0 + nx, y0:y0 + ny] += 1
        assert nx == (37 // 3 if r // 4 < 2 else 37 - 2 * (37 // 3))
        assert ny == (23 // 4 if r % 4 < 3 else 23 - 3 * (23 // 4))
    assert (seen == 1).all()


def _serial(f, rho, u, om, steps, bc):
    for _ in range(steps):
        f, rho, u = O.step(f, rho, u, om, bc)
    return f, rho, u


@pytest.mark.parametrize("p", [2, 4, 6])
def test_von_karman_channel_matches_serial_oracle(p):
    lx, ly, d = 41, 26, 8
    rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
    f = O.equilibrium(rho, u)
    bc = O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1)
    om = 1 / (3 * 0.1 * d / 100 + 0.5)
    ref = _serial(f, rho, u, om, 7, bc)
    got = MP.run(f, rho, u, om, 5, bc, processes=p, warmup=2)
    for a, b in zip(got[:3], ref):
        assert np.array_equal(a, b)
    assert got[4] == MP.get_xy_size(p)


def test_couette_and_periodic_match_serial_oracle():
    lx, ly = 20, 18
    rho, u = O.density_1_velocity_0_initial((lx, ly))
    f = O.equilibrium(rho, u)
    bc = O.couette_flow_boundary_conditions(lx, ly, 0.05, 1.0)
    for boundary in (bc, None):
        ref = _serial(f, rho, u, 1.2, 6, boundary)
        got = MP.run(f, rho, u, 1.2, 6, boundary, processes=4)
        for a, b in zip(got[:3], ref):
            assert np.array_equal(a, b)


def test_pressure_fix_is_not_decomposed():
    lx, ly = 16, 12
    rho, u = O.density_1_velocity_0_initial((lx, ly))
    bc = O.poiseuille_flow_boundary_conditions(lx, ly, 1.001 / 3, 0.999 / 3)
    with pytest.raises(RuntimeError):
        MP.run(O.equilibrium(rho, u), rho, u, 1.0, 2, bc, processes=2)
