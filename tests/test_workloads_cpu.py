"""Host-side pieces of the workloads module that need no GPU: the seeded
256-channel draw (sizes, node count, relaxation from the Reynolds number) and
the rows per block the launch autotuner tries."""

import math

from lattice_boltzmann_parallel_solver_amd import workloads as W


def test_channel_draw_is_seeded_and_sized():
    a = W.channel_jobs(256, seed=0)
    b = W.channel_jobs(256, seed=0)
    assert [(j.lx, j.ly, j.plate, j.reynolds) for j in a] == [(j.lx, j.ly, j.plate, j.reynolds) for j in b]
    assert len(a) == 256
    assert sum(j.nodes for j in a) == 66110492
    for j in a:
        assert 200 <= j.lx <= 1200 and j.ly == round(3 * j.lx / 7)
        assert j.nodes == j.lx * j.ly
        nu = 0.1 * j.plate / j.reynolds
        assert math.isclose(j.omega, 1.0 / (3.0 * nu + 0.5), rel_tol=1e-12)
        assert 0 < j.omega < 2


def test_sweep_row_candidates():
    # 8192 rows on 256 CUs: one block per CU, and slightly fewer blocks
    assert W.sweep_row_candidates(8192, 256) == [0, 1, 2, 4, 32, 34, 36]
    # small lattices: short blocks only (rows above 4 from the CU count where they exist)
    assert W.sweep_row_candidates(96, 256) == [0, 1, 2, 4, 5]
    # no device information: the fixed candidates
    assert W.sweep_row_candidates(8192, 0) == [0, 1, 2, 4]
