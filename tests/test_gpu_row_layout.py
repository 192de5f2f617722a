"""Long two-step runs in the range-interleaved row order equal the plain order.

A run of 64 or more steps with two steps per launch copies the state into an
order that stores offset t of every row range of the sweep next to each other,
sweeps there and copies it back.  States, outlet buffers and probe samples
must be bit for bit those of the same run in the plain row order, for row
counts that divide lx and that leave a shorter last range, one and two
columns per thread, column tiles, every storage and boundary kind the sweep
takes, and runs that start pending or end with an odd step.  FP64 is also
checked against the oracle.
"""

import numpy as np
import pytest

from oracle import lbm_oracle as O
from oracle import lbm_oracle_native as N

pytestmark = pytest.mark.gpu


def _lattice(lx, ly, storage, kind, interleave, rows, seed, tile=None, probe=None, steps=0):
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd import boundary_utils as BU

    bc = {"periodic": None,
          "channel": BU.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, max(2, ly // 4), walls=True, y_offset=1),
          "couette": BU.couette_flow_boundary_conditions(lx, ly, 0.05, 1.0)}[kind]
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.01 * rng.standard_normal((lx, ly))
    u = 0.05 * rng.standard_normal((lx, ly, 2))
    f = O.equilibrium(rho, u) + 1e-3 * rng.standard_normal((lx, ly, 9))
    lat = Lattice(lx, ly, boundary=bc, storage=storage)
    lat.set_state(f, rho, u)
    if tile is not None:
        lat.set_sweep_tile(*tile)
    lat.set_temporal(2, rows)
    lat.set_row_layout(interleave)
    if probe is not None:
        lat.set_probe(probe[0], probe[1], steps)
    return lat, (f, rho, u, bc)


CASES = [
    # lx, ly, storage, kind, rows, tile: the order is used for FP32 rows of two
    # columns per thread and FP64 column tiles with 16 or more rows per block
    (300, 1100, "f64", "periodic", 16, None),   # FP64 tiles; 19 ranges, the last one of 12 rows
    (257, 1800, "f32", "channel", 20, None),    # 13 ranges, the last one of 17 rows
    (259, 1100, "f32", "channel", 37, None),    # 7 ranges of 37 rows: lx = 37 * 7 exactly
    (262, 1040, "f32", "channel", 29, None),    # 10 ranges, the last one of 1 row
    (130, 2048, "f32", "couette", 16, None),
    (96, 1800, "f64", "channel", 16, None),     # FP64 tiles, lx = 16 * 6
    (120, 1500, "f32", "channel", 17, (512, 2)),  # FP32 tiles of two columns per thread
    (41, 2048, "f32", "periodic", 40, None),    # a range of 40 rows and one of 1
    (4100, 1100, "f32", "couette", 0, None),    # automatic rows (17 on 256 CUs)
    (200, 1200, "f16", "channel", 16, None),    # FP16 keeps the plain order
    (300, 700, "f32", "channel", 20, None),     # one column per thread keeps the plain order
]


@pytest.mark.parametrize("lx,ly,storage,kind,rows,tile", CASES)
def test_interleaved_runs_equal_plain_order(lx, ly, storage, kind, rows, tile):
    print(f"case: {lx}x{ly} {storage} {kind} rows {rows} tile {tile}", flush=True)
    steps = 1 + 64 + 3 + 71  # pending start, interleaved run, short plain run, interleaved run with an odd step
    probe = (lx // 3, ly // 2)
    runs = []
    for interleave in (True, False):
        lat, _ = _lattice(lx, ly, storage, kind, interleave, rows, seed=5, tile=tile, probe=probe, steps=steps)
        for k in (1, 64, 3, 71):
            lat.run(1.3, k)
        runs.append((lat.state(), lat.probe()))
        lat.close()
    (sa, pa), (sb, pb) = runs
    for a, b in zip(sa, sb):
        assert np.array_equal(a, b)
    assert np.array_equal(pa, pb)
    if storage == "f64":
        lat, (f, rho, u, _) = _lattice(lx, ly, storage, kind, True, rows, seed=5)
        lat.close()
        obc = {"periodic": None,
               "channel": O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, max(2, ly // 4), walls=True, y_offset=1),
               "couette": O.couette_flow_boundary_conditions(lx, ly, 0.05, 1.0)}[kind]
        fo, ro, uo, to = N.run(f, rho, u, 1.3, steps, obc, probe=probe)
        for a, w in zip(sa, (fo, ro, uo)):
            assert np.array_equal(a, w)
        assert np.array_equal(pa, to)


def test_row_layout_argument_checked():
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd._lib import LbmError, check, load

    lat = Lattice(16, 16)
    with pytest.raises(LbmError, match="interleave"):
        check(load().lbm_set_row_layout(lat._h, 2))
    lat.close()


def test_batch_groups_of_wide_rows_in_interleaved_order_equal_single_runs():
    """A batch whose FP32 lattices of more than 1024 columns (two columns per
    thread) sweep long runs in the interleaved row order, beside a narrow
    lattice in the plain order: every lattice equals its own plain-order run
    of single steps, over runs of 64 steps and more (pending start, odd
    steps) and a short one."""
    from lattice_boltzmann_parallel_solver_amd import Batch

    shapes = [(300, 1100, "channel"), (257, 2048, "couette"), (100, 1500, "channel"), (90, 300, "channel")]
    chunks = (1, 64, 3, 71)
    lats, refs = [], []
    for k, (lx, ly, kind) in enumerate(shapes):
        a, _ = _lattice(lx, ly, "f32", kind, True, 0, seed=11 + k, probe=(lx // 3, ly // 2), steps=sum(chunks))
        r, _ = _lattice(lx, ly, "f32", kind, False, 0, seed=11 + k, probe=(lx // 3, ly // 2), steps=sum(chunks))
        r.set_temporal(1, 0)
        lats.append(a)
        refs.append(r)
    b = Batch(lats, [1.3] * len(lats), temporal=2, rows_per_block=16)
    for chunk in chunks:
        b.run(chunk)
        for r in refs:
            r.run(1.3, chunk)
    b.close()
    for a, r, (lx, ly, _) in zip(lats, refs, shapes):
        for x, y in zip(a.state(), r.state()):
            assert np.array_equal(x, y), (lx, ly)
        assert np.array_equal(a.probe(), r.probe()), (lx, ly)
        a.close()
        r.close()
