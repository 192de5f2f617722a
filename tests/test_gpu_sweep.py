"""Two steps per launch (the row-sweep kernel) equal single steps bit for bit.

The sweep kernel keeps the intermediate step on chip (registers and LDS) and
must reproduce exactly what two single fused steps write to HBM: states,
outlet buffers and probe samples, in every storage, for lattices whose ly is
or is not a multiple of the wavefront, for one or two columns per thread, for
blocks of any row count (including more rows than the lattice has) and with
every boundary kind the kernel supports, and for rows wider than one block
(column tiles with halo columns).  FP64 is also checked against the oracle
directly.
"""

import numpy as np
import pytest

from oracle import lbm_oracle as O
from oracle import lbm_oracle_native as N

pytestmark = pytest.mark.gpu


def _boundary(kind, lx, ly):
    from lattice_boltzmann_parallel_solver_amd import boundary_utils as BU

    if kind == "periodic":
        return None
    if kind == "channel":
        return BU.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, max(2, ly // 4), walls=True, y_offset=1)
    if kind == "channel_open":
        return BU.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, max(2, ly // 4))
    if kind == "couette":
        return BU.couette_flow_boundary_conditions(lx, ly, 0.05, 1.0)
    raise ValueError(kind)


def _initial(lx, ly, seed):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.01 * rng.standard_normal((lx, ly))
    u = 0.05 * rng.standard_normal((lx, ly, 2))
    return O.equilibrium(rho, u) + 1e-3 * rng.standard_normal((lx, ly, 9)), rho, u


def _run(lx, ly, storage, kind, steps, temporal, rows=0, probe=None, omega=1.3, seed=0, tile=None):
    from lattice_boltzmann_parallel_solver_amd import Lattice

    f, rho, u = _initial(lx, ly, seed)
    lat = Lattice(lx, ly, boundary=_boundary(kind, lx, ly), storage=storage)
    lat.set_state(f, rho, u)
    if tile is not None:
        lat.set_sweep_tile(*tile)
    if temporal:
        lat.set_temporal(2 if temporal is True else int(temporal), rows)
    if probe is not None:
        lat.set_probe(probe[0], probe[1], steps)
    # uneven chunks exercise pending/running phases and odd remainders
    done = 0
    for k in (1, steps // 3, steps):
        k = min(k, steps) - done
        if k > 0:
            lat.run(omega, k)
            done += k
    out = lat.state()
    smp = lat.probe() if probe is not None else None
    lat.close()
    return out, smp


CASES = [
    # (lx, ly, storage, boundary, steps, rows)
    (24, 64, "f32", "periodic", 9, 0),
    (24, 64, "f32", "channel", 10, 5),
    (37, 100, "f32", "channel", 11, 4),        # ly not a multiple of 64
    (3, 17, "f32", "periodic", 8, 0),          # smallest lx, rows > lx
    (5, 130, "f16", "channel_open", 9, 1),     # one row per block
    (40, 1500, "f32", "channel", 7, 16),       # two columns per thread, partial waves
    (33, 2048, "f32", "channel", 6, 0),        # two columns per thread, full
    (33, 2048, "f16", "channel", 6, 8),
    (20, 96, "f64", "channel", 9, 3),
    (31, 1024, "f64", "couette", 8, 0),
    (26, 200, "f32", "couette", 10, 6),
    (26, 200, "f16", "periodic", 10, 100),
    (16, 1030, "f32", "channel", 9, 4),        # a whole wave of lanes past the last column (walls there)
    (6, 2048, "f32", "couette", 7, 1),         # two columns per thread (second pulled after the barrier), 1 row
    (10, 1500, "f32", "channel", 8, 2),        # the same with 2 rows per block
    (9, 1800, "f32", "channel", 11, 3),        # and 3 (odd rows per block, last block short)
    # column tiles: rows wider than one block, halo columns each side
    (20, 1100, "f64", "channel", 8, 3),        # two tiles of 550
    (9, 2100, "f64", "periodic", 6, 0),        # three tiles, periodic wrap at both row ends
    (11, 2049, "f64", "couette", 7, 2),        # three tiles, the last one narrower
    (12, 2200, "f32", "channel", 7, 4),        # two tiles of 1100, two columns per thread
    (7, 4500, "f16", "periodic", 6, 0),        # three tiles
    (8, 2090, "f32", "channel_open", 9, 1),    # tiles of 1045 (ly not a multiple of the line)
]


@pytest.mark.parametrize("lx,ly,storage,kind,steps,rows", CASES)
def test_two_steps_per_launch_equal_single_steps(lx, ly, storage, kind, steps, rows):
    probe = (lx - 2, ly // 2)
    (a, pa) = _run(lx, ly, storage, kind, steps, False, probe=probe)
    (b, pb) = _run(lx, ly, storage, kind, steps, True, rows=rows, probe=probe)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.array_equal(pa, pb)


TILE_CASES = [
    # (lx, ly, storage, boundary, steps, rows, (threads per tile, columns per thread))
    (20, 2048, "f32", "channel", 7, 0, (64, 2)),        # 17 tiles of 121 columns, one wave each
    (20, 2048, "f32", "channel", 7, 5, (64, 1)),        # 34 tiles of 61
    (14, 2048, "f16", "channel", 6, 3, (128, 1)),       # 17 tiles of 121, two waves each
    (12, 1500, "f32", "couette", 8, 2, (256, 2)),       # 3 tiles of 500
    (9, 1000, "f64", "periodic", 6, 2, (64, 0)),        # 17 tiles, periodic wrap at both row ends
    (15, 130, "f32", "channel_open", 9, 4, (64, 2)),    # 2 tiles of 65, the outlet in every tile
    (7, 100, "f32", "channel", 8, 0, (64, 1)),          # 2 tiles of 50
    (11, 2049, "f32", "couette", 6, 1, (256, 2)),       # 5 tiles, the last narrower
    (10, 64, "f32", "periodic", 6, 0, (64, 1)),         # one tile: whole rows
]


@pytest.mark.parametrize("lx,ly,storage,kind,steps,rows,tile", TILE_CASES)
def test_column_tiles_of_small_blocks_equal_single_steps(lx, ly, storage, kind, steps, rows, tile):
    """Narrow column tiles (blocks of one or a few waves, each tile with a halo
    column on each side) give the same states, outlet values and probe samples
    as single steps."""
    probe = (lx - 2, (ly * 2) // 3)
    (a, pa) = _run(lx, ly, storage, kind, steps, False, probe=probe, seed=ly + lx)
    (b, pb) = _run(lx, ly, storage, kind, steps, True, rows=rows, probe=probe, seed=ly + lx, tile=tile)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.array_equal(pa, pb)


KSTEP_CASES = [
    # (lx, ly, storage, boundary, steps, rows, steps per launch)
    (24, 64, "f32", "periodic", 9, 0, 3),
    (24, 64, "f32", "channel", 10, 5, 3),
    (37, 100, "f32", "channel", 11, 4, 3),        # ly not a multiple of 64
    (3, 17, "f32", "periodic", 8, 0, 3),          # smallest lx: the stage rows wrap around the lattice
    (5, 130, "f16", "channel_open", 9, 1, 3),     # one row asked for: raised so every block holds two
    (33, 1024, "f32", "channel", 7, 0, 3),        # the widest rows
    (26, 200, "f32", "couette", 10, 6, 3),
    (16, 1000, "f16", "channel", 9, 4, 3),
    (12, 514, "f32", "channel", 12, 8, 3),        # a batch-sized channel row, last block short
    (20, 96, "f32", "channel", 13, 3, 4),
    (11, 300, "f16", "channel_open", 14, 2, 4),   # two rows asked for: raised to three
    (4, 50, "f32", "channel", 12, 0, 4),          # lx = 4 with four steps
]


@pytest.mark.parametrize("lx,ly,storage,kind,steps,rows,k", KSTEP_CASES)
def test_k_steps_per_launch_equal_single_steps(lx, ly, storage, kind, steps, rows, k):
    """Three or four steps per launch (the K-stage row sweep) give the same
    states, outlet values and probe samples as single steps; runs whose
    length is not a multiple of K finish with two- and one-step launches."""
    probe = (lx - 2, ly // 3)
    (a, pa) = _run(lx, ly, storage, kind, steps, False, probe=probe, seed=lx * 31 + ly)
    (b, pb) = _run(lx, ly, storage, kind, steps, k, rows=rows, probe=probe, seed=lx * 31 + ly)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.array_equal(pa, pb)


def test_k_steps_probe_ring_wraps_like_single_steps():
    """A probe ring shorter than the run keeps the same last samples with
    three steps per launch as with single steps (slots taken modulo the
    capacity inside a launch)."""
    lx, ly, steps = 30, 100, 23
    probe = (lx - 3, 41)
    from lattice_boltzmann_parallel_solver_amd import Lattice

    outs = []
    for k in (1, 3):
        f, rho, u = _initial(lx, ly, 5)
        lat = Lattice(lx, ly, boundary=_boundary("channel", lx, ly), storage="f32")
        lat.set_state(f, rho, u)
        lat.set_temporal(k)
        lat.set_probe(probe[0], probe[1], 7)
        lat.run(1.2, steps)
        outs.append((lat.probe(), lat.state()))
        lat.close()
    assert outs[0][0].shape == (7, 3)
    assert np.array_equal(outs[0][0], outs[1][0])
    for x, y in zip(outs[0][1], outs[1][1]):
        assert np.array_equal(x, y)


def test_k_steps_rejected_where_unsupported():
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd._lib import LbmError
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import poiseuille_flow_boundary_conditions

    for lx, ly, storage, k in ((16, 64, "f64", 3), (16, 1025, "f32", 3), (16, 1024, "f32", 4), (16, 64, "f32", 5)):
        lat = Lattice(lx, ly, storage=storage)
        with pytest.raises(LbmError):
            lat.set_temporal(k)
        lat.close()
    lat = Lattice(16, 12, boundary=poiseuille_flow_boundary_conditions(16, 12, 1.001 / 3, 0.999 / 3), storage="f32")
    with pytest.raises(LbmError):
        lat.set_temporal(3)
    lat.close()


def test_column_tile_settings_checked():
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd._lib import LbmError

    lat = Lattice(8, 256, storage="f32")
    for bad in ((96, 0), (-64, 0), (2048, 0), (64, 3), (64, -1)):
        with pytest.raises(LbmError):
            lat.set_sweep_tile(*bad)
    lat.set_sweep_tile(64, 1)
    lat.set_temporal(2)
    lat.set_sweep_tile(0, 0)
    lat.close()


def test_two_steps_f64_channel_equal_oracle():
    lx, ly, d = 48, 40, 10
    rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
    f = O.equilibrium(rho, u)
    om = 1 / (3 * 0.1 * d / 100 + 0.5)
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import von_karman_boundary_conditions

    lat = Lattice(lx, ly, boundary=von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1),
                  storage="f64")
    lat.set_state(f, rho, u)
    lat.set_temporal(2, 5)
    lat.set_probe(3 * lx // 4, ly // 2, 60)
    lat.run(om, 60)
    fg, rg, ug = lat.state()
    tr = lat.probe()
    fo, ro, uo, to = N.run(f, rho, u, om, 60, O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True,
                                                                                 y_offset=1),
                           probe=(3 * lx // 4, ly // 2))
    assert np.array_equal(fg, fo) and np.array_equal(rg, ro) and np.array_equal(ug, uo)
    assert np.array_equal(tr, to)


def test_two_steps_f64_column_tiles_equal_oracle():
    """A channel wider than one FP64 block (1100 columns: two tiles)."""
    lx, ly, d = 10, 1100, 200
    rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
    f = O.equilibrium(rho, u)
    om = 1 / (3 * 0.1 * d / 100 + 0.5)
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import von_karman_boundary_conditions

    lat = Lattice(lx, ly, boundary=von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1),
                  storage="f64")
    lat.set_state(f, rho, u)
    lat.set_temporal(2, 3)
    lat.set_probe(lx - 2, 550, 12)
    lat.run(om, 12)
    fg, rg, ug = lat.state()
    tr = lat.probe()
    fo, ro, uo, to = N.run(f, rho, u, om, 12, O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True,
                                                                                 y_offset=1),
                           probe=(lx - 2, 550))
    assert np.array_equal(fg, fo) and np.array_equal(rg, ro) and np.array_equal(ug, uo)
    assert np.array_equal(tr, to)


def test_two_steps_rejected_where_unsupported():
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd._lib import LbmError
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import poiseuille_flow_boundary_conditions

    lat = Lattice(16, 12, boundary=poiseuille_flow_boundary_conditions(16, 12, 1.001 / 3, 0.999 / 3))
    with pytest.raises(LbmError):
        lat.set_temporal(2)
    lat.close()
    lat = Lattice(2, 64, storage="f32")  # fewer than three rows
    with pytest.raises(LbmError):
        lat.set_temporal(2)
    lat.set_temporal(1)
    lat.close()


def test_autotuned_launch_shape_gives_the_same_run():
    """Every launch shape the autotuner can pick (nodes per thread, 16-byte
    groups, XCD order, one or two steps per launch, rows per sweep block)
    computes the same step."""
    from lattice_boltzmann_parallel_solver_amd import workloads as W

    job = W.ChannelJob(96, 64, 16, 100.0)
    a = W.build_lattice(job, "f32")
    best = W.autotune_lattice(a, job.omega, steps=6)
    assert len(best) == 5 and best[3] in (1, 2, 3) and best[4] in W.sweep_row_candidates(96, W._compute_units(0))
    b = W.build_lattice(job, "f32")
    a.run(job.omega, 25)
    b.run(job.omega, 25)
    for x, y in zip(a.state(), b.state()):
        assert np.array_equal(x, y)
    a.close()
    b.close()


def _random_cases(n=16, seed=7):
    """Seeded shapes around the kernel's edges: rows just below or above a
    multiple of the wavefront, of one block (1024 / 2048 columns) and of the
    column-tile sizes."""
    rng = np.random.default_rng(seed)
    edges = [1, 2, 63, 65, 127, 129, 1023, 1025, 1030, 2047, 2049, 2111, 3070, 4093]
    out = []
    for _ in range(n):
        ly = int(rng.choice(edges)) + int(rng.integers(-1, 2))
        ly = max(ly, 1)
        lx = int(rng.integers(3, 24))
        storage = str(rng.choice(["f64", "f32", "f16"]))
        kind = str(rng.choice(["periodic", "channel", "couette", "channel_open"]))
        if kind != "periodic" and ly < 8:
            kind = "periodic"
        rows = int(rng.choice([0, 1, 3, 8]))
        out.append((lx, ly, storage, kind, int(rng.integers(4, 9)), rows))
    return out


@pytest.mark.parametrize("lx,ly,storage,kind,steps,rows", _random_cases())
def test_two_steps_equal_single_steps_random_shapes(lx, ly, storage, kind, steps, rows):
    probe = (lx // 2, ly // 2)
    (a, pa) = _run(lx, ly, storage, kind, steps, False, probe=probe, seed=lx * 7919 + ly)
    (b, pb) = _run(lx, ly, storage, kind, steps, True, rows=rows, probe=probe, seed=lx * 7919 + ly)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.array_equal(pa, pb)


def test_autotuner_cache_keys_and_fallback():
    """Lattices of one size bucket with and without 16-byte groups are tuned
    apart, and a cached two-step shape falls back to one step on a lattice of
    the same bucket that the sweep kernel refuses (the pressure fix)."""
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd import workloads as W
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import poiseuille_flow_boundary_conditions

    W._LATTICE_CACHE.clear()
    a = W.build_lattice(W.ChannelJob(120, 52, 12, 80.0), "f32")    # ny = 52: 16-byte groups
    b = W.build_lattice(W.ChannelJob(120, 51, 12, 80.0), "f32")    # ny = 51: none
    W.autotune_lattice(a, W.ChannelJob(120, 52, 12, 80.0).omega, steps=4)
    W.autotune_lattice(b, W.ChannelJob(120, 51, 12, 80.0).omega, steps=4)
    keys = sorted(W._LATTICE_CACHE)
    assert len(keys) == 2 and {k[2] for k in keys} == {True, False}
    # force a two-step entry for the bucket of a pressure-fix box and check the fallback
    p = Lattice(120, 52, boundary=poiseuille_flow_boundary_conditions(120, 52, 1.0015 / 3, 0.9985 / 3), storage="f32")
    p.set_equilibrium(np.ones((120, 52)), np.zeros((120, 52, 2)))
    key = ("f32", W.size_bucket(120 * 52), True)
    W._LATTICE_CACHE[key] = (1, True, False, 2, 1)
    best = W.autotune_lattice(p, 1.4, steps=4)
    assert best[3] == 1
    p.run(1.4, 5)
    for lat in (a, b, p):
        lat.close()
    W._LATTICE_CACHE.clear()
