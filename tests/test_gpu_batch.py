"""Batched step (one launch per step for many lattices, or two steps per
launch with the batched sweep kernel) against individual lattices run one by
one, bit for bit, including graph replays interleaved with eager work, odd
step counts, mixed pending/running contexts, probes and jobs the sweep kernel
does not take (the pressure fix); and against the CPU oracle: FP64 batches bit
for bit, FP32 batches within the stated tolerance of the FP64 oracle."""

import numpy as np
import pytest

from oracle import lbm_oracle as O
from oracle import lbm_oracle_native as N

pytestmark = pytest.mark.gpu

# FP32 deviation storage against the FP64 oracle over these runs (at most 42
# steps of channels with u = 0.1): max absolute difference of f, rho and u.
F32_TOL = {"f": 1e-6, "rho": 2e-6, "u": 2e-6}


def job_list(seed=0, n=6, pressure=False, wide=False):
    """Uneven channels plus a Couette and a periodic lattice (and a Poiseuille
    box): (lx, ly, kind, plate, rho, u, omega) per job."""
    from lattice_boltzmann_parallel_solver_amd.initial_values import (density_1_velocity_x_u0_velocity_y_0_initial,
                                                                      sinusoidal_velocity_x)

    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n):
        lx = int(rng.integers(40, 160))
        ly = int(round(3 * lx / 7))
        if wide and k == 0:
            lx, ly = 24, 1300  # two columns per thread in the sweep kernel
        d = max(2, int(round(rng.uniform(0.1, 0.4) * ly)))
        re = rng.uniform(40, 200)
        om = 1.0 / (3 * 0.1 * d / re + 0.5)
        rho, u = density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
        jobs.append((lx, ly, "von_karman", d, rho, u, om))
    jobs.append((33, 21, "couette", 0, np.ones((33, 21)), np.zeros((33, 21, 2)), 1.0))
    rho, u = sinusoidal_velocity_x((17, 64), 0.05)
    jobs.append((17, 64, "periodic", 0, rho, u, 0.7))
    if pressure:
        jobs.append((30, 14, "poiseuille", 0, np.ones((30, 14)), np.zeros((30, 14, 2)), 1.4))
    return jobs


def _boundary(lx, ly, kind, d, oracle=False):
    if oracle:
        mod = O
    else:
        from lattice_boltzmann_parallel_solver_amd import boundary_utils as mod
    if kind == "von_karman":
        return mod.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1)
    if kind == "couette":
        return mod.couette_flow_boundary_conditions(lx, ly, 0.05, 1.0)
    if kind == "poiseuille":
        return mod.poiseuille_flow_boundary_conditions(lx, ly, 1.0015 / 3, 0.9985 / 3)
    return None


def make_jobs(storage="f64", seed=0, n=6, pressure=False, wide=False):
    from lattice_boltzmann_parallel_solver_amd import Lattice

    lats, oms = [], []
    for lx, ly, kind, d, rho, u, om in job_list(seed, n, pressure, wide):
        lat = Lattice(lx, ly, boundary=_boundary(lx, ly, kind, d), storage=storage)
        lat.set_equilibrium(rho, u)
        lats.append(lat)
        oms.append(om)
    return lats, oms


def oracle_runs(steps, seed=0, n=6, pressure=False, wide=False, probe=None, pre=None):
    """The oracle's (f, rho, u, trace) of every job of job_list after `steps`
    steps from its equilibrium; probe(lx, ly) gives the probe node, pre[k]
    steps run before the probe starts recording (the trace covers the rest)."""
    out = []
    for k, (lx, ly, kind, d, rho, u, om) in enumerate(job_list(seed, n, pressure, wide)):
        p = probe(lx, ly) if probe else None
        first = pre[k] if pre else 0
        f, r, v, trace = N.run(O.equilibrium(rho, u), rho, u, om, steps, _boundary(lx, ly, kind, d, oracle=True),
                               probe=p)
        out.append((f, r, v, trace[first:]))
    return out


def check_oracle(lats, want, storage, probes=False):
    for lat, w in zip(lats, want):
        got = lat.state()
        for a, b, name in zip(got, w[:3], ("f", "rho", "u")):
            if storage == "f64":
                assert np.array_equal(a, b), (lat.lx, lat.ly, name, float(np.abs(a - b).max()))
            else:
                err = float(np.abs(a - b).max())
                assert err <= F32_TOL[name], (lat.lx, lat.ly, name, err)
        if probes:
            smp = lat.probe()
            if storage == "f64":
                assert np.array_equal(smp, w[3]), (lat.lx, lat.ly)
            else:
                assert np.abs(smp - w[3]).max() <= F32_TOL["u"], (lat.lx, lat.ly)


@pytest.mark.parametrize("v,graph", [(1, True), (2, True), (1, False)])
def test_batch_equals_individual_runs_fp64(v, graph):
    from lattice_boltzmann_parallel_solver_amd import Batch

    lats, oms = make_jobs()
    ref, _ = make_jobs()
    b = Batch(lats, oms, nodes_per_thread=v, use_graph=graph)
    for chunk in (7, 10, 1, 24):  # odd and even chunks, graph replays between eager reads
        b.run(chunk)
        for lat, r, om in zip(lats, ref, oms):
            r.run(om, chunk)
            got, want = lat.state(), r.state()
            for a, c in zip(got, want):
                assert np.array_equal(a, c)
    b.close()
    check_oracle(lats, oracle_runs(42), "f64")


def test_batch_with_running_and_pending_contexts_and_probes():
    from lattice_boltzmann_parallel_solver_amd import Batch

    lats, oms = make_jobs(seed=3, n=3)
    ref, _ = make_jobs(seed=3, n=3)
    lats[0].run(oms[0], 5)
    ref[0].run(oms[0], 5)
    for lat, r in zip(lats, ref):
        lat.set_probe(lat.lx // 2, lat.ly // 2, 100)
        r.set_probe(r.lx // 2, r.ly // 2, 100)
    b = Batch(lats, oms)
    b.run(12)
    b.run(9)
    for lat, r, om in zip(lats, ref, oms):
        r.run(om, 21)
        assert lat.steps == r.steps
        assert np.array_equal(lat.probe(), r.probe())
        for a, c in zip(lat.state(), r.state()):
            assert np.array_equal(a, c)
    b.close()
    want = [oracle_runs(26, seed=3, n=3, probe=lambda lx, ly: (lx // 2, ly // 2), pre=[5] + [0] * 4)[0]]
    want += oracle_runs(21, seed=3, n=3, probe=lambda lx, ly: (lx // 2, ly // 2))[1:]
    check_oracle(lats, want, "f64", probes=True)


def test_batch_fp32_matches_individual():
    from lattice_boltzmann_parallel_solver_amd import Batch

    lats, oms = make_jobs(storage="f32", seed=5)
    ref, _ = make_jobs(storage="f32", seed=5)
    b = Batch(lats, oms)
    b.run(40)
    for lat, r, om in zip(lats, ref, oms):
        r.run(om, 40)
        for a, c in zip(lat.state(), r.state()):
            assert np.array_equal(a, c)
    b.close()
    check_oracle(lats, oracle_runs(40, seed=5), "f32")


def test_batch_errors():
    from lattice_boltzmann_parallel_solver_amd import Batch, Lattice
    from lattice_boltzmann_parallel_solver_amd._lib import LbmError

    a = Lattice(8, 8, storage="f64")
    c = Lattice(8, 8, storage="f32")
    with pytest.raises(LbmError, match="share"):
        Batch([a, c], [1.0, 1.0])
    with pytest.raises(LbmError, match="omega"):
        Batch([a], [2.0])
    b = Batch([a], [1.0])
    with pytest.raises(LbmError, match="no state"):
        b.run(3)
    b.close()


@pytest.mark.parametrize("storage,wide", [("f64", False), ("f32", True), ("f16", True)])
def test_batched_sweep_equals_individual_single_steps(storage, wide):
    """Two steps per launch for the batch: channels of several block widths,
    a Couette box, a periodic lattice and a Poiseuille box (pressure fix, so
    it stays on the one-step batched kernel) against single-step runs."""
    from lattice_boltzmann_parallel_solver_amd import Batch

    lats, oms = make_jobs(storage=storage, seed=7, pressure=True, wide=wide and storage != "f64")
    ref, _ = make_jobs(storage=storage, seed=7, pressure=True, wide=wide and storage != "f64")
    for lat, r in zip(lats, ref):
        lat.set_probe(lat.lx // 3, lat.ly // 2, 200)
        r.set_probe(r.lx // 3, r.ly // 2, 200)
    b = Batch(lats, oms, temporal=2, rows_per_block=5)
    for chunk in (1, 9, 14, 3, 8):  # pending start, odd remainders, graph replays of step pairs
        b.run(chunk)
        for lat, r, om in zip(lats, ref, oms):
            r.run(om, chunk)
    for lat, r in zip(lats, ref):
        assert lat.steps == r.steps
        assert np.array_equal(lat.probe(), r.probe())
        for a, c in zip(lat.state(), r.state()):
            assert np.array_equal(a, c)
    b.close()
    if storage in ("f64", "f32"):
        w = wide and storage != "f64"
        check_oracle(lats, oracle_runs(35, seed=7, pressure=True, wide=w, probe=lambda lx, ly: (lx // 3, ly // 2)),
                     storage, probes=True)


def test_batch_switching_between_one_and_two_steps_per_launch_keeps_probes():
    """Runs alternating between the one-step and the two-step batched paths
    keep every job's probe ring slots and state equal to single runs."""
    from lattice_boltzmann_parallel_solver_amd import Batch

    lats, oms = make_jobs(storage="f32", seed=11, n=4, pressure=True)
    ref, _ = make_jobs(storage="f32", seed=11, n=4, pressure=True)
    for lat, r in zip(lats, ref):
        lat.set_probe(lat.lx // 2, lat.ly // 3, 300)
        r.set_probe(r.lx // 2, r.ly // 3, 300)
    b = Batch(lats, oms)
    for temporal, chunk in ((1, 5), (2, 12), (1, 7), (2, 9), (2, 4)):
        b.set_temporal(temporal)
        b.run(chunk)
        for r, om in zip(ref, oms):
            r.run(om, chunk)
    for lat, r in zip(lats, ref):
        assert np.array_equal(lat.probe(), r.probe())
        for a, c in zip(lat.state(), r.state()):
            assert np.array_equal(a, c)
    b.close()


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_batched_sweep_edge_widths_equal_individual_runs(storage):
    """Channels whose rows end just past a wavefront or a block (a whole wave of
    lanes past the last column at ly = 1030) and, in FP64, rows wider than a
    block, which the batch advances one step per launch beside the sweep
    groups; every lattice equals its own single-step run."""
    from lattice_boltzmann_parallel_solver_amd import Batch, Lattice
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import von_karman_boundary_conditions
    from lattice_boltzmann_parallel_solver_amd.initial_values import density_1_velocity_x_u0_velocity_y_0_initial

    shapes = [(9, 65), (7, 1030), (6, 1025), (5, 129), (8, 1100)]

    def build():
        lats, oms = [], []
        for lx, ly in shapes:
            d = max(2, ly // 4)
            lat = Lattice(lx, ly, boundary=von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1),
                          storage=storage)
            rho, u = density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
            lat.set_equilibrium(rho, u)
            lats.append(lat)
            oms.append(1.0 / (3 * 0.1 * d / 100.0 + 0.5))
        return lats, oms

    lats, oms = build()
    ref, _ = build()
    b = Batch(lats, oms, temporal=2, rows_per_block=3)
    for chunk in (9, 4):
        b.run(chunk)
        for lat, r, om in zip(lats, ref, oms):
            r.run(om, chunk)
            for a, c in zip(lat.state(), r.state()):
                assert np.array_equal(a, c), (lat.lx, lat.ly)
    b.close()
    for lat in lats + ref:
        lat.close()


@pytest.mark.parametrize("temporal", [2, 3])
def test_large_batch_groups_on_side_streams_equal_individual_runs(temporal):
    """A batch above 8 Mi nodes with several groups and a pressure-fix job runs
    its groups on side streams; states and probe traces still equal each
    lattice's own single-step run, over several runs of odd and even length.
    With three steps per launch the groups of 8 and of 15-16 waves per row
    take the three-step kernel, whose runs end with a two-step launch or a
    single step where the run's length is not a multiple of three."""
    from lattice_boltzmann_parallel_solver_amd import Batch, Lattice
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import (poiseuille_flow_boundary_conditions,
                                                                      von_karman_boundary_conditions)
    from lattice_boltzmann_parallel_solver_amd.initial_values import density_1_velocity_x_u0_velocity_y_0_initial

    shapes = [(2048, 2048), (3000, 1000), (1600, 900), (700, 300), (900, 300), (1100, 480)]
    chunks = (6, 7, 4) if temporal == 2 else (6, 7, 4, 8)

    def build():
        lats, oms = [], []
        for lx, ly in shapes:
            d = max(2, ly // 5)
            lat = Lattice(lx, ly, boundary=von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1),
                          storage="f32")
            rho, u = density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
            lat.set_equilibrium(rho, u)
            lat.set_probe(3 * lx // 4, ly // 2, 64)
            lats.append(lat)
            oms.append(1.0 / (3 * 0.1 * d / 100.0 + 0.5))
        lat = Lattice(64, 32, boundary=poiseuille_flow_boundary_conditions(64, 32, 1.0015 / 3, 0.9985 / 3), storage="f32")
        lat.set_equilibrium(np.ones((64, 32)), np.zeros((64, 32, 2)))
        lat.set_probe(32, 16, 64)
        lats.append(lat)
        oms.append(1.5)
        return lats, oms

    lats, oms = build()
    assert sum(l.lx * l.ly for l in lats) >= 8 << 20
    ref, _ = build()
    b = Batch(lats, oms, temporal=temporal)
    for chunk in chunks:
        b.run(chunk)
        for r, om in zip(ref, oms):
            r.run(om, chunk)
    b.close()
    for lat, r in zip(lats, ref):
        for a, c in zip(lat.state(), r.state()):
            assert np.array_equal(a, c), (lat.lx, lat.ly)
        assert np.array_equal(lat.probe(), r.probe()), (lat.lx, lat.ly)
    # against the FP64 oracle
    total = sum(chunks)
    want = []
    for (lx, ly), om in zip(shapes, oms):
        rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
        bc = O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, max(2, ly // 5), walls=True, y_offset=1)
        want.append(N.run(O.equilibrium(rho, u), rho, u, om, total, bc, probe=(3 * lx // 4, ly // 2)))
    rho, u = O.density_1_velocity_0_initial((64, 32))
    want.append(N.run(O.equilibrium(rho, u), rho, u, 1.5, total,
                      O.poiseuille_flow_boundary_conditions(64, 32, 1.0015 / 3, 0.9985 / 3), probe=(32, 16)))
    check_oracle(lats, want, "f32", probes=True)
    for lat in lats + ref:
        lat.close()


def test_large_single_group_batch_with_three_steps_equals_individual_runs():
    """A batch above 8 Mi nodes of one width that three steps per launch pay
    for (480 columns: 8 waves) runs it on a side stream even as a single
    group; states and probes equal single-step runs over runs whose length
    leaves no, one and two steps after the three-step launches."""
    from lattice_boltzmann_parallel_solver_amd import Batch, Lattice
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import von_karman_boundary_conditions
    from lattice_boltzmann_parallel_solver_amd.initial_values import density_1_velocity_x_u0_velocity_y_0_initial

    lx, ly, n = 1100, 480, 17

    def build():
        lats, oms = [], []
        for k in range(n):
            d = 40 + 5 * k
            lat = Lattice(lx, ly, boundary=von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1),
                          storage="f32")
            rho, u = density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
            lat.set_equilibrium(rho, u)
            lat.set_probe(3 * lx // 4 - k, ly // 2, 64)
            lats.append(lat)
            oms.append(1.0 / (3 * 0.1 * d / 100.0 + 0.5))
        return lats, oms

    lats, oms = build()
    assert sum(l.lx * l.ly for l in lats) >= 8 << 20
    ref, _ = build()
    b = Batch(lats, oms, temporal=3)
    for chunk in (7, 4, 8, 3):  # after the aligning first step: 6, 4, 8, 3 (+1 odd) steps
        b.run(chunk)
        for r, om in zip(ref, oms):
            r.run(om, chunk)
    b.close()
    for lat, r in zip(lats, ref):
        for a, c in zip(lat.state(), r.state()):
            assert np.array_equal(a, c)
        assert np.array_equal(lat.probe(), r.probe())
    for lat in lats + ref:
        lat.close()


@pytest.mark.parametrize("storage,rows", [("f32", 8), ("f32", 3), ("f16", 5)])
def test_packed_launch_equals_grouped_launches(storage, rows):
    """One packed launch per step pair (row ranges of many widths side by side
    in 16-wave blocks; one block per packed block, and a persistent grid)
    against one launch per group of equal width: states and
    probe traces equal bit for bit over runs of odd and even length, with a
    job wider than 1024 columns (two columns per thread, kept in its own
    group), a job shorter than a row range and a pressure-fix job (one-step
    kernel) in the same batch; and the FP32 result against the FP64 oracle."""
    from lattice_boltzmann_parallel_solver_amd import Batch, Lattice
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import (poiseuille_flow_boundary_conditions,
                                                                      von_karman_boundary_conditions)
    from lattice_boltzmann_parallel_solver_amd.initial_values import density_1_velocity_x_u0_velocity_y_0_initial

    shapes = [(120, 51), (97, 300), (2, 64), (64, 514), (33, 129), (5, 1300), (40, 86), (200, 257)]

    def build():
        lats, oms = [], []
        for lx, ly in shapes:
            d = max(2, ly // 4)
            lat = Lattice(lx, ly, boundary=von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1)
                          if lx >= 3 else None, storage=storage)
            rho, u = density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
            lat.set_equilibrium(rho, u)
            lat.set_probe(lx // 2, ly // 3, 64)
            lats.append(lat)
            oms.append(1.0 / (3 * 0.1 * d / 100.0 + 0.5))
        lat = Lattice(30, 14, boundary=poiseuille_flow_boundary_conditions(30, 14, 1.0015 / 3, 0.9985 / 3),
                      storage=storage)
        lat.set_equilibrium(np.ones((30, 14)), np.zeros((30, 14, 2)))
        lats.append(lat)
        oms.append(1.4)
        return lats, oms

    print(f"case: packed (one block per packed block, persistent) vs grouped batch, {storage}, {rows} rows per range",
          flush=True)
    out = []
    for packed in (1, 0, 2):
        lats, oms = build()
        b = Batch(lats, oms, temporal=2, rows_per_block=rows, packed=packed)
        for chunk in (5, 8, 3, 10):
            b.run(chunk)
        b.close()
        out.append(([lat.state() for lat in lats], [lat.probe() for lat in lats[:-1]]))
        for lat in lats:
            lat.close()
    for (sa, pa), (sb, pb) in zip([out[0], out[2]], [out[1], out[1]]):
        for k, (a, c) in enumerate(zip(sa, sb)):
            for x, y in zip(a, c):
                assert np.array_equal(x, y), k
        for k, (a, c) in enumerate(zip(pa, pb)):
            assert np.array_equal(a, c), k
    if storage == "f32":
        for k, (lx, ly) in enumerate(shapes):
            if lx < 3:
                continue
            d = max(2, ly // 4)
            rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
            bc = O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1)
            want = N.run(O.equilibrium(rho, u), rho, u, 1.0 / (3 * 0.1 * d / 100.0 + 0.5), 26, bc)
            for a, w, name in zip(out[0][0][k], want[:3], ("f", "rho", "u")):
                assert np.abs(a - w).max() <= F32_TOL[name], (k, name)


SMALL_GROUPS = [(60, 480), (50, 700), (41, 200), (30, 96)]


@pytest.mark.parametrize("temporal,shapes,rows", [
    (2, [(40, 200), (33, 200), (57, 200)], 0),  # one group: graph-captured pairs, one row per block
    (3, SMALL_GROUPS, 0),                       # several groups, too few rows for three steps per launch
    (3, SMALL_GROUPS, 2),                       # the same with three steps for the 8- and 11-wave groups
])
def test_small_batch_rows_and_side_streams_equal_individual_runs(temporal, shapes, rows):
    """Small batches: with automatic rows per block (few rows per block, so
    that the batch's few rows still fill the GPU) one width group replays
    captured step pairs on the batch stream and several groups run on side
    streams with two steps per launch; with rows given, the FP32 groups of 8
    and of 11 waves per row take three steps per launch on their side
    streams.  States and probes equal each lattice's own single-step run, and
    the FP32 oracle's after the last chunk."""
    from lattice_boltzmann_parallel_solver_amd import Batch, Lattice
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import von_karman_boundary_conditions
    from lattice_boltzmann_parallel_solver_amd.initial_values import density_1_velocity_x_u0_velocity_y_0_initial

    chunks = (5, 8, 7, 4)

    def build():
        lats, oms = [], []
        for lx, ly in shapes:
            d = max(2, ly // 5)
            lat = Lattice(lx, ly, boundary=von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1),
                          storage="f32")
            rho, u = density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
            lat.set_equilibrium(rho, u)
            lat.set_probe(3 * lx // 4, ly // 2, 64)
            lats.append(lat)
            oms.append(1.0 / (3 * 0.1 * d / 100.0 + 0.5))
        return lats, oms

    lats, oms = build()
    ref, _ = build()
    b = Batch(lats, oms, temporal=temporal, rows_per_block=rows)
    for chunk in chunks:
        b.run(chunk)
        for r, om in zip(ref, oms):
            r.run(om, chunk)
    b.close()
    for lat, r in zip(lats, ref):
        for a, c in zip(lat.state(), r.state()):
            assert np.array_equal(a, c), (lat.lx, lat.ly)
        assert np.array_equal(lat.probe(), r.probe()), (lat.lx, lat.ly)
    total = sum(chunks)
    want = []
    for (lx, ly), om in zip(shapes, oms):
        rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
        bc = O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, max(2, ly // 5), walls=True, y_offset=1)
        want.append(N.run(O.equilibrium(rho, u), rho, u, om, total, bc, probe=(3 * lx // 4, ly // 2)))
    check_oracle(lats, want, "f32", probes=True)
    for lat in lats + ref:
        lat.close()
