"""Velocity-magnitude snapshots and checkpoint/resume on the device.

* Lattice.speed() is numpy's norm over the last axis of the oracle's
  velocity, bit for bit (FP64).
* speed_snapshots takes |u| after steps 1, every+1, ... as the reference's
  loop does (experiments.py:699-707) and writes save_mpiio-layout files.
* A run checkpointed, restored into a fresh context and continued equals the
  uninterrupted run bit for bit, in every storage and with one or two steps
  per launch; pending (not yet stepped) states round-trip too; mismatched or
  corrupt checkpoints are rejected.
"""

import numpy as np
import pytest

from oracle import lbm_oracle as O
from oracle import lbm_oracle_native as N

pytestmark = pytest.mark.gpu


def _channel(lx, ly, storage, temporal=1):
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import von_karman_boundary_conditions

    lat = Lattice(lx, ly, boundary=von_karman_boundary_conditions(lx, ly, 1.0, 0.1, ly // 4, walls=True, y_offset=1),
                  storage=storage)
    rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
    lat.set_equilibrium(rho, u)
    if temporal > 1:
        lat.set_temporal(temporal)
    return lat


def test_speed_field_is_numpys_norm_of_the_oracle_velocity():
    lx, ly, d, steps = 60, 40, 10, 37
    lat = _channel(lx, ly, "f64")
    lat.run(1 / 1.7, steps)
    sp = lat.speed()
    rho, u = O.density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), 0.1)
    _, _, uo, _ = N.run(O.equilibrium(rho, u), rho, u, 1 / 1.7, steps,
                        O.von_karman_boundary_conditions(lx, ly, 1.0, 0.1, d, walls=True, y_offset=1))
    assert sp.shape == (lx, ly)
    assert np.array_equal(sp, np.linalg.norm(uo, axis=-1))
    lat.close()


def test_speed_snapshots_steps_and_files(tmp_path):
    from lattice_boltzmann_parallel_solver_amd.snapshots import speed_snapshots

    lx, ly = 48, 32
    lat = _channel(lx, ly, "f32")
    seen = {}
    taken = speed_snapshots(lat, 1 / 1.7, 250, every=100, directory=str(tmp_path),
                            on_snapshot=lambda s, f: seen.__setitem__(s, f.copy()))
    assert taken == [1, 101, 201]
    steps = lat.steps
    assert steps == 250
    for s in taken:
        ref = _channel(lx, ly, "f32")
        ref.run(1 / 1.7, s)
        want = ref.speed()
        assert np.array_equal(seen[s], want)
        assert np.array_equal(np.load(str(tmp_path / f"vel_norm_{s}.npy")), want)
        ref.close()
    lat.close()


@pytest.mark.parametrize("storage,temporal", [("f64", 1), ("f32", 1), ("f32", 2), ("f16", 2), ("f64", 2), ("f32", 3),
                                              ("f16", 3)])
def test_checkpoint_resume_is_bit_exact(storage, temporal, tmp_path):
    lx, ly = 52, 36
    om = 1 / 1.7
    a = _channel(lx, ly, storage, temporal)
    a.run(om, 301)
    want = a.state()
    b = _channel(lx, ly, storage, temporal)
    b.run(om, 137)
    path = str(tmp_path / "ck.bin")
    b.save_checkpoint(path)
    b.close()
    c = _channel(lx, ly, storage, temporal)   # fresh context, same boundary
    c.run(om, 5)                               # its own state is overwritten
    c.load_checkpoint(path)
    assert c.steps == 137
    c.run(om, 164)
    for x, y in zip(c.state(), want):
        assert np.array_equal(x, y)
    a.close()
    c.close()


def test_pending_checkpoint_and_rejections():
    from lattice_boltzmann_parallel_solver_amd._lib import LbmError

    a = _channel(20, 12, "f32")
    blob = a.checkpoint()              # pending: the step input itself
    b = _channel(20, 12, "f32")
    b.run(1.1, 3)
    b.restore(blob)
    assert b.steps == 0
    a.run(1.1, 9)
    b.run(1.1, 9)
    for x, y in zip(a.state(), b.state()):
        assert np.array_equal(x, y)
    wrong = _channel(20, 14, "f32")
    with pytest.raises(LbmError):
        wrong.restore(blob)
    other = _channel(20, 12, "f64")
    with pytest.raises(LbmError):
        other.restore(blob)
    bad = bytearray(blob)
    bad[0:8] = b"NOTACKPT"
    with pytest.raises(LbmError):
        b.restore(bytes(bad))
    with pytest.raises(LbmError):
        b.restore(blob[:100])
    # format version 1 (rows not padded to 128-byte lines) is refused
    assert int.from_bytes(blob[8:12], "little") == 2
    old = bytearray(blob)
    old[8:12] = (1).to_bytes(4, "little")
    with pytest.raises(LbmError):
        b.restore(bytes(old))
    for lat in (a, b, wrong, other):
        lat.close()
