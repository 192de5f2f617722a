"""Batched channel workloads, the sequential baseline and the launch autotuner.

The reference runs its simulations one after another (main.py:150-207, one
x_strouhal call per Reynolds number / length / blockage ratio).  Here many
uneven channels advance together: ``run_batched`` puts them in one Batch (one
kernel launch per step for all of them, step pairs replayed from a HIP
graph), ``run_sequential`` solves them one at a time with the plain
single-lattice path, for comparison.

``autotune_lattice`` / ``autotune_batch`` pick the launch shape of the fused
step per size bucket by timing the candidates on scratch lattices of the same
shape and boundary (the real state is never advanced by tuning); results are
cached per (storage, size bucket), and for single lattices also per whether ny
allows 16-byte groups.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .boundary_utils import von_karman_boundary_conditions
from .initial_values import density_1_velocity_x_u0_velocity_y_0_initial
from .solver import Batch, Lattice


@dataclass
class ChannelJob:
    """One channel: plate of blockage d/ly at x = lx//4 (+1 y offset), rigid
    walls, inlet u = 0.1."""
    lx: int
    ly: int
    plate: int
    reynolds: float
    u_in: float = 0.1

    @property
    def nu(self) -> float:
        return self.u_in * self.plate / self.reynolds

    @property
    def omega(self) -> float:
        return 1.0 / (3.0 * self.nu + 0.5)

    @property
    def nodes(self) -> int:
        return self.lx * self.ly


def channel_jobs(n: int = 256, seed: int = 0, nx_range=(200, 1200), blockage=(0.1, 0.4),
                 reynolds=(40.0, 200.0)) -> List[ChannelJob]:
    """Seeded draw: nx ~ U[200, 1200], ny = round(3 nx / 7), B ~ U[0.1, 0.4],
    d = round(B ny), Re ~ U[40, 200]."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n):
        lx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ly = int(round(3 * lx / 7))
        d = max(2, int(round(rng.uniform(*blockage) * ly)))
        re = float(rng.uniform(*reynolds))
        jobs.append(ChannelJob(lx, ly, d, re))
    return jobs


def build_lattice(job: ChannelJob, storage: str = "f32", device: int = 0) -> Lattice:
    lat = Lattice(job.lx, job.ly,
                  boundary=von_karman_boundary_conditions(job.lx, job.ly, 1.0, job.u_in, job.plate, walls=True,
                                                          y_offset=1),
                  storage=storage, device=device)
    rho, u = density_1_velocity_x_u0_velocity_y_0_initial((job.lx, job.ly), job.u_in)
    lat.set_equilibrium(rho, u)
    return lat


def _sync_all(lattices):
    for lat in lattices:
        lat.synchronize()


def run_sequential(jobs: Sequence[ChannelJob], steps: int, storage: str = "f32", tune: bool = False,
                   temporal: int = 1):
    """One channel at a time (create, run `steps`, finish), as the reference's
    sweeps do, with one or two steps per launch (or the autotuned launch shape
    if `tune`).  Returns (seconds, aggregate MLUPS)."""
    from ._lib import LbmError

    total = 0.0
    for job in jobs:
        lat = build_lattice(job, storage)
        if tune:
            autotune_lattice(lat, job.omega)
        elif temporal == 2:
            try:
                lat.set_temporal(2)
            except LbmError:
                pass
        lat.synchronize()
        t0 = time.perf_counter()
        lat.run(job.omega, steps)
        lat.synchronize()
        total += time.perf_counter() - t0
        lat.close()
    return total, sum(j.nodes for j in jobs) * steps / total / 1e6


def run_batched(jobs: Sequence[ChannelJob], steps: int, storage: str = "f32", nodes_per_thread: Optional[int] = None,
                use_graph: bool = True, chunk: int = 1000, keep: bool = False, temporal: int = 2,
                rows_per_block: int = 0, packed: int = 0):
    """All channels in one Batch, two steps per launch by default (one launch
    per group of equal row width; packed=1 one packed launch per step pair,
    packed=2 persistent).  Returns (seconds, aggregate MLUPS, lattices or None)."""
    lats = [build_lattice(j, storage) for j in jobs]
    oms = [j.omega for j in jobs]
    if nodes_per_thread is None:
        nodes_per_thread = autotune_batch(jobs, storage) if temporal == 1 else 1
    b = Batch(lats, oms, nodes_per_thread=nodes_per_thread, use_graph=use_graph, temporal=temporal,
              rows_per_block=rows_per_block, packed=packed)
    b.run(1)  # first collisions, tables, graph capture
    _sync_all(lats)
    t0 = time.perf_counter()
    done = 1
    while done < steps:
        k = min(chunk, steps - done)
        b.run(k)
        done += k
    sec = time.perf_counter() - t0
    b.close()
    mlups = sum(j.nodes for j in jobs) * (steps - 1) / sec / 1e6
    if not keep:
        for lat in lats:
            lat.close()
        lats = None
    return sec, mlups, lats


# ----------------------------------------------------------------------------- autotuning
_LATTICE_CACHE: Dict[Tuple[str, int, bool], Tuple[int, bool, bool, int, int]] = {}
_BATCH_CACHE: Dict[Tuple[str, int], int] = {}


def size_bucket(nodes: int) -> int:
    """Half-octave bucket of a node count."""
    return int(round(2 * math.log2(max(1, nodes))))


def lattice_candidates(ly: int, storage: str):
    vec = {"f64": 2, "f32": 4, "f16": 8}[storage]
    cands = [(v, xcd, False) for v in (1, 2, 4) for xcd in (False, True)]
    if ly % vec == 0:
        cands += [(0, False, False), (0, True, False)]
    return cands


def _compute_units(device: int) -> int:
    """CUs of the device (0 where torch cannot tell)."""
    try:
        import torch

        return int(torch.cuda.get_device_properties(device).multi_processor_count)
    except Exception:
        return 0


def sweep_row_candidates(lx: int, cus: int) -> List[int]:
    """Rows per block to try for the sweep kernels: the occupancy-derived
    count (0), a few short blocks (small lattices fill the chip only with
    those), and for large lattices one block per CU and slightly fewer blocks
    than CUs (8192x2048 FP32, two steps: 137.4 GLUPS at 32 rows, 139.5 at
    34, 139.3 at 36; FP16 and three steps were fastest at 32)."""
    rows = [0, 1, 2, 4]
    if cus > 0:
        per_cu = -(-lx // cus)
        rows += [r for r in (per_cu, per_cu + 2, per_cu + 4) if r > 4]
    return list(dict.fromkeys(rows))


def _time_run(lat, omega, steps) -> float:
    lat.synchronize()
    t0 = time.perf_counter()
    lat.run(omega, steps)
    lat.synchronize()
    return time.perf_counter() - t0


def autotune_lattice(lat: Lattice, omega: float, steps: int = 30) -> Tuple[int, bool, bool, int, int]:
    """Fastest launch shape (nodes per thread, XCD order, non-temporal stores,
    steps per launch, rows per block of the sweep) for a lattice of this
    storage and size bucket, timed on a scratch lattice with the same shape
    and boundary; applied to `lat` and cached.  Two and three steps per launch
    are tried where the lattice supports them, with the occupancy-derived row
    count (0) and with a few rows per block: lattices of a few hundred rows
    fill the chip only with very short blocks."""
    from ._lib import LbmError

    # 16-byte groups exist only where ny is a multiple of their node count, so
    # lattices with and without that property are tuned apart
    key = (lat.storage, size_bucket(lat.lx * lat.ly), lat.ly % {"f64": 2, "f32": 4, "f16": 8}[lat.storage] == 0)
    best = _LATTICE_CACHE.get(key)
    if best is None:
        scratch = Lattice(lat.lx, lat.ly, boundary=lat._program, storage=lat.storage, device=lat.device)
        rho = np.ones((lat.lx, lat.ly))
        u = np.zeros((lat.lx, lat.ly, 2))
        u[..., 0] = 0.05
        scratch.set_equilibrium(rho, u)
        scratch.run(omega, 2)
        timings = []
        for cand in lattice_candidates(lat.ly, lat.storage):
            scratch.set_launch(*cand)
            _time_run(scratch, omega, 3)
            timings.append((min(_time_run(scratch, omega, steps) for _ in range(2)), cand + (1, 0)))
        scratch.set_launch(1, True, False)
        rows_list = sweep_row_candidates(lat.lx, _compute_units(lat.device))
        for k, rows in [(k, r) for k in (2, 3) for r in rows_list]:
            try:
                scratch.set_temporal(k, rows)
            except LbmError:
                continue
            _time_run(scratch, omega, 6)
            timings.append((min(_time_run(scratch, omega, steps) for _ in range(2)), (1, True, False, k, rows)))
        scratch.close()
        best = min(timings)[1]
        _LATTICE_CACHE[key] = best
    lat.set_launch(*best[:3])
    # a lattice of the same size bucket may not take the cached shape: fall
    # back to two steps, then to one
    for k in range(best[3], 0, -1):
        try:
            lat.set_temporal(k, best[4] if k == best[3] else 0)
            return best if k == best[3] else best[:3] + (k, 0)
        except LbmError:
            continue
    return best


def autotune_batch(jobs: Sequence[ChannelJob], storage: str = "f32", sample: int = 16, steps: int = 20) -> int:
    """Nodes per thread of the batched kernel (1 or 2) for this storage and
    mean job size, timed on a batch of scratch lattices of the first
    `sample` job shapes; cached."""
    key = (storage, size_bucket(int(np.mean([j.nodes for j in jobs]))))
    best = _BATCH_CACHE.get(key)
    if best is None:
        sub = list(jobs[:sample])
        lats = [build_lattice(j, storage) for j in sub]
        b = Batch(lats, [j.omega for j in sub])
        b.run(2)
        timings = []
        for v in (1, 2):
            b.set_launch(v, True)
            b.run(2)
            t0 = time.perf_counter()
            b.run(steps)
            timings.append((time.perf_counter() - t0, v))
        b.close()
        for lat in lats:
            lat.close()
        best = min(timings)[1]
        _BATCH_CACHE[key] = best
    return best


def _main(argv=None):
    """Batched channel workload: `python -m lattice_boltzmann_parallel_solver_amd.workloads
    --jobs 256 --steps 20000` prints aggregate MLUPS of the batch (two steps
    per launch in per-width groups, in one packed launch and in a persistent
    packed grid, and one step per launch), and of the same channels solved one at a time, each with the
    autotuned launch shape of its size bucket and with one step per launch."""
    import argparse
    import json

    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--jobs", type=int, default=256)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--storage", default="f32")
    ap.add_argument("--sequential-jobs", type=int, default=0, help="solve only the first N one at a time (0: all)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batch with three and with two steps per launch (for kernel profiles)")
    a = ap.parse_args(argv)
    jobs = channel_jobs(a.jobs, seed=a.seed)
    res = {"jobs": a.jobs, "steps": a.steps, "seed": a.seed, "storage": a.storage,
           "nodes": sum(j.nodes for j in jobs), "nx_range": [min(j.lx for j in jobs), max(j.lx for j in jobs)]}
    print(json.dumps(res), flush=True)
    if a.batched_only:
        for temporal, name in ((3, "batched_three_step"), (2, "batched_two_step")):
            sec, ml, _ = run_batched(jobs, a.steps, a.storage, temporal=temporal)
            res[name] = {"seconds": sec, "mlups": ml}
            print(json.dumps({name: res[name]}), flush=True)
        return
    sec, ml, _ = run_batched(jobs, a.steps, a.storage, temporal=2)
    res["batched_two_step"] = {"seconds": sec, "mlups": ml,
                               "launch": "one launch per group of equal row width, groups on side streams"}
    print(json.dumps({"batched_two_step": res["batched_two_step"]}), flush=True)
    sec, ml, _ = run_batched(jobs, a.steps, a.storage, temporal=3)
    res["batched_three_step"] = {"seconds": sec, "mlups": ml,
                                 "launch": "as batched_two_step; FP32 groups of 385-512 columns take three steps "
                                           "per launch"}
    print(json.dumps({"batched_three_step": res["batched_three_step"]}), flush=True)
    for mode, name in ((1, "batched_two_step_packed"), (2, "batched_two_step_persistent")):
        sec, ml, _ = run_batched(jobs, a.steps, a.storage, temporal=2, packed=mode)
        res[name] = {"seconds": sec, "mlups": ml,
                     "launch": "one packed launch per step pair, captured in a HIP graph on one stream"
                               + (", persistent grid" if mode == 2 else "")}
        print(json.dumps({name: res[name]}), flush=True)
    sec, ml, _ = run_batched(jobs, a.steps, a.storage, temporal=1)
    res["batched_one_step"] = {"seconds": sec, "mlups": ml}
    print(json.dumps({"batched_one_step": res["batched_one_step"]}), flush=True)
    seq = jobs[:a.sequential_jobs] if a.sequential_jobs else jobs
    sec, ml = run_sequential(seq, a.steps, a.storage, tune=True)
    res["sequential"] = {"jobs": len(seq), "seconds": sec, "mlups": ml,
                         "launch": "autotuned per size bucket (one step per launch, or two or three with a few rows per block)",
                         "shapes_chosen": {f"{k[0]}/bucket{k[1]}/{'v16' if k[2] else 'ragged'}": list(v)
                                           for k, v in sorted(_LATTICE_CACHE.items())}}
    res["speedup_two_step_vs_sequential"] = res["batched_two_step"]["mlups"] / ml
    res["speedup_three_step_vs_sequential"] = res["batched_three_step"]["mlups"] / ml
    res["speedup_one_step_vs_sequential"] = res["batched_one_step"]["mlups"] / ml
    print(json.dumps({"sequential": {"mlups": ml}}), flush=True)
    # the plain solve: every lattice alone, one step per launch
    sec, mlp = run_sequential(seq, a.steps, a.storage, tune=False, temporal=1)
    res["sequential_plain"] = {"jobs": len(seq), "seconds": sec, "mlups": mlp,
                               "launch": "one lattice at a time, one step per launch"}
    res["speedup_two_step_vs_sequential_plain"] = res["batched_two_step"]["mlups"] / mlp
    res["speedup_three_step_vs_sequential_plain"] = res["batched_three_step"]["mlups"] / mlp
    text = json.dumps(res, indent=1)
    print(text, flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(text)


if __name__ == "__main__":
    _main()
