"""Continuous-batching serving of simulation jobs on one GPU.

Jobs (von Karman channels, Couette and Poiseuille boxes of any size and step
budget) arrive over time.  The server admits waiting jobs while their HBM
footprint fits the budget, keeps every admitted lattice resident, and
advances all of them together with a Batch (one kernel launch per step for
the whole active set), in chunks.  When jobs finish, their results are read,
their memory is freed and waiting jobs are admitted at the next chunk
boundary, so new work overlaps with work already on the device.  The reference
ran one simulation after another (main.py:150-207).

Reported per job: arrival, admission and completion times (time-to-result is
completion - arrival), the result (probe trace and Strouhal numbers for a
channel, the u_x profile for Couette / Poiseuille), and whether the FP64
moment sums stayed finite.
"""

from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import LbmError
from .boundary_utils import (couette_flow_boundary_conditions, poiseuille_flow_boundary_conditions,
                             von_karman_boundary_conditions)
from .experiments import speed, strouhal_interpolated, strouhal_reference
from .initial_values import density_1_velocity_0_initial, density_1_velocity_x_u0_velocity_y_0_initial
from .solver import Batch, Lattice

STORAGE_BYTES = {"f64": 8, "f32": 4, "f16": 2}


@dataclass
class JobSpec:
    kind: str                 # "von_karman", "couette" or "poiseuille"
    lx: int
    ly: int
    steps: int
    storage: str = "f32"
    plate: int = 400          # von Karman
    reynolds: float = 100.0   # von Karman
    u_in: float = 0.1         # von Karman inlet / Couette lid
    omega: float = 1.0        # Couette / Poiseuille relaxation
    delta_p: float = 1e-3     # Poiseuille

    @property
    def nodes(self) -> int:
        return self.lx * self.ly

    @property
    def relaxation(self) -> float:
        if self.kind == "von_karman":
            return 1.0 / (3.0 * self.u_in * self.plate / self.reynolds + 0.5)
        return self.omega

    def footprint(self) -> int:
        """Device bytes of the job's context, an upper bound: two state buffers
        (rows padded to 128-byte lines), outlet buffers, the sweep's scratch
        area, FP64 staging, reduction partials, boundary tables, probe ring,
        and 2 MiB of rounding for each of its twelve allocations."""
        s = STORAGE_BYTES[self.storage]
        lx, ly, n = self.lx, self.ly, self.nodes
        pitch = -(-ly * s // 128) * 128 // s
        waves = lx * -(-ly // 64)
        planes = 2 * 9 * lx * pitch * s
        sides = 2 * 3 * ly * max(s, 4)
        scratch = 18 * 1024 * 8
        staging = 96 * n
        partials = 24 * max(-(-waves // 4), -(-n // 256)) + 24
        special = 2 * lx + 2 * ly + 2 * self.plate + 16
        tables = 2 * n + n // 64 + waves + 32 + 80 * 64 + 4 * special + 5 * 256
        probe = 24 * self.steps if self.kind == "von_karman" else 0
        return planes + sides + scratch + staging + partials + tables + probe + 12 * (2 << 20)

    def boundary(self):
        if self.kind == "von_karman":
            return von_karman_boundary_conditions(self.lx, self.ly, 1.0, self.u_in, self.plate, walls=True,
                                                  y_offset=1)
        if self.kind == "couette":
            return couette_flow_boundary_conditions(self.lx, self.ly, self.u_in, 1.0)
        if self.kind == "poiseuille":
            drho = 3 * self.delta_p
            return poiseuille_flow_boundary_conditions(self.lx, self.ly, (1 + drho / 2) / 3, (1 - drho / 2) / 3)
        raise ValueError(self.kind)

    def start(self, device: int = 0, storage: Optional[str] = None) -> Lattice:
        lat = Lattice(self.lx, self.ly, boundary=self.boundary(), storage=storage or self.storage, device=device)
        if self.kind == "von_karman":
            rho, u = density_1_velocity_x_u0_velocity_y_0_initial((self.lx, self.ly), self.u_in)
            lat.set_probe(3 * self.lx // 4, self.ly // 2, self.steps)
        else:
            rho, u = density_1_velocity_0_initial((self.lx, self.ly))
        lat.set_equilibrium(rho, u)
        return lat

    def result(self, lat: Lattice) -> dict:
        """The job's answer, read from the device."""
        out = {}
        try:
            m, jx, jy = lat.moments_sum()
            out["mass"] = m
            out["finite"] = True
        except LbmError:
            out["finite"] = False
            return out
        if self.kind == "von_karman":
            samples = lat.probe()
            out["samples"] = samples
            start = len(samples) // 2
            if len(samples) - start > 16:
                out["strouhal_uy"] = strouhal_interpolated(samples[:, 2], self.plate, self.u_in, start)
                out["strouhal_speed"] = strouhal_interpolated(speed(samples), self.plate, self.u_in, start)
                out["strouhal_uy_fft"] = strouhal_reference(samples[:, 2], self.plate, self.u_in, start)[0]
        else:
            _, _, u = lat.state(want=("u",))
            out["ux_mid"] = u[self.lx // 2, :, 0].copy()
        return out


@dataclass
class JobRecord:
    job_id: int
    spec: JobSpec
    arrival: float
    admitted: float = math.nan
    finished: float = math.nan
    result: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def c5_jobs(n: int, seed: int = 0, rate: float = 50.0, vk_size=(8192, 2048), square=(128, 2048),
            budget=(5000, 100000), storage: str = "f32") -> List[Tuple[float, JobSpec]]:
    """Seeded serving workload: Poisson arrivals at `rate` jobs/s; 60 % von
    Karman channels of size vk_size (plate ly/5, Re 100), 20 % Couette and 20 %
    Poiseuille on squares with side ~ U[square]; step budgets log-uniform in
    `budget`."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        steps = int(round(math.exp(rng.uniform(math.log(budget[0]), math.log(budget[1])))))
        r = rng.random()
        if r < 0.6:
            lx, ly = vk_size
            spec = JobSpec("von_karman", lx, ly, steps, storage, plate=max(2, ly // 5), reynolds=100.0)
        else:
            side = int(rng.integers(square[0], square[1] + 1))
            if r < 0.8:
                spec = JobSpec("couette", side, side, steps, storage, u_in=0.05, omega=1.0)
            else:
                spec = JobSpec("poiseuille", side, side, steps, storage, omega=1.5, delta_p=1e-3)
        out.append((t, spec))
    return out


class Server:
    """Continuous batching with admission by HBM footprint.

    policy "fifo": jobs are admitted in arrival order while their footprint
    fits, and every admitted job advances in every chunk.
    policy "srpt": waiting jobs are admitted smallest remaining work first, and
    each chunk advances only the admitted jobs with the least remaining work
    (node steps), as many as it takes to reach `fill_nodes` nodes, enough to
    keep the GPU near full rate.  Short jobs then finish without sharing the
    GPU with long ones: lower median time-to-result for the same total work.
    On the 40-job stream of 8192x2048 channels and squares, srpt with 96 Mi
    nodes gave p50 24.7 s / p99 104.4 s at 118.2 GLUPS, against fifo's
    40.6 s / 104.5 s at 119.5 GLUPS.
    Results do not depend on the policy (a job's steps are the same in any
    batch).  Under srpt a long job can wait while shorter ones keep arriving;
    a finite stream always drains."""

    def __init__(self, hbm_budget: int = 160 << 30, chunk_steps: int = 2000, max_active: int = 256,
                 device: int = 0, nodes_per_thread: int = 1, temporal: int = 2, progress=None,
                 policy: str = "srpt", fill_nodes: int = 96 << 20):
        if policy not in ("fifo", "srpt"):
            raise ValueError(f"unknown policy {policy!r}")
        self.policy = policy
        self.fill_nodes = int(fill_nodes)
        self.hbm_budget = int(hbm_budget)
        self.chunk_steps = int(chunk_steps)
        self.max_active = int(max_active)
        self.device = device
        self.nodes_per_thread = nodes_per_thread
        self.temporal = temporal  # steps per launch of the batched step (2: sweep kernel where it applies)
        self.progress = progress  # called after every chunk with (seconds, active, waiting, finished)
        # bring up the HIP runtime and load the kernels before the first job arrives
        warm = Lattice(8, 8, device=device)
        warm.set_equilibrium(np.ones((8, 8)), np.zeros((8, 8, 2)))
        warm.run(1.0, 2)
        warm.synchronize()
        warm.close()

    def run(self, arrivals: Sequence[Tuple[float, JobSpec]], realtime: bool = True) -> List[JobRecord]:
        """Serve the jobs; arrival times are seconds after the call (with
        realtime=False every job is available at once).  Returns the records in
        submission order."""
        records = [JobRecord(i, spec, t if realtime else 0.0) for i, (t, spec) in enumerate(arrivals)]
        for r in records:
            if r.spec.footprint() > self.hbm_budget:
                raise ValueError(f"job {r.job_id} needs {r.spec.footprint()} B, more than the HBM budget")
        incoming = deque(sorted(records, key=lambda r: r.arrival))
        waiting: deque = deque()
        active: Dict[int, Tuple[JobRecord, Lattice, int]] = {}  # id -> (record, lattice, steps left)
        used = 0
        batch: Optional[Batch] = None
        members: Tuple[int, ...] = ()
        t0 = time.perf_counter()
        self.busy = 0.0
        while incoming or waiting or active:
            now = time.perf_counter() - t0
            while incoming and incoming[0].arrival <= now:
                waiting.append(incoming.popleft())
            # admission while the footprint fits: arrival order (fifo) or least work first (srpt)
            if self.policy == "srpt" and len(waiting) > 1:
                waiting = deque(sorted(waiting, key=lambda r: (r.spec.nodes * r.spec.steps, r.job_id)))
            while waiting and len(active) < self.max_active and used + waiting[0].spec.footprint() <= self.hbm_budget:
                rec = waiting.popleft()
                lat = rec.spec.start(self.device)
                rec.admitted = time.perf_counter() - t0
                active[rec.job_id] = (rec, lat, rec.spec.steps)
                used += rec.spec.footprint()
            if not active:
                if incoming:
                    time.sleep(max(0.0, incoming[0].arrival - (time.perf_counter() - t0)))
                continue
            if self.policy == "srpt":
                # the jobs with the least remaining node steps, up to fill_nodes nodes
                chosen, nodes = [], 0
                for i in sorted(active, key=lambda i: (active[i][0].spec.nodes * active[i][2], i)):
                    chosen.append(i)
                    nodes += active[i][0].spec.nodes
                    if nodes >= self.fill_nodes:
                        break
                ids = tuple(sorted(chosen))
            else:
                ids = tuple(sorted(active))
            if ids != members:
                if batch is not None:
                    batch.close()
                batch = Batch([active[i][1] for i in ids], [active[i][0].spec.relaxation for i in ids],
                              nodes_per_thread=self.nodes_per_thread, temporal=self.temporal)
                members = ids
            k = min(self.chunk_steps, min(active[i][2] for i in ids))
            tb = time.perf_counter()
            batch.run(k)
            self.busy += time.perf_counter() - tb
            if self.progress is not None:
                self.progress(time.perf_counter() - t0, len(active), len(waiting) + len(incoming),
                              sum(1 for r in records if not math.isnan(r.finished)))
            finished = []
            for i in ids:
                rec, lat, left = active[i]
                active[i] = (rec, lat, left - k)
                if left - k == 0:
                    finished.append(i)
            if finished:
                batch.close()
                batch, members = None, ()
                for i in finished:
                    rec, lat, _ = active.pop(i)
                    rec.result = rec.spec.result(lat)
                    rec.finished = time.perf_counter() - t0
                    lat.close()
                    used -= rec.spec.footprint()
        if batch is not None:
            batch.close()
        self.makespan = time.perf_counter() - t0
        return records


def summarize(records: Sequence[JobRecord], makespan: float) -> dict:
    lat = np.array([r.latency for r in records])
    work = sum(r.spec.nodes * r.spec.steps for r in records)
    span = max(r.arrival for r in records) if records else 0.0
    return {"jobs": len(records), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "makespan_s": makespan, "aggregate_mlups": work / makespan / 1e6,
            "offered_mlups": work / span / 1e6 if span > 0 else None,
            "unstable": sum(1 for r in records if not r.result.get("finite", False))}


def accuracy_vs_fp64(spec: JobSpec, result: dict, device: int = 0) -> dict:
    """Rerun `spec` alone with FP64 populations (the FP64 path computes the
    oracle's step bit for bit) and compare: von Karman jobs by the probe trace
    and the Strouhal numbers, boxes by the u_x profile at x = lx/2."""
    ref_spec = JobSpec(**{**spec.__dict__, "storage": "f64"})
    lat = ref_spec.start(device)
    lat.run(ref_spec.relaxation, ref_spec.steps)
    ref = ref_spec.result(lat)
    lat.close()
    out = {"kind": spec.kind, "storage": spec.storage, "steps": spec.steps}
    if spec.kind == "von_karman":
        a, b = result["samples"], ref["samples"]
        out["probe_max_abs_err_uy"] = float(np.abs(a[:, 2] - b[:, 2]).max())
        out["probe_max_abs_uy_fp64"] = float(np.abs(b[:, 2]).max())
        for key in ("strouhal_uy", "strouhal_uy_fft"):
            if key in result and key in ref:
                out[key] = result[key]
                out[key + "_fp64"] = ref[key]
    else:
        a, b = result["ux_mid"], ref["ux_mid"]
        out["profile_max_rel_err"] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    return out


def _main(argv=None):
    """Serving run: `python -m lattice_boltzmann_parallel_solver_amd.serving --jobs 40`
    serves a seeded Poisson job stream and prints p50/p99 time-to-result,
    aggregate MLUPS and, for `--check` jobs of each kind, the error against the
    same job run alone in FP64."""
    import argparse
    import json

    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--jobs", type=int, default=40)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--rate", type=float, default=50.0)
    ap.add_argument("--vk", type=int, nargs=2, default=[8192, 2048])
    ap.add_argument("--square", type=int, nargs=2, default=[128, 2048])
    ap.add_argument("--budget", type=int, nargs=2, default=[5000, 100000])
    ap.add_argument("--storage", default="f32")
    ap.add_argument("--chunk", type=int, default=2000)
    ap.add_argument("--check", type=int, default=0, help="jobs of each kind to check against FP64")
    ap.add_argument("--policy", default="srpt", choices=["fifo", "srpt"])
    ap.add_argument("--fill-nodes", type=int, default=96 << 20, help="srpt: nodes advanced per chunk (at least)")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    arrivals = c5_jobs(a.jobs, seed=a.seed, rate=a.rate, vk_size=tuple(a.vk), square=tuple(a.square),
                       budget=tuple(a.budget), storage=a.storage)
    work = sum(s.nodes * s.steps for _, s in arrivals)
    print(json.dumps({"jobs": a.jobs, "node_steps": work, "von_karman": sum(s.kind == "von_karman" for _, s in arrivals)}),
          flush=True)
    last = [0.0]

    def progress(t, active, waiting, finished):
        if t - last[0] >= 10.0:
            last[0] = t
            print(json.dumps({"t_s": round(t, 1), "active": active, "waiting": waiting, "finished": finished}),
                  flush=True)

    server = Server(chunk_steps=a.chunk, progress=progress, policy=a.policy, fill_nodes=a.fill_nodes)
    recs = server.run(arrivals)
    summary = summarize(recs, server.makespan)
    summary["busy_s"] = server.busy
    summary["config"] = {"jobs": a.jobs, "seed": a.seed, "rate": a.rate, "vk": a.vk, "square": a.square,
                         "budget": a.budget, "storage": a.storage, "chunk": a.chunk, "policy": a.policy,
                         "fill_nodes": a.fill_nodes if a.policy == "srpt" else None}
    print(json.dumps(summary), flush=True)
    jobs = [{"id": r.job_id, "kind": r.spec.kind, "size": [r.spec.lx, r.spec.ly], "steps": r.spec.steps,
             "arrival_s": r.arrival, "admitted_s": r.admitted, "finished_s": r.finished, "latency_s": r.latency,
             "strouhal_uy": r.result.get("strouhal_uy")} for r in recs]
    checks = []
    if a.check:
        for kind in ("von_karman", "couette", "poiseuille"):
            for r in [r for r in recs if r.spec.kind == kind][:a.check]:
                c = accuracy_vs_fp64(r.spec, r.result)
                c["id"] = r.job_id
                checks.append(c)
                print(json.dumps(c), flush=True)
    out = {"summary": summary, "jobs": jobs, "fp64_checks": checks}
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(out, fh, indent=1)


if __name__ == "__main__":
    _main()
