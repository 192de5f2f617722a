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
import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import LbmError, LbmUnstable
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

    def strouhal_window(self) -> int:
        """Steps in two shedding periods of a plate of size d at inlet speed
        u_in, taking St = f d / u = 0.2 (the u_y shedding frequency of these
        channels, 0.21-0.22 measured): 2 d / (0.2 u) = 10 d / u.  Strouhal
        numbers are reported only from at least this many samples."""
        return int(math.ceil(10 * self.plate / self.u_in))

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
            if len(samples) - start >= self.strouhal_window():
                out["strouhal_uy"] = strouhal_interpolated(samples[:, 2], self.plate, self.u_in, start)
                out["strouhal_speed"] = strouhal_interpolated(speed(samples), self.plate, self.u_in, start)
                out["strouhal_uy_fft"] = strouhal_reference(samples[:, 2], self.plate, self.u_in, start)[0]
            else:
                out["strouhal_note"] = (f"not reported: the analysed second half ({len(samples) - start} steps) is "
                                        f"shorter than two shedding periods ({self.strouhal_window()} steps)")
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
    Results do not depend on the policy (a job's steps are the same in any
    batch).  Under srpt a long job can wait while shorter ones keep arriving;
    a finite stream always drains.

    A chunk runs on a worker thread; meanwhile the server keeps admitting
    arrivals whose footprint fits, so their lattices are allocated and
    initialised (on their own streams) while the GPU advances the others.

    After every chunk each job that advanced is checked: FP64 sums of mass and
    momentum over its state (the reference's only guard is the assertion
    0 < omega < 2, lattice_boltzmann_method.py:212; an unstable BGK run grows
    without bound).  A job whose sums are not finite, whose mass moved more than
    `mass_tolerance` from its initial value (rho = 1 on every node) or whose
    mean velocity |j|/m passed `max_mean_speed` (beyond 0.5 the lattice's
    velocities are far past the sound speed 1/sqrt(3)) is marked failed at
    once: its result records why and after how many steps, and its memory and
    batch slot are freed for the next job."""

    def __init__(self, hbm_budget: int = 160 << 30, chunk_steps: int = 2000, max_active: int = 256,
                 device: int = 0, nodes_per_thread: int = 1, temporal: int = 3, progress=None,
                 policy: str = "srpt", fill_nodes: int = 96 << 20, on_result=None, mass_tolerance: float = 0.1,
                 max_mean_speed: float = 0.5, poll_s: float = 0.002):
        if policy not in ("fifo", "srpt"):
            raise ValueError(f"unknown policy {policy!r}")
        self.policy = policy
        self.fill_nodes = int(fill_nodes)
        self.hbm_budget = int(hbm_budget)
        self.chunk_steps = int(chunk_steps)
        self.max_active = int(max_active)
        self.device = device
        self.nodes_per_thread = nodes_per_thread
        # steps per launch of the batched step: 2, the sweep kernel where it
        # applies; 3 (default), also the three-step kernel for the FP32 widths
        # where it pays in large batches
        self.temporal = temporal
        self.progress = progress  # called after every chunk with (seconds, active, waiting, finished)
        # called as on_result(record, lattice) when a job finishes, before its
        # lattice is freed (e.g. to read the final fields); not called for
        # failed jobs
        self.on_result = on_result
        self.mass_tolerance = float(mass_tolerance)
        self.max_mean_speed = float(max_mean_speed)
        self.poll_s = float(poll_s)
        # bring up the HIP runtime and load the kernels before the first job arrives
        warm = Lattice(8, 8, device=device)
        warm.set_equilibrium(np.ones((8, 8)), np.zeros((8, 8, 2)))
        warm.run(1.0, 2)
        warm.synchronize()
        warm.close()

    def health(self, spec: JobSpec, lat: Lattice) -> Optional[str]:
        """None if the job's state is sound, else the reason it failed."""
        try:
            m, jx, jy = lat.moments_sum()
        except LbmUnstable:
            return "non-finite mass or momentum"
        if abs(m / spec.nodes - 1.0) > self.mass_tolerance:
            return f"mass {m / spec.nodes:.6g} per node, more than {self.mass_tolerance} from 1"
        if m <= 0 or math.hypot(jx, jy) / m > self.max_mean_speed:
            return f"mean velocity {math.hypot(jx, jy) / m if m > 0 else math.inf:.6g} above {self.max_mean_speed}"
        return None

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
        self.admitted_during_chunks = 0

        def arrive_and_admit():
            nonlocal waiting, used
            now = time.perf_counter() - t0
            while incoming and incoming[0].arrival <= now:
                waiting.append(incoming.popleft())
            # admission while the footprint fits: arrival order (fifo) or least work first (srpt)
            if self.policy == "srpt" and len(waiting) > 1:
                waiting = deque(sorted(waiting, key=lambda r: (r.spec.nodes * r.spec.steps, r.job_id)))
            n = 0
            while waiting and len(active) < self.max_active and used + waiting[0].spec.footprint() <= self.hbm_budget:
                rec = waiting.popleft()
                lat = rec.spec.start(self.device)
                rec.admitted = time.perf_counter() - t0
                active[rec.job_id] = (rec, lat, rec.spec.steps)
                used += rec.spec.footprint()
                n += 1
            return n

        def retire(i, result):
            nonlocal used
            rec, lat, _ = active.pop(i)
            rec.result = result
            rec.finished = time.perf_counter() - t0
            if self.on_result is not None and not result.get("failed"):
                self.on_result(rec, lat)
            lat.close()
            used -= rec.spec.footprint()

        while incoming or waiting or active:
            arrive_and_admit()
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
            # the chunk on a worker thread (the library call releases the
            # interpreter); arrivals are admitted meanwhile
            err = []

            def chunk(b=batch, k=k):
                try:
                    b.run(k)
                except BaseException as e:  # re-raised below, on the serving thread
                    err.append(e)

            tb = time.perf_counter()
            worker = threading.Thread(target=chunk, name="lbm-chunk")
            worker.start()
            while worker.is_alive():
                worker.join(self.poll_s)
                if worker.is_alive():
                    self.admitted_during_chunks += arrive_and_admit()
            self.busy += time.perf_counter() - tb
            if err:
                raise err[0]
            changed = False
            for i in ids:
                rec, lat, left = active[i]
                left -= k
                active[i] = (rec, lat, left)
                why = self.health(rec.spec, lat)
                if why is not None:
                    retire(i, {"finite": False, "failed": True, "reason": why,
                               "failed_after_steps": rec.spec.steps - left})
                    changed = True
                elif left == 0:
                    retire(i, rec.spec.result(lat))
                    changed = True
            if changed:
                batch.close()
                batch, members = None, ()
            if self.progress is not None:
                self.progress(time.perf_counter() - t0, len(active), len(waiting) + len(incoming),
                              sum(1 for r in records if not math.isnan(r.finished)))
        if batch is not None:
            batch.close()
        self.makespan = time.perf_counter() - t0
        return records


def summarize(records: Sequence[JobRecord], makespan: float) -> dict:
    """p50/p99 time-to-result over the jobs that completed, aggregate rate of
    the work they completed, and the count of failed jobs."""
    done = [r for r in records if not r.result.get("failed")]
    lat = np.array([r.latency for r in done]) if done else np.array([math.nan])
    work = sum(r.spec.nodes * r.spec.steps for r in done)
    span = max(r.arrival for r in records) if records else 0.0
    offered = sum(r.spec.nodes * r.spec.steps for r in records)
    failed = sum(1 for r in records if r.result.get("failed"))
    return {"jobs": len(records), "completed": len(done), "failed": failed,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "makespan_s": makespan, "aggregate_mlups": work / makespan / 1e6,
            "offered_mlups": offered / span / 1e6 if span > 0 else None,
            "unstable": failed}


def fp64_results(specs: Sequence[JobSpec], device: int = 0, **server_kw) -> List[dict]:
    """Every job run again with FP64 populations (the FP64 path computes the
    oracle's step bit for bit), all of them served together as one batch
    stream; returns their results in order."""
    twins = [JobSpec(**{**s.__dict__, "storage": "f64"}) for s in specs]
    server = Server(device=device, policy="fifo", **server_kw)
    recs = server.run([(0.0, t) for t in twins], realtime=False)
    return [r.result for r in recs]


def compare_to_fp64(spec: JobSpec, result: dict, ref: dict) -> dict:
    """Error of a job's result against its FP64 twin: von Karman jobs by the
    probe's u_y trace and the Strouhal numbers (where both report them), boxes
    by the u_x profile at x = lx/2."""
    out = {"kind": spec.kind, "storage": spec.storage, "steps": spec.steps}
    if result.get("failed") or ref.get("failed"):
        out["failed"] = True
        out["fp64_failed"] = bool(ref.get("failed"))
        return out
    if spec.kind == "von_karman":
        a, b = result["samples"], ref["samples"]
        out["probe_max_abs_err_uy"] = float(np.abs(a[:, 2] - b[:, 2]).max())
        out["probe_max_abs_uy_fp64"] = float(np.abs(b[:, 2]).max())
        for key in ("strouhal_uy", "strouhal_uy_fft"):
            if key in result and key in ref:
                out[key] = result[key]
                out[key + "_fp64"] = ref[key]
        if "strouhal_uy" in result and "strouhal_uy" in ref:
            out["strouhal_uy_rel_err"] = result["strouhal_uy"] / ref["strouhal_uy"] - 1.0
    else:
        a, b = result["ux_mid"], ref["ux_mid"]
        out["profile_max_rel_err"] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    return out


def accuracy_vs_fp64(spec: JobSpec, result: dict, device: int = 0) -> dict:
    """compare_to_fp64 against the job rerun alone in FP64."""
    return compare_to_fp64(spec, result, fp64_results([spec], device)[0])


def _main(argv=None):
    """Serving run: `python -m lattice_boltzmann_parallel_solver_amd.serving --jobs 40`
    serves a seeded Poisson job stream and prints p50/p99 time-to-result,
    aggregate MLUPS, the failed-job count and, for `--check` jobs of each kind
    (`--check-all`: every job), the error against the job's FP64 twin."""
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
    ap.add_argument("--check-all", action="store_true", help="check every job against its FP64 twin")
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
    summary["admitted_during_chunks"] = server.admitted_during_chunks
    summary["config"] = {"jobs": a.jobs, "seed": a.seed, "rate": a.rate, "vk": a.vk, "square": a.square,
                         "budget": a.budget, "storage": a.storage, "chunk": a.chunk, "policy": a.policy,
                         "fill_nodes": a.fill_nodes if a.policy == "srpt" else None}
    print(json.dumps(summary), flush=True)
    jobs = [{"id": r.job_id, "kind": r.spec.kind, "size": [r.spec.lx, r.spec.ly], "steps": r.spec.steps,
             "arrival_s": r.arrival, "admitted_s": r.admitted, "finished_s": r.finished, "latency_s": r.latency,
             "strouhal_uy": r.result.get("strouhal_uy"), "strouhal_note": r.result.get("strouhal_note"),
             "failed": r.result.get("failed", False), "reason": r.result.get("reason")} for r in recs]
    checks = []
    if a.check or a.check_all:
        if a.check_all:
            pick = list(recs)
        else:
            pick = [r for kind in ("von_karman", "couette", "poiseuille")
                    for r in [r for r in recs if r.spec.kind == kind][:a.check]]
        last[0] = 0.0

        def progress64(t, active, waiting, finished):
            if t - last[0] >= 10.0:
                last[0] = t
                print(json.dumps({"fp64_twins_t_s": round(t, 1), "active": active, "finished": finished}),
                      flush=True)

        refs = fp64_results([r.spec for r in pick], progress=progress64)
        for r, ref in zip(pick, refs):
            c = compare_to_fp64(r.spec, r.result, ref)
            c["id"] = r.job_id
            checks.append(c)
            print(json.dumps(c), flush=True)
    if checks:
        ok = [c for c in checks if not c.get("failed")]

        def worst(key):
            vals = [abs(c[key]) for c in ok if c.get(key) is not None]
            return max(vals) if vals else None

        summary["fp64_error"] = {"jobs_checked": len(checks), "probe_max_abs_err_uy": worst("probe_max_abs_err_uy"),
                                 "profile_max_rel_err": worst("profile_max_rel_err"),
                                 "strouhal_uy_max_abs_rel_err": worst("strouhal_uy_rel_err")}
        print(json.dumps({"fp64_error": summary["fp64_error"]}), flush=True)
    out = {"summary": summary, "jobs": jobs, "fp64_checks": checks}
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(out, fh, indent=1)


if __name__ == "__main__":
    _main()
