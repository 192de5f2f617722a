"""Benchmark: fused D2Q9 step on the 8192x2048 von Karman channel, FP32 vs FP64.

Workload (the precision configuration of the reference's channel scaled up):
8192x2048 lattice, plate of 400 nodes at x = 2048 centred in y with a +1 node
offset, inlet equilibrium u = (0.1, 0), rho = 1, outlet extrapolation, rigid
top and bottom walls, nu = 0.1*400/100 (Re = 100) so omega = 1/1.7.

value      : MLUPS of the FP32-deviation step (inputs resident in HBM), two steps
             per launch (the row-sweep kernel keeps the intermediate step on chip),
             at the launch shape the library's autotuner picks for this lattice
             (rows per block of the sweep; config.launch_shape), tuned outside
             the timed region; --no-autotune times the default shape.
roofline   : HBM bound; achieved = algorithmic bytes per launch (every population
             read and written once: 18 accesses x 4 B per node) / average launch
             time over the timed region (HIP events on the stream the kernels run
             on); traffic = HBM bytes per launch measured with rocprofv3
             FETCH_SIZE/WRITE_SIZE, copy-calibrated
             (profiles/fused_step_8192x2048_hbm_traffic.json).
             single_step_equivalent_gbs: the bandwidth a one-step-per-launch kernel
             would need for the same MLUPS (72 B per node step); copy_gbs: a
             1 GiB streaming copy (16 bytes per lane) timed the same way, the
             practical ceiling, and frac_of_copy = achieved / copy_gbs.
             row_sweep_pattern_gbs: the two-step kernel's access pattern with no
             work (lbm_row_sweep_probe_ordered: the pull's rows, one row
             prefetched, a barrier per row) at the headline's rows per block,
             in the range-interleaved row order its long runs use: the ceiling
             of a row sweep on this layout, and frac_of_row_sweep_pattern =
             achieved / that; row_sweep_pattern_plain_order_gbs: the same in
             the plain order.
single_step: the same workload with one step per launch (k_fused); its
             periodic_lattice entry times the same kernel on a fully periodic
             8192x2048 lattice (no boundary nodes).
three_steps_per_launch: two against three steps per launch (k_sweepk) on an
             8192x1024 FP32 channel (three steps take rows of up to 1024).
other_workloads_fp64: GPU rates of the shear wave (300x300, its whole 3000
             steps), the 4200x1800 channel and the 2048x2048 Couette and
             Poiseuille boxes in FP64.
fp64, f16  : the same workload with FP64 populations / FP16 deviation storage,
             two steps per launch (FP64 rows of 2048 split into three column
             tiles); fp64_single_step: FP64 with one step per launch.
latency    : p50/p99 time-to-result of the continuous-batching server on a seeded
             Poisson job stream (scaled serving mix, stated in the line).
batching   : aggregate MLUPS of 256 seeded uneven channels advanced together (one
             two-step launch per group of equal row width per step pair, the
             groups on side streams) versus the same channels solved one at a
             time, each at its autotuned launch shape.
f16        : the rate of FP16 deviation storage, with its measured error against
             FP64 (it does not reproduce this channel's shedding).
cpu_baseline: the oracle (numpy FP64) on windows of at least 10 s of the same
             workload: P processes in the reference's block decomposition (value),
             one process, and the C restatement with OpenMP threads; plus the
             300x300 shear wave.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "MLUPS & %HBM roofline at 8192x2048 FP32 vs FP64; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
LX, LY, PLATE = 8192, 2048, 400
U_IN, RE = 0.1, 100.0
BYTES_PER_NODE = {"f64": 144, "f32": 72, "f16": 36}
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "fused_step_8192x2048_hbm_traffic.json")


def measured_traffic(storage, steps_per_launch=1):
    """HBM bytes per launch of the step kernel from the committed rocprofv3
    counter runs (keys "f32", "f64", "f16" for one step per launch,
    "f32_sweep2", ... for the two-step kernel)."""
    key = storage if steps_per_launch == 1 else f"{storage}_sweep2"
    try:
        with open(TRAFFIC_JSON) as fh:
            return float(json.load(fh)[key]["traffic_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def omega():
    nu = U_IN * PLATE / RE
    return 1.0 / (3.0 * nu + 0.5)


def build_lattice(storage, torch, temporal=1):
    """The bench channel, advanced `temporal` steps per launch where the
    two-step kernel applies (falls back to one)."""
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import von_karman_boundary_conditions

    bc = von_karman_boundary_conditions(LX, LY, 1.0, U_IN, PLATE, walls=True, y_offset=1)
    lat = Lattice(LX, LY, boundary=bc, storage=storage, stream=torch.cuda.current_stream().cuda_stream)
    rho = torch.ones((LX, LY), dtype=torch.float64, device="cuda")
    u = torch.zeros((LX, LY, 2), dtype=torch.float64, device="cuda")
    u[..., 0] = U_IN
    lat.set_equilibrium(rho, u)
    del rho, u
    lat.steps_per_launch = 1
    if temporal == 2:
        from lattice_boltzmann_parallel_solver_amd._lib import LbmError

        try:
            lat.set_temporal(2)
            lat.steps_per_launch = 2
        except LbmError:
            pass
    return lat


def copy_gbs(torch, nbytes=1 << 30, reps=10):
    """HBM bandwidth of the library's streaming copy (16 bytes per lane) over
    nbytes, read + write counted, timed with events on the current stream:
    the practical ceiling a streaming kernel is measured against beside the
    8 TB/s peak.  (torch's copy_ goes through the runtime's blit kernel,
    which reaches less than the step kernels do.)"""
    import ctypes

    from lattice_boltzmann_parallel_solver_amd import _lib

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream().cuda_stream
    L = _lib.load()

    def once():
        _lib.check(L.lbm_copy_probe(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), nbytes,
                                    ctypes.c_void_p(stream)))

    once()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        once()
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end)
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes * reps / (ms / 1e3) / 1e9


def row_sweep_gbs(torch, rows=32, reps=10, interleaved=True):
    """HBM bandwidth of the two-step kernel's access pattern with no work
    (the bench lattice's shape: nine planes of 2048 FP32 values per row,
    `rows` rows per block, each plane read from the row the pull reads it
    from, one row prefetched, a barrier per row), read + write counted, with
    the rows in the range-interleaved order of long two-step runs or in the
    plain order: the ceiling a row sweep reaches on this layout."""
    import ctypes

    from lattice_boltzmann_parallel_solver_amd import _lib

    pitch = LY
    a = torch.zeros(9 * LX * pitch, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream().cuda_stream
    L = _lib.load()

    def once():
        mode = 2 | (1 if interleaved else 0)  # the pull's rows, in the chosen row order
        _lib.check(L.lbm_row_sweep_probe_ordered(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), LX,
                                                 pitch, rows, mode, ctypes.c_void_p(stream)))

    once()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        once()
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end)
    del a, b
    torch.cuda.empty_cache()
    return 2 * 9 * LX * pitch * 4 * reps / (ms / 1e3) / 1e9


def three_step_leg(torch, steps=240, ly=1024):
    """Two against three steps per launch on a von Karman channel of
    LX x 1024 (plate ly/5 at x = LX/4, +1 y offset, walls, inlet U_IN, Re
    RE): three steps per launch apply to rows of up to 1024 columns."""
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd.boundary_utils import von_karman_boundary_conditions

    plate = ly // 5
    bc = von_karman_boundary_conditions(LX, ly, 1.0, U_IN, plate, walls=True, y_offset=1)
    lat = Lattice(LX, ly, boundary=bc, storage="f32", stream=torch.cuda.current_stream().cuda_stream)
    rho = torch.ones((LX, ly), dtype=torch.float64, device="cuda")
    u = torch.zeros((LX, ly, 2), dtype=torch.float64, device="cuda")
    u[..., 0] = U_IN
    lat.set_equilibrium(rho, u)
    del rho, u
    om = 1.0 / (3.0 * U_IN * plate / RE + 0.5)
    out = {"workload": f"von Karman channel {LX}x{ly}, plate {plate}, walls, f32"}
    for k in (2, 3):
        lat.set_temporal(k)
        lat.run(om, 24)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        lat.run(om, steps)
        end.record()
        torch.cuda.synchronize()
        sec = start.elapsed_time(end) / 1e3
        out[f"mlups_{k}_steps_per_launch"] = LX * ly * steps / sec / 1e6
        out[f"roofline_frac_{k}_steps_per_launch"] = 72 * LX * ly / (sec / (steps / k)) / 1e9 / HBM_PEAK_GBS
    lat.close()
    torch.cuda.empty_cache()
    return out


def config_legs(torch):
    """Rates of the other workloads on the GPU, FP64 (the oracle's arithmetic
    bit for bit), each lattice resident, with the launch mode it supports
    fastest (two steps per launch except the pressure-driven box):
    the 300x300 shear wave (omega 1, epsilon 0.05) over its whole 3000-step
    run, the 4200x1800 von Karman channel (plate 400 at x = 1050, +1 y offset,
    walls, inlet 0.1, Re 100), the 2048x2048 Couette box (lid 0.05, omega 1)
    and the 2048x2048 Poiseuille box (delta p 1.5e-3, omega 1.5)."""
    from lattice_boltzmann_parallel_solver_amd import Lattice
    from lattice_boltzmann_parallel_solver_amd import boundary_utils as BU
    from lattice_boltzmann_parallel_solver_amd.initial_values import sinusoidal_velocity_x

    def timed(lat, om, steps):
        lat.run(om, 4)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        lat.run(om, steps)
        end.record()
        torch.cuda.synchronize()
        return start.elapsed_time(end) / 1e3

    def equilibrium_at_rest(lat, u_x=0.0):
        rho = torch.ones((lat.lx, lat.ly), dtype=torch.float64, device="cuda")
        u = torch.zeros((lat.lx, lat.ly, 2), dtype=torch.float64, device="cuda")
        u[..., 0] = u_x
        lat.set_equilibrium(rho, u)

    out = {}
    stream = torch.cuda.current_stream().cuda_stream
    lat = Lattice(300, 300, storage="f64", stream=stream)
    rho, u = sinusoidal_velocity_x((300, 300), 0.05)
    lat.set_equilibrium(rho, u)
    lat.set_temporal(2)
    sec = timed(lat, 1.0, 3000)
    out["shear_wave_300x300"] = {"steps": 3000, "seconds": sec, "mlups": 90000 * 3000 / sec / 1e6,
                                 "steps_per_launch": 2, "storage": "f64"}
    lat.close()
    cases = (("channel_4200x1800", 4200, 1800,
              BU.von_karman_boundary_conditions(4200, 1800, 1.0, U_IN, 400, walls=True, y_offset=1),
              1.0 / (3.0 * U_IN * 400 / RE + 0.5), U_IN, 2),
             ("couette_2048x2048", 2048, 2048, BU.couette_flow_boundary_conditions(2048, 2048, 0.05, 1.0), 1.0, 0.0, 2),
             ("poiseuille_2048x2048", 2048, 2048,
              BU.poiseuille_flow_boundary_conditions(2048, 2048, (1 + 1.5e-3) / 3, (1 - 1.5e-3) / 3), 1.5, 0.0, 1))
    for name, lx, ly, bc, om, u_x, spl in cases:
        lat = Lattice(lx, ly, boundary=bc, storage="f64", stream=stream)
        equilibrium_at_rest(lat, u_x)
        lat.set_temporal(spl)
        sec = timed(lat, om, 200)
        out[name] = {"steps": 200, "seconds": sec, "mlups": lx * ly * 200 / sec / 1e6, "steps_per_launch": spl,
                     "storage": "f64"}
        lat.close()
    torch.cuda.empty_cache()
    return out


def periodic_single_step_leg(torch, storage="f32", steps=300):
    """The one-step kernel on a fully periodic 8192x2048 lattice (no boundary
    nodes; uniform flow u = (0.05, 0)), the same launch shape: the fused
    stream+collide step alone, beside the channel's legs."""
    from lattice_boltzmann_parallel_solver_amd import Lattice

    lat = Lattice(LX, LY, storage=storage, stream=torch.cuda.current_stream().cuda_stream)
    rho = torch.ones((LX, LY), dtype=torch.float64, device="cuda")
    u = torch.zeros((LX, LY, 2), dtype=torch.float64, device="cuda")
    u[..., 0] = 0.05
    lat.set_equilibrium(rho, u)
    del rho, u
    ms = time_steps(lat, steps, 30, torch)
    lat.close()
    torch.cuda.empty_cache()
    ach = BYTES_PER_NODE[storage] * LX * LY / (ms / 1e3 / steps) / 1e9
    return {"mlups": LX * LY * steps / (ms / 1e3) / 1e6, "ms_per_step": ms / steps, "steps": steps,
            "roofline_achieved_gbs": ach, "roofline_frac": ach / HBM_PEAK_GBS,
            "workload": f"periodic {LX}x{LY}, {storage}, uniform flow, one step per launch"}


def time_steps(lat, steps, warmup, torch):
    lat.run(omega(), warmup)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    lat.run(omega(), steps)
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end)  # ms


def serving_leg(n=24, seed=0):
    """Continuous batching on a seeded Poisson stream (50 jobs/s): 60 % von
    Karman 2048x512 channels, 20 % Couette, 20 % Poiseuille on squares of side
    U[128, 1024], budgets log-uniform in [5k, 20k] steps, FP32."""
    from lattice_boltzmann_parallel_solver_amd.serving import Server, c5_jobs, summarize

    arrivals = c5_jobs(n, seed=seed, vk_size=(2048, 512), square=(128, 1024), budget=(5000, 20000))
    server = Server(chunk_steps=2000)
    recs = server.run(arrivals)
    s = summarize(recs, server.makespan)
    s["workload"] = (f"{n} jobs, Poisson 50/s, 60% von Karman 2048x512, 20% Couette, 20% Poiseuille "
                     "squares U[128,1024], budgets log-uniform [5k,20k], f32, seed 0, "
                     f"policy {server.policy}")
    return s


def batching_leg(n=256, steps=1000, seed=0):
    """The 256 seeded uneven channels of the batching workload (nx U[200,1200],
    ny = 3nx/7, blockage U[0.1,0.4], Re U[40,200], FP32), 1000 steps: one
    batch (three steps per launch for the groups of 385-512 columns, two for
    the others) versus the same channels solved one at a time, each at the
    autotuned launch shape of its size bucket."""
    from lattice_boltzmann_parallel_solver_amd import workloads as W

    jobs = W.channel_jobs(n, seed=seed)
    _, batched, _ = W.run_batched(jobs, steps, "f32", temporal=3)
    _, sequential = W.run_sequential(jobs, steps, "f32", tune=True)
    return {"jobs": n, "steps": steps, "nodes": sum(j.nodes for j in jobs), "batched_mlups": batched,
            "sequential_mlups": sequential, "speedup": batched / sequential,
            "workload": f"{n} channels, seed {seed}, {steps} steps, f32; full-length run (20k steps) in "
                        "profiles/batching_c4_256_channels.json"}


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo where present)."""
    import platform

    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(steps=4, mp_steps=30, c_steps=120, c1_numpy_steps=3000, c1_steps=3000):
    """The reference's numpy path restated (the oracle, FP64) on this host's
    cores, each leg sized for a window of at least 10 s at the rates measured
    on the GPU machines' hosts (channel: about 5 MLUPS on one core, 40 MLUPS on
    16 processes, 180 MLUPS in C with 16 threads):

    value          : P processes with the reference's MPI block decomposition
                     (oracle/lbm_oracle_mp.py; P = host cores, at most 16),
                     mp_steps steps of the 8192x2048 channel
    single_process : the numpy oracle on one core, `steps` steps of the channel
    c_openmp       : the C restatement of the oracle with OpenMP threads,
                     c_steps steps of the channel
    shear_wave_300 : the 300x300 shear wave (omega 1): the numpy oracle on one
                     core over c1_numpy_steps steps, and the C restatement over
                     the whole c1_steps-step run
    MLUPS = nodes x steps / seconds / 1e6 (the reference's own timer,
    experiments.py:765-772, divides nanoseconds by 10e9, which makes its stored
    numbers 10x LUPS; here seconds are seconds).

    Runs before anything initialises the GPU (the multi-process leg forks)."""
    import numpy as np

    from oracle import lbm_oracle as O
    from oracle import lbm_oracle_mp as MP
    from oracle import lbm_oracle_native as NAT

    rho0, u0 = O.density_1_velocity_x_u0_velocity_y_0_initial((LX, LY), U_IN)
    f0 = O.equilibrium(rho0, u0)
    bc = O.von_karman_boundary_conditions(LX, LY, 1.0, U_IN, PLATE, walls=True, y_offset=1)
    N = LX * LY

    f, rho, u = O.step(f0, rho0, u0, omega(), bc)
    t0 = time.perf_counter_ns()
    for _ in range(steps):
        f, rho, u = O.step(f, rho, u, omega(), bc)
    dt1 = (time.perf_counter_ns() - t0) / 1e9
    assert np.isfinite(rho).all()
    del f, rho, u

    procs = MP.default_processes()
    _, rho_mp, _, dtp, grid = MP.run(f0, rho0, u0, omega(), mp_steps, bc, processes=procs, warmup=1)
    assert np.isfinite(rho_mp).all()
    del rho_mp

    NAT.run(f0, rho0, u0, omega(), 1, bc)  # build/load, first touch
    t0 = time.perf_counter()
    NAT.run(f0, rho0, u0, omega(), 1, bc)
    t1 = time.perf_counter()
    _, rho_c, _, _ = NAT.run(f0, rho0, u0, omega(), 1 + c_steps, bc)
    t2 = time.perf_counter()
    assert np.isfinite(rho_c).all()
    dtc = (t2 - t1) - (t1 - t0)
    threads = int(NAT.lib().oracle_threads())
    del f0, rho0, u0, rho_c

    # the 300x300 shear wave
    rs, us = O.sinusoidal_velocity_x((300, 300), 0.05)
    fs = O.equilibrium(rs, us)
    f, rho, u = O.step(fs, rs, us, 1.0)
    t0 = time.perf_counter()
    for _ in range(c1_numpy_steps):
        f, rho, u = O.step(f, rho, u, 1.0)
    dts = time.perf_counter() - t0
    t0 = time.perf_counter()
    NAT.run(fs, rs, us, 1.0, c1_steps)
    dtsc = time.perf_counter() - t0

    host = f"{cpu_model()}, numpy {np.__version__}"
    return {"value": N * mp_steps / dtp / 1e6, "unit": "MLUPS", "cores": procs, "kind": "port",
            "sample": f"{mp_steps} steps (after 1 warm-up) of the {LX}x{LY} FP64 channel with the numpy oracle "
                      f"on {procs} processes in the reference's {grid[0]}x{grid[1]} block decomposition, "
                      f"halos through shared memory ({dtp:.1f} s); host {host}",
            "cpu_model": cpu_model(), "numpy": np.__version__,
            "single_process": {"value": N * steps / dt1 / 1e6, "unit": "MLUPS", "cores": 1,
                               "sample": f"{steps} steps of the channel after 1 warm-up ({dt1:.1f} s)"},
            "c_openmp": {"value": N * c_steps / dtc / 1e6, "unit": "MLUPS", "cores": threads,
                         "sample": f"{c_steps} steps of the channel with the C restatement of the oracle, OpenMP "
                                   f"({dtc:.1f} s, set-up subtracted by a 1-step run)"},
            "shear_wave_300": {"numpy_mlups": 90000 * c1_numpy_steps / dts / 1e6, "numpy_cores": 1,
                               "numpy_sample": f"{c1_numpy_steps} steps after 1 warm-up ({dts:.1f} s)",
                               "numpy_projected_3000_steps_s": dts / c1_numpy_steps * 3000,
                               "c_openmp_mlups": 90000 * c1_steps / dtsc / 1e6, "c_openmp_cores": threads,
                               "c_openmp_sample": f"the whole {c1_steps}-step run ({dtsc:.1f} s)"}}


def f16_error_note():
    """The FP16 storage's measured error against the FP64 run of the same
    channel (the committed precision study)."""
    path = os.path.join(ROOT, "profiles", "precision_study_8192x2048.json")
    try:
        with open(path) as fh:
            ps = json.load(fh)
        cps = ps["checkpoints"]
        out = {f"l2_u_rel_{k}": v["f16"]["l2_u"] for k, v in cps.items()}
        out.update({f"l2_rho_rel_{k}": v["f16"]["l2_rho"] for k, v in cps.items()})
        out["strouhal_uy_fp64"] = ps.get("strouhal", {}).get("f64", {}).get("strouhal_uy_interp")
        out["strouhal_uy"] = None  # no shedding: no frequency to report
    except (OSError, KeyError, ValueError):
        out = {}
    out["note"] = ("FP16 deviation storage does not reproduce this channel: the per-step change of a population "
                   "(about 1e-6) is below the FP16 spacing of the deviations (1.5e-5), so the wake stagnates and "
                   "no vortices shed; a rate for the storage, not a valid result for this configuration")
    out["source"] = "profiles/precision_study_8192x2048.json"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "f64"])
    ap.add_argument("--single-steps", type=int, default=200, help="steps of the one-step-per-launch leg")
    ap.add_argument("--fp64-steps", type=int, default=200)
    ap.add_argument("--f16-steps", type=int, default=200)
    ap.add_argument("--temporal", type=int, default=2, choices=[1, 2],
                    help="steps per kernel launch (2: row-sweep kernel where it applies)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-batching", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the headline lattice at the default launch shape instead of the autotuned one")
    ap.add_argument("--rows", type=int, default=0,
                    help="rows per block of the headline's two-step sweep (implies --no-autotune)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the three-step leg and the other workloads' rates (profiling the main legs)")
    ap.add_argument("--three-step-only", action="store_true",
                    help="only the two- against three-step leg on the 8192x1024 channel (profiling k_sweepk)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 3 device-to-device copies of BYTES (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the lattice is not split across devices)", file=sys.stderr)
        sys.exit(2)
    cpu = None if (args.no_cpu_baseline or args.copy_calibration or args.three_step_only) else cpu_baseline()

    import torch

    if not torch.cuda.is_available():
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.three_step_only:
        print(json.dumps({"three_steps_per_launch": three_step_leg(torch)}))
        return
    if args.copy_calibration:
        a = torch.empty(args.copy_calibration // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": a.numel() * 4, "copies": 3}))
        return
    N = LX * LY

    lat = build_lattice(args.storage, torch, args.temporal)
    tuned = None
    if args.rows > 0 and lat.steps_per_launch == 2:
        lat.set_temporal(2, args.rows)
        tuned = {"steps_per_launch": 2, "rows_per_block": args.rows}
    elif not args.no_autotune and args.temporal == 2:
        # the library's autotuner (launch shape, steps per launch and rows per
        # block timed on a scratch lattice of the same shape), outside the timed region
        from lattice_boltzmann_parallel_solver_amd import workloads as W

        # (trials of 200 steps: long enough for the interleaved row order)
        best = W.autotune_lattice(lat, omega(), steps=200)
        lat.steps_per_launch = best[3]
        tuned = {"nodes_per_thread": best[0], "xcd_order": best[1], "nt_stores": best[2],
                 "steps_per_launch": best[3], "rows_per_block": best[4]}
        torch.cuda.empty_cache()
    spl = lat.steps_per_launch
    steps = -(-args.steps // spl) * spl  # whole launches
    ms = time_steps(lat, steps, args.warmup, torch)
    m, jx, _ = lat.moments_sum()
    assert m == m and abs(m / N - 1.0) < 0.1, "unstable run"
    lat.close()
    del lat
    torch.cuda.empty_cache()
    sec = ms / 1e3
    mlups = N * steps / sec / 1e6
    launch_s = sec / (steps / spl)
    # one launch reads and writes every population once, whatever its step count
    bytes_launch = BYTES_PER_NODE[args.storage] * N
    achieved = bytes_launch / launch_s / 1e9

    def leg(storage, steps, temporal):
        if steps <= 0:
            return None
        steps += steps % 2
        lat = build_lattice(storage, torch, temporal)
        spl_ = lat.steps_per_launch
        ms_ = time_steps(lat, steps, max(20, args.warmup), torch)
        lat.close()
        torch.cuda.empty_cache()
        ach = BYTES_PER_NODE[storage] * N / (ms_ / 1e3 / (steps / spl_)) / 1e9
        return {"mlups": N * steps / (ms_ / 1e3) / 1e6, "ms_per_step": ms_ / steps, "steps": steps,
                "steps_per_launch": spl_, "roofline_achieved_gbs": ach, "roofline_frac": ach / HBM_PEAK_GBS,
                "bytes_per_node_step": BYTES_PER_NODE[storage] // spl_, "traffic": measured_traffic(storage, spl_)}

    single = leg(args.storage, args.single_steps, 1) if spl == 2 else None
    if single is not None and not args.no_extra:
        single["periodic_lattice"] = periodic_single_step_leg(torch, args.storage)
    fp64 = leg("f64", args.fp64_steps, args.temporal) if args.storage != "f64" else None
    fp64_single = leg("f64", args.fp64_steps, 1) if args.storage != "f64" and args.temporal == 2 else None
    f16 = leg("f16", args.f16_steps, args.temporal) if args.storage != "f16" else None
    if f16 is not None:
        f16["error_vs_fp64"] = f16_error_note()

    copy = copy_gbs(torch)
    # the headline's rows per block (automatic: 32 on 256 CUs), in the row
    # order its long two-step run uses and in the plain one
    sweep_rows = (tuned or {}).get("rows_per_block") or 32
    sweep_pattern = row_sweep_gbs(torch, rows=sweep_rows, interleaved=True)
    sweep_pattern_plain = row_sweep_gbs(torch, rows=sweep_rows, interleaved=False)
    three = None if args.no_extra else three_step_leg(torch)
    configs = None if args.no_extra else config_legs(torch)
    latency = None if args.no_serving else serving_leg()
    batching = None if args.no_batching else batching_leg()

    line = {
        "metric": METRIC,
        "value": mlups,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32", "f16": "f16-storage/f32", "f64": "f64"}[args.storage],
        "data": "synthetic",
        "config": {"workload": f"von Karman channel {LX}x{LY}, plate {PLATE} at x={LX // 4} (+1 y offset), "
                               f"rigid walls, inlet u={U_IN}, Re={RE:g}, {args.storage} population storage",
                   "lx": LX, "ly": LY, "storage": args.storage, "omega": omega(), "steps_per_launch": spl,
                   "launch_shape": tuned or "default",
                   "parallelism": "single GPU (whole lattice resident)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.storage, spl),
                     "kernel": "k_sweep2 (two steps per launch)" if spl == 2 else "k_fused (one step per launch)",
                     "bytes_per_node_launch": BYTES_PER_NODE[args.storage],
                     "bytes_per_node_step": BYTES_PER_NODE[args.storage] // spl,
                     "single_step_equivalent_gbs": mlups * 1e6 * BYTES_PER_NODE[args.storage] / 1e9,
                     "copy_gbs": copy, "frac_of_copy": achieved / copy,
                     "row_sweep_pattern_gbs": sweep_pattern, "frac_of_row_sweep_pattern": achieved / sweep_pattern,
                     "row_sweep_pattern_plain_order_gbs": sweep_pattern_plain},
        "single_step": single,
        "three_steps_per_launch": three,
        "other_workloads_fp64": configs,
        "fp64": fp64,
        "fp64_single_step": fp64_single,
        "f16": f16,
        "latency": latency,
        "batching": batching,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
