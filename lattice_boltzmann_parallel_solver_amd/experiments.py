"""Experiment drivers of the reference, on the device-resident lattice.

Counterparts of src/experiments.py and src/visualizations_utils.py without
the plotting: each function runs the reference's scenario with the fused GPU
step and returns the numbers the reference plotted or wrote to CSV.

  shear_wave_viscosity     plot_measured_viscosity_vs_omega, velocity variant
                           (experiments.py:147-223)
  couette_profile          plot_couette_flow_vel_vectors (experiments.py:304-374)
  poiseuille_profile       plot_poiseuille_flow_vel_vectors (experiments.py:377-504)
  von_karman_probe         x_strouhal (experiments.py:650-720), optionally with the
                           rigid walls and plate offset of the large channel runs
  strouhal_reference       the FFT peak procedure (visualizations_utils.py:149-167)
  strouhal_interpolated,
  strouhal_zero_crossing   sub-bin estimators (the reference's peak is quantised
                           to one FFT bin, d/(u N) in Strouhal units)
  step_rates               node rate with one and with two steps per launch on a
                           lattice of any size (no counterpart: a measurement)
"""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from .boundary_utils import (couette_flow_boundary_conditions, poiseuille_flow_boundary_conditions,
                             von_karman_boundary_conditions)
from .initial_values import (density_1_velocity_0_initial, density_1_velocity_x_u0_velocity_y_0_initial,
                             sinusoidal_velocity_x)
from ._lib import LbmError
from .solver import Lattice


def omega_from_viscosity(nu: float) -> float:
    """omega = 1/(3 nu + 1/2) (experiments.py:672)."""
    return float(np.reciprocal(3 * nu + 0.5))


def viscosity_from_omega(omega: float) -> float:
    """nu = (1/omega - 1/2)/3 (experiments.py:210)."""
    return (1 / 3) * (1 / omega - 0.5)


def speed(samples: np.ndarray) -> np.ndarray:
    """|u| of probe rows (rho, ux, uy) as sqrt(ux^2 + uy^2)."""
    ux, uy = samples[:, 1], samples[:, 2]
    return np.sqrt(ux * ux + uy * uy)


# ----------------------------------------------------------------------------- Strouhal
def strouhal_reference(trace, L: float, u: float, start: int = 0) -> Tuple[float, float, float]:
    """(St, peak frequency, bin width in St) by the reference's procedure:
    remove the mean of trace[start:], take |FFT| and its argmax
    (visualizations_utils.py:157-167)."""
    v = np.array(trace[start:], dtype=np.float64)
    v -= np.mean(v)
    yf = np.fft.fft(v)
    freq = np.fft.fftfreq(len(v), 1)
    f = float(np.abs(freq[np.argmax(np.abs(yf))]))
    return f * L / u, f, L / (u * len(v))


def strouhal_interpolated(trace, L: float, u: float, start: int = 0) -> float:
    """St from a Hann-windowed spectrum with the peak located by a parabola
    through the log magnitudes of the three bins around it (quadratic
    interpolation of the Gaussian-like Hann main lobe)."""
    v = np.array(trace[start:], dtype=np.float64)
    v -= np.mean(v)
    n = len(v)
    mag = np.abs(np.fft.rfft(v * np.hanning(n)))
    k = int(np.argmax(mag[1:])) + 1
    if 1 <= k < len(mag) - 1:
        a, b, c = np.log(mag[k - 1] + 1e-300), np.log(mag[k] + 1e-300), np.log(mag[k + 1] + 1e-300)
        den = a - 2 * b + c
        delta = 0.5 * (a - c) / den if den != 0 else 0.0
    else:
        delta = 0.0
    return (k + delta) / n * L / u


def strouhal_zero_crossing(trace, L: float, u: float, start: int = 0) -> float:
    """St from the mean period between upward crossings of the mean (crossing
    times linearly interpolated)."""
    v = np.array(trace[start:], dtype=np.float64)
    v -= np.mean(v)
    idx = np.flatnonzero((v[:-1] < 0) & (v[1:] >= 0))
    if len(idx) < 2:
        return float("nan")
    t = idx + v[idx] / (v[idx] - v[idx + 1])
    period = (t[-1] - t[0]) / (len(t) - 1)
    return L / (period * u)


# ----------------------------------------------------------------------------- drivers
def von_karman_probe(lx: int, ly: int, plate_size: int, steps: int, nu: float = 0.04, u_in: float = 0.1,
                     rho_in: float = 1.0, storage: str = "f64", walls: bool = False, y_offset: int = 0,
                     probe: Optional[Tuple[int, int]] = None, device: int = 0, chunk: int = 100000,
                     progress=None):
    """x_strouhal (experiments.py:650-720) on one GPU: uniform inflow start,
    probe (rho, ux, uy) at (3 lx//4, ly//2) after every step.

    Returns (samples[steps, 3], lattice)."""
    probe = probe or (3 * lx // 4, ly // 2)
    lat = Lattice(lx, ly, boundary=von_karman_boundary_conditions(lx, ly, rho_in, u_in, plate_size, walls=walls,
                                                                  y_offset=y_offset),
                  storage=storage, device=device)
    rho, u = density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), u_in)
    lat.set_equilibrium(rho, u)
    lat.set_probe(probe[0], probe[1], steps)
    om = omega_from_viscosity(nu)
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        lat.run(om, k)
        done += k
        if progress:
            lat.synchronize()
            progress(done)
    lat.synchronize()
    return lat.probe(), lat


def shear_wave_viscosity(lx: int, ly: int, omega: float, epsilon: float = 0.05, steps: int = 3000,
                         sample_every: int = 50, storage: str = "f64", device: int = 0):
    """Decay of ux = eps sin(2 pi y / ly) (experiments.py:147-223, velocity
    variant): amplitude max |u| sampled every `sample_every` steps, fitted to
    eps exp(-nu k^2 t).  Returns (nu_fit, nu_theory, t, amplitude)."""
    from scipy.optimize import curve_fit

    lat = Lattice(lx, ly, storage=storage, device=device)
    rho, u = sinusoidal_velocity_x((lx, ly), epsilon)
    lat.set_equilibrium(rho, u)
    t, amp = [], []
    done = 0
    while done < steps:
        k = min(sample_every, steps - done)
        lat.run(omega, k)
        done += k
        _, _, uu = lat.state(want=("u",))
        t.append(done)
        amp.append(max(abs(uu.min()), abs(uu.max())))
    t = np.array(t, dtype=np.float64)
    amp = np.array(amp)
    kk = 2 * np.pi / ly
    popt, _ = curve_fit(lambda tt, v: epsilon * np.exp(-v * kk * kk * tt), t, amp, p0=[viscosity_from_omega(omega)])
    return float(popt[0]), viscosity_from_omega(omega), t, amp


def couette_analytic(ly: int, U: float, nu: float, t: float, terms: int = 4000) -> np.ndarray:
    """Start-up Couette flow u_x(y, t) between a wall moving with U at
    y = -1/2 and a rigid wall at y = ly - 1/2 (the bounce-back walls sit half a
    node outside the boundary nodes), from rest: steady line minus the
    Fourier series of the initial deficit."""
    h = float(ly)
    eta = np.arange(ly) + 0.5  # distance from the moving wall
    u = U * (1 - eta / h)
    for n in range(1, terms + 1):
        u -= (2 * U / (n * np.pi)) * np.sin(n * np.pi * eta / h) * np.exp(-nu * (n * np.pi / h) ** 2 * t)
    return u


def couette_profile(lx: int, ly: int, U: float = 0.05, omega: float = 1.0, steps: int = 5000,
                    storage: str = "f64", device: int = 0):
    """plot_couette_flow_vel_vectors (experiments.py:304-363): u_x(y) at
    x = lx/2 after `steps` steps, its linear regression, and the analytic
    start-up profile.  Returns dict."""
    from scipy.stats import linregress

    lat = Lattice(lx, ly, boundary=couette_flow_boundary_conditions(lx, ly, U, 1.0), storage=storage,
                  device=device)
    rho, u = density_1_velocity_0_initial((lx, ly))
    lat.set_equilibrium(rho, u)
    lat.run(omega, steps)
    _, rho, u = lat.state(want=("rho", "u"))
    ux = u[int(lx / 2), :, 0]
    r = linregress(np.arange(ly), ux)
    ana = couette_analytic(ly, U, viscosity_from_omega(omega), steps)
    return {"ux": ux, "slope": r.slope, "intercept": r.intercept, "rvalue": r.rvalue, "analytic": ana,
            "max_abs_error": float(np.abs(ux - ana).max()), "rho": rho}


def poiseuille_analytic(ly: int, dpdx: float, mu: float) -> np.ndarray:
    """Steady Poiseuille profile -dp/dx/(2 mu) y (h - y) at y + 1/2, walls half a
    node outside the boundary nodes (experiments.py:430-437)."""
    y = np.arange(ly) + 0.5
    return -np.reciprocal(2 * mu) * dpdx * y * (ly - y)


def poiseuille_startup_analytic(ly: int, accel: float, nu: float, t: float, terms: int = 4001) -> np.ndarray:
    """Start-up plane Poiseuille flow from rest under a uniform driving
    acceleration G = -dp/dx / rho between walls at y = -1/2 and y = ly - 1/2:
    u = G/(2 nu) eta (h - eta) - sum over odd n of
    4 G h^2/(nu pi^3 n^3) sin(n pi eta/h) exp(-nu (n pi/h)^2 t), eta = y + 1/2."""
    h = float(ly)
    eta = np.arange(ly) + 0.5
    u = accel / (2 * nu) * eta * (h - eta)
    for n in range(1, terms + 1, 2):
        u -= (4 * accel * h * h / (nu * np.pi ** 3 * n ** 3)) * np.sin(n * np.pi * eta / h) * \
            np.exp(-nu * (n * np.pi / h) ** 2 * t)
    return u


def poiseuille_profile(lx: int, ly: int, delta_p: float = 0.001, omega: float = 1.5, steps: int = 40000,
                       storage: str = "f64", device: int = 0):
    """plot_poiseuille_flow_vel_vectors (experiments.py:377-504): pressure
    difference split +-delta_p/2 around rho = 1; returns the profiles at
    x = 1 and x = lx//2, the flux ratio, a quadratic fit and the analytic
    profile."""
    from scipy.optimize import curve_fit

    drho = delta_p * 3
    p_in, p_out = (1 + drho / 2) / 3, (1 - drho / 2) / 3
    lat = Lattice(lx, ly, boundary=poiseuille_flow_boundary_conditions(lx, ly, p_in, p_out), storage=storage,
                  device=device)
    rho, u = density_1_velocity_0_initial((lx, ly))
    lat.set_equilibrium(rho, u)
    lat.run(omega, steps)
    _, rho, u = lat.state(want=("rho", "u"))
    vx = u[..., 0]
    trapz = getattr(np, "trapezoid", None) or np.trapz
    a_in = trapz(vx[1, :], np.arange(ly))
    a_mid = trapz(vx[lx // 2, :], np.arange(ly))
    popt, _ = curve_fit(lambda y, a, b, c: a * y ** 2 + b * y + c, np.arange(ly), vx[lx // 2, :])
    mu = viscosity_from_omega(omega) * np.mean(rho[lx // 2, :])
    ana = poiseuille_analytic(ly, (p_out - p_in) / lx, mu)
    rho_mid = float(np.mean(rho[lx // 2, :]))
    startup = poiseuille_startup_analytic(ly, -(p_out - p_in) / lx / rho_mid, viscosity_from_omega(omega), steps)
    return {"ux_inlet": vx[1, :], "ux_mid": vx[lx // 2, :], "area_inlet": a_in, "area_middle": a_mid,
            "ratio": a_in / a_mid, "popt": popt, "analytic": ana,
            "max_rel_error": float(np.abs(vx[lx // 2, :] - ana).max() / ana.max()),
            "analytic_startup": startup,
            "max_rel_error_startup": float(np.abs(vx[lx // 2, :] - startup).max() / np.abs(startup).max()),
            "rho": rho}


# ----------------------------------------------------------------------------- precision study
def precision_study(lx: int = 8192, ly: int = 2048, plate: int = 400, reynolds: float = 100.0, u_in: float = 0.1,
                    checkpoints=(10000, 50000), strouhal_steps: int = 0, storages=("f32", "f16"),
                    walls: bool = True, y_offset: int = 1, device: int = 0, chunk: int = 20000,
                    progress=None) -> dict:
    """Reduced-precision storage against the FP64 run of the same channel
    (the FP64 device path equals the CPU oracle bit for bit).

    At each checkpoint: relative L2 errors ||x - x64|| / ||x64|| of the
    velocity and density fields (density also against rho64 - 1).  If
    strouhal_steps > 0 every run continues to that many steps and the
    Strouhal numbers of the probe's u_y over the second half are compared.
    Returns a dict of results."""
    nu = u_in * plate / reynolds
    om = omega_from_viscosity(nu)
    bc = von_karman_boundary_conditions(lx, ly, 1.0, u_in, plate, walls=walls, y_offset=y_offset)
    rho0, u0 = density_1_velocity_x_u0_velocity_y_0_initial((lx, ly), u_in)
    total = max(max(checkpoints), strouhal_steps)
    runs = {}
    for st in ("f64",) + tuple(storages):
        lat = Lattice(lx, ly, boundary=bc, storage=st, device=device)
        lat.set_equilibrium(rho0, u0)
        lat.set_probe(3 * lx // 4, ly // 2, total)
        try:
            lat.set_temporal(2)  # identical results, half the traffic
        except LbmError:
            pass
        runs[st] = lat
    out = {"lattice": [lx, ly], "plate": plate, "reynolds": reynolds, "omega": om, "checkpoints": {}}
    done = 0

    def advance(to):
        nonlocal done
        while done < to:
            k = min(chunk, to - done)
            for lat in runs.values():
                lat.run(om, k)
            for lat in runs.values():
                lat.synchronize()
            done += k
            if progress:
                progress(done)

    for cp in sorted(checkpoints):
        advance(cp)
        _, r64, u64 = runs["f64"].state(want=("rho", "u"))
        row = {}
        for st in storages:
            _, r, u = runs[st].state(want=("rho", "u"))
            row[st] = {"l2_u": float(np.linalg.norm(u - u64) / np.linalg.norm(u64)),
                       "l2_rho": float(np.linalg.norm(r - r64) / np.linalg.norm(r64)),
                       "l2_rho_dev": float(np.linalg.norm(r - r64) / max(np.linalg.norm(r64 - 1.0), 1e-300)),
                       "max_abs_u": float(np.abs(u - u64).max())}
        out["checkpoints"][str(cp)] = row
    if strouhal_steps:
        advance(strouhal_steps)
        st_out = {}
        for st, lat in runs.items():
            smp = lat.probe()
            half = len(smp) // 2
            fft_st, _, fft_bin = strouhal_reference(smp[:, 2], plate, u_in, half)
            st_out[st] = {"strouhal_uy_interp": strouhal_interpolated(smp[:, 2], plate, u_in, half),
                          "strouhal_uy_zero_crossing": strouhal_zero_crossing(smp[:, 2], plate, u_in, half),
                          "strouhal_uy_fft": fft_st, "fft_bin": fft_bin,
                          "strouhal_speed_fft": strouhal_reference(speed(smp), plate, u_in, half)[0]}
        for st in storages:
            st_out[st]["rel_error_vs_f64"] = (st_out[st]["strouhal_uy_interp"] /
                                             st_out["f64"]["strouhal_uy_interp"] - 1.0)
        out["strouhal"] = st_out
        out["strouhal_steps"] = strouhal_steps
    for lat in runs.values():
        lat.close()
    return out


def channel_checks(lx: int = 4200, ly: int = 1800, plate: int = 400, steps: int = 50000, box: int = 2048,
                   box_steps: int = 50000, u_in: float = 0.1, reynolds: float = 100.0, device: int = 0,
                   progress=None) -> dict:
    """FP64 channel workload and the two box checks of the reference's
    experiments at a large size:

    * von Karman channel lx x ly, plate `plate` at lx//4 (+1 y offset), rigid
      walls, Re from nu = u_in*plate/Re; probe u_y at (3 lx//4, ly//2);
      Strouhal by the reference's FFT peak and by the sub-bin estimators over
      the second half of the trace.
    * Couette box x box (lid u = 0.05, omega 1) against the start-up Fourier
      series, Poiseuille box x box (delta p = 1e-3, omega 1.5) against the
      start-up series and the flux balance between x = 1 and x = box//2."""
    out = {}
    nu = u_in * plate / reynolds
    smp, lat = von_karman_probe(lx, ly, plate, steps, nu=nu, u_in=u_in, walls=True, y_offset=1, device=device,
                                chunk=10000, progress=(lambda n: progress("channel", n)) if progress else None)
    m, jx, jy = lat.moments_sum()
    lat.close()
    half = steps // 2
    st_fft, f_peak, bin_w = strouhal_reference(smp[:, 2], plate, u_in, half)
    out["channel"] = {"lattice": [lx, ly], "plate": plate, "steps": steps, "omega": omega_from_viscosity(nu),
                      "probe": [3 * lx // 4, ly // 2], "strouhal_uy_fft": st_fft, "fft_bin": bin_w,
                      "strouhal_uy_interp": strouhal_interpolated(smp[:, 2], plate, u_in, half),
                      "strouhal_uy_zero_crossing": strouhal_zero_crossing(smp[:, 2], plate, u_in, half),
                      "strouhal_speed_fft": strouhal_reference(speed(smp), plate, u_in, half)[0],
                      "mass": m, "finite": bool(np.isfinite(smp).all())}
    if progress:
        progress("couette", 0)
    c = couette_profile(box, box, U=0.05, omega=1.0, steps=box_steps, device=device)
    out["couette"] = {"lattice": [box, box], "steps": box_steps, "max_abs_error": c["max_abs_error"],
                      "max_rel_error": c["max_abs_error"] / 0.05}
    if progress:
        progress("poiseuille", 0)
    p = poiseuille_profile(box, box, delta_p=1e-3, omega=1.5, steps=box_steps, device=device)
    out["poiseuille"] = {"lattice": [box, box], "steps": box_steps, "flux_ratio_inlet_middle": p["ratio"],
                         "max_rel_error_startup": p["max_rel_error_startup"],
                         "u_centre": float(p["ux_mid"][box // 2]),
                         "u_centre_startup": float(p["analytic_startup"][box // 2])}
    return out


def step_rates(lx: int, ly: int, boundary: str = "couette", storages=("f32", "f64"), steps: int = 200,
               omega: float = 1.0, device: int = 0) -> dict:
    """Node rate (GLUPS) of lbm_run with one and with two steps per launch.

    The lattice starts at equilibrium with rho = 1, u = (0.05, 0); boundary is
    "periodic", "couette" (lid 0.05) or "channel" (von Karman channel with
    walls, plate ly/5).  Each rate is the best of two timed runs of `steps`
    steps after 10 untimed ones; `state_buffer_bytes` is the size of one of
    the two population buffers (rows padded to 128 bytes)."""
    import time

    bcs = {"periodic": lambda: None,
           "couette": lambda: couette_flow_boundary_conditions(lx, ly, 0.05, 1.0),
           "channel": lambda: von_karman_boundary_conditions(lx, ly, 1.0, 0.1, max(2, ly // 5), walls=True,
                                                             y_offset=1)}
    if boundary not in bcs:
        raise ValueError(f"boundary must be one of {sorted(bcs)}")
    rho = np.ones((lx, ly))
    u = np.zeros((lx, ly, 2))
    u[..., 0] = 0.05
    out = {"lattice": [lx, ly], "boundary": boundary, "steps": steps, "omega": omega, "state_buffer_bytes": {},
           "glups": {}}
    for st in storages:
        elem = {"f64": 8, "f32": 4, "f16": 2}[st]
        pitch = -(-ly * elem // 128) * 128 // elem
        lat = Lattice(lx, ly, boundary=bcs[boundary](), storage=st, device=device)
        lat.set_equilibrium(rho, u)
        rates = {}
        for spl in (1, 2):
            lat.set_temporal(spl)
            best = 0.0
            for _ in range(2):
                lat.run(omega, 10)
                lat.synchronize()
                t0 = time.perf_counter()
                lat.run(omega, steps)
                lat.synchronize()
                best = max(best, lx * ly * steps / (time.perf_counter() - t0) / 1e9)
            rates[f"{spl}_per_launch"] = best
        lat.close()
        out["glups"][st] = rates
        out["state_buffer_bytes"][st] = 9 * lx * pitch * elem
    return out


def _main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="experiment drivers on the GPU")
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("precision", help="FP32/FP16 storage against FP64 on a von Karman channel")
    p.add_argument("--lx", type=int, default=8192)
    p.add_argument("--ly", type=int, default=2048)
    p.add_argument("--plate", type=int, default=400)
    p.add_argument("--checkpoints", type=int, nargs="+", default=[10000, 50000])
    p.add_argument("--strouhal-steps", type=int, default=0)
    p.add_argument("--out", default=None)
    c = sub.add_parser("channel-checks", help="FP64 von Karman Strouhal plus Couette and Poiseuille profiles")
    c.add_argument("--lx", type=int, default=4200)
    c.add_argument("--ly", type=int, default=1800)
    c.add_argument("--plate", type=int, default=400)
    c.add_argument("--steps", type=int, default=50000)
    c.add_argument("--box", type=int, default=2048)
    c.add_argument("--box-steps", type=int, default=50000)
    c.add_argument("--out", default=None)
    r = sub.add_parser("rates", help="node rate with one and two steps per launch on a lattice of any size")
    r.add_argument("--lx", type=int, default=8192)
    r.add_argument("--ly", type=int, default=2048)
    r.add_argument("--boundary", default="couette", choices=["periodic", "couette", "channel"])
    r.add_argument("--storages", nargs="+", default=["f32", "f64"], choices=["f64", "f32", "f16"])
    r.add_argument("--steps", type=int, default=200)
    r.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "rates":
        res = step_rates(a.lx, a.ly, a.boundary, a.storages, a.steps)
        text = json.dumps(res, indent=1)
        print(text)
        if a.out:
            with open(a.out, "w") as fh:
                fh.write(text)
    if a.cmd == "channel-checks":
        res = channel_checks(a.lx, a.ly, a.plate, a.steps, a.box, a.box_steps,
                             progress=lambda tag, n: print(f"{tag} step {n}", flush=True))
        text = json.dumps(res, indent=1)
        print(text)
        if a.out:
            with open(a.out, "w") as fh:
                fh.write(text)
    if a.cmd == "precision":
        res = precision_study(a.lx, a.ly, a.plate, checkpoints=a.checkpoints, strouhal_steps=a.strouhal_steps,
                              progress=lambda n: print(f"step {n}", flush=True))
        text = json.dumps(res, indent=1)
        print(text)
        if a.out:
            with open(a.out, "w") as fh:
                fh.write(text)


if __name__ == "__main__":
    _main()
