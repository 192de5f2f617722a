"""Write the compact golden fixtures of tests/golden/ from the reference's data files.

Runs in the development container only (it reads /root/reference, which the
GPU machines do not have); the committed outputs are data:

  vel_at_p.npy                       the reference's own 12-value probe fixture
                                     (tests/von_karman_vortex_shedding/vel_at_p.npy),
                                     copied unchanged.
  von_karman_420x180_reference.npz   digests and subsamples of the reference's
                                     420x180, Re=100 von Karman results:
      serial run, 100k steps (milestoneQuickFunctionCalls.py:293-341):
        figures/von_karman_vortex_shedding/vel_at_p.py.npy   |u| probe trace
        figures/von_karman_vortex_shedding/density.py.npy    final density
        figures/von_karman_vortex_shedding/velocity.py.npy   final velocity
      MPI run, 200k steps (experiments.py:650-720):
        figures/von_karman_vortex_shedding/reynold_strouhal/vel_at_p_100.npy
  von_karman_sweeps.json             every probe trace of the reference's Strouhal
                                     sweeps (figures/von_karman_vortex_shedding/
                                     {reynold,nx,blockage}_strouhal/vel_at_p_*.npy,
                                     main.py:150-207): run parameters, length,
                                     SHA-256, every 1000th sample and the Strouhal
                                     number by the reference's FFT procedure
                                     (visualizations_utils.py:135-250).
  reference_csv.json                 the Couette / Poiseuille fits the reference
                                     published (figures/couette_flow/linregress.csv,
                                     figures/poiseuille_flow/{areas,curve_fit}.csv).

Digests are SHA-256 of the arrays' little-endian float64 bytes in C order.
"""

import csv
import glob
import hashlib
import json
import os
import shutil

import numpy as np

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
VK = os.path.join(REF, "figures", "von_karman_vortex_shedding")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def main():
    shutil.copyfile(os.path.join(REF, "tests", "von_karman_vortex_shedding", "vel_at_p.npy"),
                    os.path.join(OUT, "vel_at_p.npy"))

    serial = np.load(os.path.join(VK, "vel_at_p.py.npy"))
    rho = np.load(os.path.join(VK, "density.py.npy"))
    vel = np.load(os.path.join(VK, "velocity.py.npy"))
    par = np.load(os.path.join(VK, "reynold_strouhal", "vel_at_p_100.npy"))
    pts = np.array([[315, 90], [105, 90], [106, 70], [419, 0], [0, 179], [210, 45], [300, 130]])
    np.savez_compressed(
        os.path.join(OUT, "von_karman_420x180_reference.npz"),
        serial_steps=np.array(len(serial) - 1),
        serial_trace_sha=np.array(sha(serial)),
        serial_trace_head=serial[:6001],
        serial_trace_every100=serial[::100],
        serial_density_sha=np.array(sha(rho)),
        serial_velocity_sha=np.array(sha(vel)),
        serial_points=pts,
        serial_density_at=rho[pts[:, 0], pts[:, 1]],
        serial_velocity_at=vel[pts[:, 0], pts[:, 1]],
        serial_density_sum=np.array(rho.sum()),
        parallel_steps=np.array(len(par) - 1),
        parallel_trace_sha=np.array(sha(par)),
        parallel_trace_head=par[:6001],
        parallel_trace_every100=par[::100],
    )

    sweeps = []
    for folder in ("reynold_strouhal", "nx_strouhal", "blockage_strouhal"):
        for path in sorted(glob.glob(os.path.join(VK, folder, "vel_at_p_*.npy"))):
            tag = path[path.rfind("_") + 1:path.rfind(".npy")]
            tr = np.load(path)
            lx, ly, nu, u0 = 420, 180, 0.04, 0.1
            if folder == "reynold_strouhal":
                re = int(tag)
                if re > 130:   # main.py:154-165
                    nu = 0.03
                    u0 = re * nu / 40
                else:
                    nu = 40 * 0.1 / re
                window = 150000 if re == 40 else (125000 if re == 70 else 70000)  # visualizations_utils.py:150-155
            elif folder == "nx_strouhal":
                lx = int(tag)
                window = {260: 125000, 300: 90000, 350: 90000, 700: 90000, 100: 175000}.get(lx, 75000)
            else:
                ly = int(round(40 / float(tag)))
                window = 90000
            v = tr[window:].copy()
            v -= np.mean(v)
            freq = np.fft.fftfreq(len(v), 1)
            f_peak = float(np.abs(freq[np.argmax(np.abs(np.fft.fft(v)))]))
            sweeps.append({"folder": folder, "file": os.path.basename(path), "lx": lx, "ly": ly, "plate": 40,
                           "nu": nu, "u0": u0, "steps": len(tr) - 1, "probe": [3 * lx // 4, ly // 2],
                           "sha256": sha(tr), "every1000": tr[::1000].tolist(), "window": window,
                           "peak_frequency": f_peak, "strouhal": f_peak * 40 / u0})
    with open(os.path.join(OUT, "von_karman_sweeps.json"), "w") as fh:
        json.dump(sweeps, fh, indent=1)

    with open(os.path.join(REF, "figures", "couette_flow", "linregress.csv")) as fh:
        couette = next(csv.DictReader(fh))
    with open(os.path.join(REF, "figures", "poiseuille_flow", "areas.csv")) as fh:
        areas = next(csv.DictReader(fh))
    with open(os.path.join(REF, "figures", "poiseuille_flow", "curve_fit.csv")) as fh:
        fit = next(csv.DictReader(fh))
    popt = [float(v) for v in fit["popt"].strip("[]").split()]
    data = {
        "couette": {"lattice": [20, 30], "omega": 1.0, "U": 0.05, "steps": 5000,
                    "slope": float(couette["slope"]), "intercept": float(couette["intercept"]),
                    "rvalue": float(couette["rvalue"])},
        "poiseuille": {"lattice": [200, 60], "omega": 1.5, "delta_p": 0.001, "steps": 40000,
                       "area_inlet": float(areas["inlet"]), "area_middle": float(areas["middle"]),
                       "ratio": float(areas["relative_difference"]), "popt": popt},
    }
    with open(os.path.join(OUT, "reference_csv.json"), "w") as fh:
        json.dump(data, fh, indent=1)


if __name__ == "__main__":
    main()
