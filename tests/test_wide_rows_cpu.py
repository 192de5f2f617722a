"""The batch runtime's choice of rows per block for its wide two-step groups
(pick_wide_rows in csrc/lbm_batch_tables.cpp), on the host: known answers for
the batches measured on MI355X (256 CUs) and, for random batches, the
minimum of rounds x (rows + halo) recomputed here with exact fractions.
Host code only (g++); nothing runs on a GPU."""

import os
import random
import shutil
import subprocess
from fractions import Fraction
from math import ceil

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lattice_boltzmann_parallel_solver_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <vector>

#include "lbm_batch_tables.h"

int main() {
    int n, cus, max_rows, halo;
    while (std::scanf("%d %d %d %d", &n, &cus, &max_rows, &halo) == 4) {
        std::vector<lbm::WideLattice> v(n);
        for (int k = 0; k < n; k++) std::scanf("%u %d", &v[k].lx, &v[k].per_cu);
        std::printf("%d\n", lbm::pick_wide_rows(v, cus, max_rows, halo));
    }
    return 0;
}
"""


def _build(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ missing")
    src = tmp_path / "wide_rows.cpp"
    src.write_text(DRIVER)
    exe = str(tmp_path / "wide_rows")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", CSRC, "-o", exe, str(src),
                    os.path.join(CSRC, "lbm_batch_tables.cpp")], check=True)
    return exe


def _run(exe, cases):
    text = ""
    for lats, cus, max_rows, halo in cases:
        text += f"{len(lats)} {cus} {max_rows} {halo}\n" + "".join(f"{lx} {pc}\n" for lx, pc in lats)
    out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout.split()
    return [int(x) for x in out]


def _cost(lats, cus, r, halo):
    share = sum(Fraction(ceil(lx / r), pc) for lx, pc in lats)
    return ceil(share / cus) * (r + halo)


def test_known_batches(tmp_path):
    exe = _build(tmp_path)
    cases = [
        ([(8192, 1)], 256, 128, 3),                                  # one 8192x2048 channel: 256 blocks, one round
        ([(8192, 1)] * 3, 256, 128, 3),                              # three: 97 rows, 255 blocks
        ([(4200, 1)] * 4, 256, 128, 3),                              # four 4200x1800: 66 rows, 256 blocks
        ([(2048, 1)], 256, 128, 3),                                  # a 2048x2048 box: 8 rows, 256 blocks
        ([(8192, 1), (8192, 1), (2048, 1), (1500, 1)], 256, 128, 3),  # serving-like mix: 79 rows, 253 blocks
        ([(1200, 3)], 256, 128, 3),                                  # one 514-column channel, 3 blocks per CU
        ([], 256, 128, 3),
        ([(100, 1)], 0, 128, 3),
    ]
    assert _run(exe, cases) == [32, 97, 66, 8, 79, 2, 0, 0]


def test_random_batches_minimise_rounds_times_rows(tmp_path):
    exe = _build(tmp_path)
    rng = random.Random(7)
    cases = []
    for _ in range(300):
        lats = [(rng.randint(3, 20000), rng.randint(1, 4)) for _ in range(rng.randint(1, 12))]
        cases.append((lats, rng.choice([8, 64, 256, 304]), rng.choice([16, 64, 128]), rng.randint(0, 4)))
    for (lats, cus, max_rows, halo), r in zip(cases, _run(exe, cases)):
        assert 1 <= r <= max_rows
        costs = [_cost(lats, cus, q, halo) for q in range(1, max_rows + 1)]
        assert costs[r - 1] == min(costs)
        assert r == 1 + costs.index(min(costs))  # ties go to the fewest rows
