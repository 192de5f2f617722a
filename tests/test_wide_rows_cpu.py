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


GROUP_DRIVER = r"""
#include <cstdio>

#include "lbm_batch_tables.h"

int main() {
    int threads, k, base, autorows, cus, wide;
    long long total;
    while (std::scanf("%d %d %d %d %lld %d %d", &threads, &k, &base, &autorows, &total, &cus, &wide) == 7)
        std::printf("%d %d\n", lbm::batch_group_rows(threads, k, base, autorows != 0, total, cus, wide),
                    lbm::batch_three_steps_fill(total, cus) ? 1 : 0);
    return 0;
}
"""


def test_group_rows_of_large_small_and_wide_batches(tmp_path):
    """Rows per block asked of a batch group: the 256-channel batch (180,000
    rows on 256 CUs) keeps 8 rows for narrow groups, 32 for wide ones and 32
    for three-step groups; four channels of 2462 rows in all get 2 rows (four
    blocks per CU) and too few rows for three steps per launch; wide groups
    of a batch they dominate take the whole-round count; rows given by the
    caller are only raised (32 for wide rows, 4x for three steps).  Groups of
    512 threads or more count as wide (about two blocks per CU)."""
    if shutil.which("g++") is None:
        pytest.skip("g++ missing")
    src = tmp_path / "group_rows.cpp"
    src.write_text(GROUP_DRIVER)
    exe = str(tmp_path / "group_rows")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", CSRC, "-o", exe, str(src),
                    os.path.join(CSRC, "lbm_batch_tables.cpp")], check=True)
    cases = [
        # threads, k, base rows, automatic, rows in the batch, CUs, wide rows -> rows, three steps fill
        (256, 2, 8, 1, 180000, 256, 0, 8, 1),
        (576, 2, 8, 1, 180000, 256, 0, 32, 1),
        (512, 3, 8, 1, 180000, 256, 0, 32, 1),
        (256, 2, 8, 1, 2462, 256, 0, 2, 0),
        (512, 3, 8, 1, 2462, 256, 0, 4, 0),    # 512 threads count as wide: two blocks per CU
        (576, 2, 8, 1, 5819, 256, 0, 11, 0),
        (1024, 2, 8, 1, 8192, 256, 32, 32, 1),
        (1024, 2, 8, 1, 24576, 256, 97, 97, 1),
        (256, 2, 5, 0, 100, 1, 0, 5, 1),
        (1024, 2, 5, 0, 100, 1, 0, 32, 1),
        (256, 3, 5, 0, 100, 1, 0, 20, 1),
        (256, 2, 8, 1, 8191, 256, 0, 7, 0),   # just under 8 rows per block at four blocks per CU
        (256, 2, 8, 1, 8192, 256, 0, 8, 1),
    ]
    text = "".join(f"{t} {k} {b} {a} {tot} {c} {w}\n" for t, k, b, a, tot, c, w, _, _ in cases)
    out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout.split()
    got = [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(len(cases))]
    assert got == [(r, f) for *_, r, f in cases]
