"""The batch runtime's host-side tables (csrc/lbm_batch_tables.cpp) under
AddressSanitizer and UndefinedBehaviorSanitizer on the CPU.

* Layout: thousands of random group, packed and one-step table sets packed
  into one reused image (growing and shrinking, empty tables included): every
  table on a 256-byte boundary inside the image, no two overlapping, every
  byte equal to its input.
* Packing plan: random row ranges (widths 1-16 waves, LDS sizes up to the
  budget) placed in 16-wave blocks: every range exactly once, no block over
  16 waves or over the LDS budget, each block's ranges on consecutive waves
  and consecutive LDS areas, r_block the block's longest range; ranges wider
  than a block or larger than the budget refused.
Host code only (g++); nothing runs on a GPU."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lattice_boltzmann_parallel_solver_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "lbm_batch_tables.h"

using namespace lbm;

static int fail(const char *what, int it) {
    std::fprintf(stderr, "iteration %d: %s\n", it, what);
    return 1;
}

int main() {
    std::mt19937_64 rng(12345);
    TableImage img;
    for (int it = 0; it < 3000; it++) {
        const int ngroups = (int)(rng() % 12);
        std::vector<SweepGroupTables> groups(ngroups);
        for (auto &g : groups) {
            const int nj = (int)(rng() % 40);
            for (int p = 0; p < 2; p++) {
                g.jobs[p].resize(nj);
                for (auto &j : g.jobs[p]) {
                    j.rows = (uint32_t)rng();
                    j.block0 = (uint32_t)rng();
                    j.omega = (double)(rng() % 1000) / 1000.0;
                    j.probe_offset = (int64_t)rng();
                }
            }
            const int nb = (int)(rng() % (it % 7 == 0 ? 5000 : 300));
            for (int k = 0; k < nb; k++) g.block_job.push_back((uint32_t)rng());
        }
        PackedTables pk;
        const int np = (int)(rng() % 20);
        for (int p = 0; p < 2; p++) {
            pk.jobs[p].resize(np);
            for (auto &j : pk.jobs[p]) j.rows = (uint32_t)rng();
        }
        pk.subs.resize(rng() % 700);
        for (auto &q : pk.subs) q.xa = (uint32_t)rng();
        pk.blocks.resize(rng() % 300);
        for (auto &b : pk.blocks) b.lds = (uint32_t)rng();
        StepTables st;
        const int nf = (int)(rng() % 30);
        for (int p = 0; p < 2; p++) {
            st.jobs[p].resize(nf);
            for (auto &j : st.jobs[p]) {
                j.block0 = (uint32_t)rng();
                j.probe_offset = (int64_t)rng();
            }
        }
        const int nfb = (int)(rng() % 2000);
        for (int k = 0; k < nfb; k++) st.block_job.push_back((uint32_t)rng());
        pack_batch_tables(groups, pk, st, img);

        struct Span { size_t off, len; const void *src; };
        std::vector<Span> spans;
        if (img.group_jobs[0].size() != groups.size() || img.group_blocks.size() != groups.size())
            return fail("offset tables do not match the group count", it);
        for (size_t i = 0; i < groups.size(); i++) {
            for (int p = 0; p < 2; p++)
                spans.push_back({img.group_jobs[p][i], groups[i].jobs[p].size() * sizeof(SweepJob),
                                 groups[i].jobs[p].data()});
            spans.push_back({img.group_blocks[i], groups[i].block_job.size() * sizeof(uint32_t),
                             groups[i].block_job.data()});
        }
        for (int p = 0; p < 2; p++)
            spans.push_back({img.packed_jobs[p], pk.jobs[p].size() * sizeof(SweepJob), pk.jobs[p].data()});
        spans.push_back({img.packed_subs, pk.subs.size() * sizeof(PackedSub), pk.subs.data()});
        spans.push_back({img.packed_blocks, pk.blocks.size() * sizeof(PackedBlock), pk.blocks.data()});
        for (int p = 0; p < 2; p++) spans.push_back({img.step_jobs[p], st.jobs[p].size() * sizeof(BatchJob), st.jobs[p].data()});
        spans.push_back({img.step_blocks, st.block_job.size() * sizeof(uint32_t), st.block_job.data()});
        if (img.bytes.size() % kTableAlign) return fail("image size is not a whole number of alignment units", it);
        for (size_t a = 0; a < spans.size(); a++) {
            const Span &s = spans[a];
            if (s.off % kTableAlign) return fail("table not aligned", it);
            if (s.off + s.len > img.bytes.size()) return fail("table past the end of the image", it);
            if (s.len && std::memcmp(img.bytes.data() + s.off, s.src, s.len)) return fail("table bytes differ", it);
            for (size_t b = a + 1; b < spans.size(); b++) {
                const Span &t = spans[b];
                if (s.len && t.len && s.off < t.off + t.len && t.off < s.off + s.len) return fail("tables overlap", it);
            }
        }
    }
    std::printf("packed and checked 3000 table sets\n");

    // packing plans
    for (int it = 0; it < 400; it++) {
        const uint32_t budget = 40000 + (uint32_t)(rng() % 60000);
        std::vector<PackItem> items;
        const int njobs = 1 + (int)(rng() % 60);
        for (int j = 0; j < njobs; j++) {
            const uint32_t waves = 1 + (uint32_t)(rng() % 16), lx = 3 + (uint32_t)(rng() % 400);
            const uint32_t rows = 1 + (uint32_t)(rng() % 12), lds = 16 * (1 + (uint32_t)(rng() % (budget / 16)));
            for (uint32_t xa = 0; xa < lx; xa += rows)
                items.push_back({(uint32_t)j, xa, lx - xa < rows ? lx - xa : rows, waves, lds});
        }
        PackedTables t;
        if (!plan_packed_blocks(items, 16, budget, t)) return fail("plan refused valid items", it);
        if (t.subs.size() != items.size()) return fail("plan lost or duplicated ranges", it);
        std::vector<int> seen(items.size(), 0);
        size_t next_sub = 0;
        for (const PackedBlock &b : t.blocks) {
            if (b.sub0 != next_sub || b.nsub == 0) return fail("block sub ranges not consecutive", it);
            next_sub += b.nsub;
            uint32_t wave = 0, lds = 0, rmax = 0;
            for (uint32_t k = 0; k < b.nsub; k++) {
                const PackedSub &q = t.subs[b.sub0 + k];
                if (q.wave0 != wave || q.lds_off != lds) return fail("waves or LDS not consecutive", it);
                wave += q.waves;
                int found = -1;
                for (size_t i = 0; i < items.size(); i++)
                    if (!seen[i] && items[i].job == q.job && items[i].xa == q.xa && items[i].rows == q.rows &&
                        items[i].waves == q.waves) {
                        found = (int)i;
                        break;
                    }
                if (found < 0) return fail("sub matches no unplaced range", it);
                seen[(size_t)found] = 1;
                lds = q.lds_off + items[(size_t)found].lds;
                rmax = q.rows > rmax ? q.rows : rmax;
            }
            if (wave > 16) return fail("block over 16 waves", it);
            if (lds != b.lds || lds > budget) return fail("block LDS wrong or over the budget", it);
            if (b.r_block != rmax) return fail("r_block is not the longest range", it);
        }
        if (next_sub != t.subs.size()) return fail("subs outside every block", it);
    }
    std::vector<PackItem> wide = {{0, 0, 4, 17, 64}};
    PackedTables t;
    if (plan_packed_blocks(wide, 16, 1 << 16, t)) return fail("a range wider than a block was accepted", 0);
    std::vector<PackItem> fat = {{0, 0, 4, 2, 1 << 20}};
    if (plan_packed_blocks(fat, 16, 1 << 16, t)) return fail("a range over the LDS budget was accepted", 0);
    std::printf("planned and checked 400 packings\n");
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_batch_table_packing_under_asan(tmp_path):
    src = tmp_path / "pack_driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "pack_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-I", CSRC, "-I", os.path.join(ROOT, "include"), str(src),
           os.path.join(CSRC, "lbm_batch_tables.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "packed and checked 3000 table sets" in run.stdout
    assert "planned and checked 400 packings" in run.stdout
