"""The batch runtime's host-side table layout (csrc/lbm_batch_tables.cpp)
under AddressSanitizer and UndefinedBehaviorSanitizer on the CPU: a driver
packs thousands of random group/job/block table sets into one reused image
(growing and shrinking, empty groups and empty one-step tables included) and
checks that every table lands on a 256-byte boundary inside the image, that
no two tables overlap, and that every byte equals its input.  Host code only
(g++); nothing runs on a GPU."""

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
        pack_batch_tables(groups, st, img);

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
