// Host-only layout of the device tables of one batch run: the job records and
// block -> job tables of every sweep group (one set per pair parity) and of
// the one-step jobs (one set per step parity), packed into one byte image
// that is staged through pinned memory and copied to the device in one
// transfer.  Plain C++ (no HIP), so that a host build with AddressSanitizer
// can exercise it (tests/test_batch_tables_asan.py).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "lbm_types.h"

namespace lbm {

// The records of one group of jobs advanced by one sweep launch per pair.
struct SweepGroupTables {
    std::vector<SweepJob> jobs[2];    // per pair parity, one record per member job
    std::vector<uint32_t> block_job;  // launch block -> index into jobs[p]
};

// The records of the jobs advanced by one-step launches.
struct StepTables {
    std::vector<BatchJob> jobs[2];    // per step parity
    std::vector<uint32_t> block_job;  // launch block -> index into jobs[p]
};

// Byte offsets of each table in the image; every table starts on a 256-byte
// boundary.  Empty tables get an offset too (nothing is stored there).
struct TableImage {
    std::vector<char> bytes;
    std::vector<size_t> group_jobs[2];  // per group, per pair parity
    std::vector<size_t> group_blocks;   // per group
    size_t step_jobs[2] = {0, 0};
    size_t step_blocks = 0;
};

constexpr size_t kTableAlign = 256;

// Lay the tables out in `out` (its previous contents are replaced).
void pack_batch_tables(const std::vector<SweepGroupTables> &groups, const StepTables &step, TableImage &out);

}  // namespace lbm
