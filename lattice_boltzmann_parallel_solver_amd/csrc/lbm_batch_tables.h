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

// The records of the packed sweep launch (one per pair parity), its subs and blocks.
struct PackedTables {
    std::vector<SweepJob> jobs[2];
    std::vector<PackedSub> subs;
    std::vector<PackedBlock> blocks;
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
    size_t packed_jobs[2] = {0, 0};
    size_t packed_subs = 0, packed_blocks = 0;
    size_t step_jobs[2] = {0, 0};
    size_t step_blocks = 0;
};

constexpr size_t kTableAlign = 256;

// Lay the tables out in `out` (its previous contents are replaced).
void pack_batch_tables(const std::vector<SweepGroupTables> &groups, const PackedTables &packed, const StepTables &step,
                       TableImage &out);

// One row range to place in the packed sweep launch: `waves` waves of the
// block and `lds` bytes of its LDS (16-byte multiple).
struct PackItem {
    uint32_t job = 0, xa = 0, rows = 0, waves = 0, lds = 0;
};

// Place the items in blocks of at most `block_waves` waves and `lds_budget`
// bytes of LDS: items in order of decreasing width (then job, then row), each
// into the open block with the fewest free waves that still holds it and has
// the LDS, else a new block.  Consecutive row ranges of a job thus share a
// block where they fit.  Items wider than a block, or with more LDS than the
// budget, are refused (returns false).  Fills tables.subs and tables.blocks.
bool plan_packed_blocks(const std::vector<PackItem> &items, uint32_t block_waves, uint32_t lds_budget,
                        PackedTables &tables);

// One lattice of the wide two-step groups of a batch: its rows and how many
// blocks of its group one CU holds at once (1 to 8).
struct WideLattice {
    uint32_t lx = 0;
    int per_cu = 1;
};

// Rows per block shared by the wide groups: the r in 1..max_rows that
// minimises rounds(r) x (r + halo), where rounds(r) = ceil(sum of
// ceil(lx / r) / per_cu over the lattices, over cus) is the number of rounds
// of resident blocks the launches take; ties go to the smaller r.  0 when
// there is no lattice or cus < 1.
int pick_wide_rows(const std::vector<WideLattice> &lattices, int cus, int max_rows, int halo);

// Rows per block asked of a group's launches (before the kernels' own
// clamps): base_rows (the batch's setting) for rows of fewer than 512
// threads, at least 32 for wider ones, at least 4 base_rows for three steps
// per launch (k = 3).  With automatic rows (rows_auto), wide two-step groups
// take wide_rows where it is set, and the others no more than leave the
// batch's rows_total rows about four blocks per CU (two for wide rows).
int batch_group_rows(int threads, int k, int base_rows, bool rows_auto, int64_t rows_total, int cus, int wide_rows);

// With automatic rows, three steps per launch only where the batch's rows
// make at least 8 per block at four blocks per CU.
bool batch_three_steps_fill(int64_t rows_total, int cus);

}  // namespace lbm
