// Packing of a batch run's device tables (lbm_batch_tables.h).
#include "lbm_batch_tables.h"

#include <cstring>

namespace lbm {

namespace {

// Append `bytes` bytes at the next aligned offset of the image; returns the offset.
size_t put(std::vector<char> &img, const void *data, size_t bytes) {
    const size_t off = (img.size() + kTableAlign - 1) / kTableAlign * kTableAlign;
    img.resize(off + bytes);
    if (bytes) std::memcpy(img.data() + off, data, bytes);
    return off;
}

}  // namespace

void pack_batch_tables(const std::vector<SweepGroupTables> &groups, const StepTables &step, TableImage &out) {
    out.bytes.clear();
    for (int p = 0; p < 2; p++) out.group_jobs[p].assign(groups.size(), 0);
    out.group_blocks.assign(groups.size(), 0);
    for (size_t i = 0; i < groups.size(); i++) {
        const SweepGroupTables &g = groups[i];
        for (int p = 0; p < 2; p++)
            out.group_jobs[p][i] = put(out.bytes, g.jobs[p].data(), g.jobs[p].size() * sizeof(SweepJob));
        out.group_blocks[i] = put(out.bytes, g.block_job.data(), g.block_job.size() * sizeof(uint32_t));
    }
    for (int p = 0; p < 2; p++) out.step_jobs[p] = put(out.bytes, step.jobs[p].data(), step.jobs[p].size() * sizeof(BatchJob));
    out.step_blocks = put(out.bytes, step.block_job.data(), step.block_job.size() * sizeof(uint32_t));
    // a whole number of alignment units, so that the image can be copied as is
    out.bytes.resize((out.bytes.size() + kTableAlign - 1) / kTableAlign * kTableAlign);
}

}  // namespace lbm
