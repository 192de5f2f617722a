// Packing of a batch run's device tables (lbm_batch_tables.h).
#include "lbm_batch_tables.h"

#include <algorithm>
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

void pack_batch_tables(const std::vector<SweepGroupTables> &groups, const PackedTables &packed, const StepTables &step,
                       TableImage &out) {
    out.bytes.clear();
    for (int p = 0; p < 2; p++) out.group_jobs[p].assign(groups.size(), 0);
    out.group_blocks.assign(groups.size(), 0);
    for (size_t i = 0; i < groups.size(); i++) {
        const SweepGroupTables &g = groups[i];
        for (int p = 0; p < 2; p++)
            out.group_jobs[p][i] = put(out.bytes, g.jobs[p].data(), g.jobs[p].size() * sizeof(SweepJob));
        out.group_blocks[i] = put(out.bytes, g.block_job.data(), g.block_job.size() * sizeof(uint32_t));
    }
    for (int p = 0; p < 2; p++)
        out.packed_jobs[p] = put(out.bytes, packed.jobs[p].data(), packed.jobs[p].size() * sizeof(SweepJob));
    out.packed_subs = put(out.bytes, packed.subs.data(), packed.subs.size() * sizeof(PackedSub));
    out.packed_blocks = put(out.bytes, packed.blocks.data(), packed.blocks.size() * sizeof(PackedBlock));
    for (int p = 0; p < 2; p++) out.step_jobs[p] = put(out.bytes, step.jobs[p].data(), step.jobs[p].size() * sizeof(BatchJob));
    out.step_blocks = put(out.bytes, step.block_job.data(), step.block_job.size() * sizeof(uint32_t));
    // a whole number of alignment units, so that the image can be copied as is
    out.bytes.resize((out.bytes.size() + kTableAlign - 1) / kTableAlign * kTableAlign);
}

bool plan_packed_blocks(const std::vector<PackItem> &items, uint32_t block_waves, uint32_t lds_budget,
                        PackedTables &tables) {
    tables.subs.clear();
    tables.blocks.clear();
    for (const PackItem &it : items)
        if (it.waves == 0 || it.waves > block_waves || it.lds > lds_budget || it.rows == 0) return false;
    std::vector<uint32_t> order(items.size());
    for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const PackItem &x = items[a], &y = items[b];
        if (x.waves != y.waves) return x.waves > y.waves;
        if (x.job != y.job) return x.job < y.job;
        return x.xa < y.xa;
    });
    // open blocks by free waves; each block's subs collected before the
    // table is laid out in block order
    struct Open {
        uint32_t used_waves = 0, lds = 0, r_block = 0;
        std::vector<uint32_t> members;  // indices into items
    };
    std::vector<Open> blocks;
    std::vector<std::vector<uint32_t>> by_free(block_waves + 1);  // free waves -> open block ids
    for (uint32_t idx : order) {
        const PackItem &it = items[idx];
        int chosen = -1;
        for (uint32_t f = it.waves; f <= block_waves && chosen < 0; f++) {
            std::vector<uint32_t> &lst = by_free[f];
            for (size_t k = lst.size(); k-- > 0;) {
                const uint32_t b = lst[k];
                if (blocks[b].lds + it.lds <= lds_budget) {
                    chosen = (int)b;
                    lst.erase(lst.begin() + (ptrdiff_t)k);
                    break;
                }
            }
        }
        if (chosen < 0) {
            chosen = (int)blocks.size();
            blocks.emplace_back();
        }
        Open &o = blocks[(size_t)chosen];
        o.members.push_back(idx);
        o.used_waves += it.waves;
        o.lds += it.lds;
        o.r_block = std::max(o.r_block, it.rows);
        const uint32_t free_waves = block_waves - o.used_waves;
        if (free_waves > 0) by_free[free_waves].push_back((uint32_t)chosen);
    }
    for (const Open &o : blocks) {
        PackedBlock pb;
        pb.sub0 = (uint32_t)tables.subs.size();
        pb.nsub = (uint32_t)o.members.size();
        pb.r_block = o.r_block;
        uint32_t wave = 0, lds = 0;
        for (uint32_t idx : o.members) {
            const PackItem &it = items[idx];
            PackedSub q;
            q.job = it.job;
            q.xa = it.xa;
            q.rows = it.rows;
            q.wave0 = wave;
            q.waves = it.waves;
            q.lds_off = lds;
            wave += it.waves;
            lds += it.lds;
            tables.subs.push_back(q);
        }
        pb.lds = lds;
        tables.blocks.push_back(pb);
    }
    return true;
}

int pick_wide_rows(const std::vector<WideLattice> &lattices, int cus, int max_rows, int halo) {
    if (lattices.empty() || cus < 1 || max_rows < 1) return 0;
    // exact arithmetic: CU shares in units of 1 / 840 (every per_cu of 1..8 divides 840)
    constexpr uint64_t kUnit = 840;
    int best_r = 0;
    uint64_t best = 0;
    for (int r = 1; r <= max_rows; r++) {
        uint64_t share = 0;
        for (const WideLattice &w : lattices) {
            const uint64_t per_cu = (uint64_t)(w.per_cu >= 1 && w.per_cu <= 8 ? w.per_cu : 1);
            share += ((uint64_t)w.lx + (uint64_t)r - 1) / (uint64_t)r * (kUnit / per_cu);
        }
        const uint64_t per_round = (uint64_t)cus * kUnit;
        const uint64_t rounds = (share + per_round - 1) / per_round;
        const uint64_t cost = rounds * (uint64_t)(r + halo);
        if (best_r == 0 || cost < best) {
            best = cost;
            best_r = r;
        }
    }
    return best_r;
}

int batch_group_rows(int threads, int k, int base_rows, bool rows_auto, int64_t rows_total, int cus, int wide_rows) {
    int rows = threads >= 512 ? std::max(base_rows, 32) : base_rows;
    if (k == 3) rows = std::max(rows, 4 * base_rows);
    if (rows_auto) {
        if (k == 2 && threads >= 512 && wide_rows > 0) {
            rows = wide_rows;
        } else {
            const int64_t slots = (int64_t)(threads >= 512 ? 2 : 4) * std::max(cus, 1);
            rows = (int)std::min<int64_t>(rows, std::max<int64_t>(1, rows_total / slots));
        }
    }
    return rows;
}

bool batch_three_steps_fill(int64_t rows_total, int cus) { return rows_total / (4 * (int64_t)std::max(cus, 1)) >= 8; }

}  // namespace lbm
