// Plain data shared by the kernels, their launchers and host-only code (the
// batch table builder): boundary views, geometry, state views, launch
// settings and the per-job records of batched launches.  No HIP types, so
// host C++ compilers can include it.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "lbm_internal.h"

namespace lbm {

// Boundary tables as the kernels see them (device pointers).
struct BcView {
    const uint16_t *codes = nullptr;       // class per node, N
    const uint8_t *chunk = nullptr;        // nonzero if any node of a 64-node chunk is special
    const uint8_t *wave = nullptr;         // nonzero if any node of wave (x, y/64) is special, x*wpr + y/64
    const ClassEntry *classes = nullptr;
    uint32_t nclasses = 0;                 // entries of classes (0 with no program)
    const uint32_t *special = nullptr;     // list of special nodes
    int64_t nspecial = 0;
    double rho_in = 0, rho_out = 0;        // pressure fix
    uint32_t d1 = 0, d2 = 0;
    uint32_t flags = 0;                    // kProg*
};

// Geometry with a precomputed divisor for node -> (x, y).  Node n = x*ly + y
// is the logical index (boundary codes, chunk flags, AoS host arrays); a
// population plane row starts every `pitch` elements, ly rounded up to whole
// 128-byte lines so that every row of every plane starts on a line.
struct Geom {
    uint32_t lx = 0, ly = 0, N = 0;
    uint32_t pitch = 0;
    uint32_t wpr = 0;        // 64-node waves per row, ceil(ly / 64)
    uint64_t ly_magic = 0;   // floor(2^64 / ly) + 1 (exact 32-bit division by ly)
    uint64_t wpr_magic = 0;  // the same for wpr
};

// The state buffers of one lattice: nine planes (of the storage type) and the
// outlet side buffers (3*ly of the compute type), each double-buffered.
struct StateView {
    void *plane[2] = {nullptr, nullptr};
    void *side[2] = {nullptr, nullptr};
    void *scratch = nullptr;  // kSweepScratchBytes of write-only scratch for the sweep kernel
};

// Write-only scratch of a context: the sweep kernel's entry-path stores land
// here (18 values per thread of a 1024-thread block at most).
constexpr size_t kSweepScratchBytes = 18 * 1024 * 8;
// Most threads of one block of the two-step sweep kernel.
constexpr int kSweepThreadsMax = 1024;

struct ProbeView {
    int64_t cell = -1;       // node index, -1 = off
    double *ring = nullptr;  // capacity * 3 doubles (rho, ux, uy)
    int64_t capacity = 0;
};

struct LaunchCfg {
    int cells_per_thread = 0;  // 0: 16-byte kernel (needs ly % (16/sizeof) == 0), else 1, 2, 4
    int xcd_remap = 1;
    int nt_stores = 0;
    int steps_per_launch = 1;  // 2: two-step row-sweep kernel where supported
    int sweep_rows = 0;        // output rows per block of the sweep kernel (0: automatic)
    int sweep_tile = 0;        // threads per column tile of the sweep kernel (0: whole rows up to one block)
    int sweep_v = 0;           // columns per thread of column tiles (0: two for FP32/FP16, one for FP64)
    int row_interleave = 1;    // long two-step runs move the state to the range-interleaved row order
    int sweep_il = 0;          // (per launch) the buffers are in that order for sweep_rows-row ranges
};

// Launch shape of the sweep kernel for rows of ly nodes: columns per thread,
// column tiles (1: whole rows; else tiles of tw output columns plus a halo
// column each side, lw = tw + 2 computed columns), threads per block.
struct SweepShape {
    int v = 1;
    uint32_t ntiles = 1, tw = 0, lw = 0, threads = 0;
};

// One lattice of a batch, with its buffers already resolved for the step
// parity of the launch (src -> dst).
struct BatchJob {
    const void *src = nullptr;
    void *dst = nullptr;
    const void *side_prev = nullptr;
    void *side_next = nullptr;
    Geom g;
    BcView bc;
    double omega = 1.0;
    int64_t probe_cell = -1;
    double *probe_ring = nullptr;
    int64_t probe_cap = 0;
    int64_t probe_offset = 0;  // ring slot = batch step counter + probe_offset
    uint32_t block0 = 0;       // first block of the job in the batch grid
    uint32_t pad = 0;
};

// One lattice of a batch advanced by the two-step sweep kernel, buffers
// resolved for the pair parity of the launch (src -> dst).
struct SweepJob {
    const void *src = nullptr;
    void *dst = nullptr;
    const void *side_prev = nullptr;
    void *side_next = nullptr;
    void *scratch = nullptr;
    Geom g;
    BcView bc;
    double omega = 1.0;
    int64_t probe_cell = -1;
    double *probe_ring = nullptr;
    int64_t probe_cap = 0;
    int64_t probe_offset = 0;  // step-1 ring slot = batch step counter + probe_offset
    uint32_t rows = 0;         // output rows per block
    uint32_t block0 = 0;       // first block of the job in its bucket's grid
    // row order of the state buffers (lbm_rowlayout.h; il_rows = 0: plain)
    uint32_t il_rows = 0, il_ranges = 0, il_whole = 0;
    uint64_t il_magic = 0;
};

// Packed sweep launch: each block of kPackedWaves waves sweeps several row
// ranges (of one or more lattices) side by side, each on its own run of
// waves and its own part of the block's LDS.
struct PackedSub {
    uint32_t job = 0;     // index into the launch's job records
    uint32_t xa = 0;      // first output row
    uint32_t rows = 0;    // output rows (clipped at the lattice's end)
    uint32_t wave0 = 0;   // first wave of the block that sweeps it
    uint32_t waves = 0;   // waves that sweep it (the row's width in 64-node waves)
    uint32_t lds_off = 0; // byte offset of its LDS area in the block's
};

struct PackedBlock {
    uint32_t sub0 = 0;     // first of its subs in the launch's sub table
    uint32_t nsub = 0;
    uint32_t r_block = 0;  // rows of its longest sub: every wave runs r_block + 2 row iterations
    uint32_t lds = 0;      // LDS bytes its subs use
};

}  // namespace lbm
