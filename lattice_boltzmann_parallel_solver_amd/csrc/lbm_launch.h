// Kernel argument blocks and host-side launch wrappers (lbm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
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

// `align`: elements per 128-byte line of the plane type (1 for AoS-only uses).
Geom make_geom(uint32_t lx, uint32_t ly, uint32_t align = 1);

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
};

// One fused step: planes[src] -> planes[dst], side[src] -> side[dst].
// probe_slot < 0: no probe sample; partials != nullptr: per-block moment sums.
hipError_t launch_fused(int storage, const Geom &g, const StateView &st, int src, const BcView &bc,
                        double omega, const ProbeView &pr, int64_t probe_slot, double *partials,
                        const LaunchCfg &cfg, hipStream_t s, int *nblocks_out);
int fused_blocks(const Geom &g, const LaunchCfg &cfg);

// Two fused steps in one launch (row sweep with the intermediate state on
// chip): planes[src] -> planes[src^1], side[src] -> side[src^1].  probe_slot1
// and probe_slot2 (< 0: none) receive the samples of the two steps.  Returns
// hipErrorInvalidValue if sweep_supported() is false.
hipError_t launch_sweep2(int storage, const Geom &g, const StateView &st, int src, const BcView &bc,
                         double omega, const ProbeView &pr, int64_t probe_slot1, int64_t probe_slot2,
                         const LaunchCfg &cfg, hipStream_t s);
// The sweep kernel needs no pressure fix, at most 128 boundary classes,
// lx >= 3, the state buffer (9 lx pitch values) addressable with 32-bit byte
// offsets, and its LDS (12 (cols + 1) storage values, 6 cols compute values,
// the class table, one byte per row and wave column; cols = ly or a column
// tile's width plus 2) within 160 KiB.
bool sweep_supported(int storage, const Geom &g, const BcView &bc);
// The batched sweep takes whole rows only (no column tiles).
bool sweep_batch_supported(int storage, const Geom &g, const BcView &bc);
// Launch shape of the sweep kernel for rows of ly nodes: columns per thread,
// column tiles (1: whole rows; else tiles of tw output columns plus a halo
// column each side, lw = tw + 2 computed columns), threads per block.
struct SweepShape {
    int v = 1;
    uint32_t ntiles = 1, tw = 0, lw = 0, threads = 0;
};
SweepShape sweep_shape(int storage, uint32_t ly);
// Output rows per block of the sweep kernel: cfg.sweep_rows if set, else
// spread over whole waves of `slots` resident blocks (at least 8 rows).
int sweep_rows_for(const Geom &g, const LaunchCfg &cfg, int slots);
// Nodes per thread of the 16-byte kernel for a storage type.
int vec16_nodes(int storage);

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

// One fused step of every job of a batch (block b runs job block_job[b]),
// followed by a one-thread kernel that advances the batch step counter *step.
hipError_t launch_fused_batch(int storage, int nodes_per_thread, const BatchJob *jobs, const uint32_t *block_job,
                              unsigned nblocks, int64_t *step, unsigned *done, hipStream_t s);
// Blocks of one job in the batched one-step launch (its rows' waves, kBlock/64
// waves and nodes_per_thread nodes a thread).
int64_t batch_job_blocks(const Geom &g, int nodes_per_thread);

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
    uint32_t pad = 0;
};

// Two steps of every job of a bucket (all jobs with `threads` threads per
// block and v columns per thread; lds = the largest layout among them).
// Probe slots count from *step, or from step_now where step is null.
// big: the bucket's state buffers are 4 GiB or more (sweep_big).
hipError_t launch_sweep2_batch(int storage, int v, bool big, const SweepJob *jobs, const uint32_t *block_job,
                               unsigned nblocks, unsigned threads, size_t lds, const int64_t *step, int64_t step_now,
                               hipStream_t s);
// Allow `lds` bytes of dynamic LDS for the batched sweep kernel (call before
// launching or capturing it).
hipError_t prepare_sweep2_batch(int storage, int v, bool big, size_t lds);
// Whether a lattice's state buffers need 64-bit row bases in the sweep
// kernels (4 GiB or more each).
bool sweep_big(int storage, const Geom &g);
// LDS bytes of the sweep kernel for a lattice of this ly, class count and rows per block.
size_t sweep_lds_bytes(int storage, uint32_t ly, uint32_t nclasses, uint32_t rows);
// *step += k (one thread).
hipError_t launch_add_step(int64_t *step, int64_t k, hipStream_t s);

// First collision from AoS FP64 (f, rho, u) into planes[dst] / side[dst].
hipError_t launch_collide_first(int storage, const Geom &g, const double *f, const double *rho,
                                const double *u, const StateView &st, int dst, const BcView &bc,
                                double omega, hipStream_t s);

// Stream + boundary + moments from planes[src] into AoS FP64 outputs (any may
// be null); partials != nullptr adds the per-block moment sums.
hipError_t launch_finalize(int storage, const Geom &g, const StateView &st, int src, const BcView &bc,
                           double *f, double *rho, double *u, double *partials, hipStream_t s,
                           int *nblocks_out);

// Moments of one node into probe.ring[slot].
hipError_t launch_probe_cell(int storage, const Geom &g, const StateView &st, int src, const BcView &bc,
                             const ProbeView &pr, int64_t slot, hipStream_t s);

// |u| = sqrt(ux*ux + uy*uy) of n AoS FP64 velocities.
hipError_t launch_speed(const double *u, double *out, int64_t n, hipStream_t s);

// Sum nblocks rows of 3 partials into out[3] in a fixed order.
hipError_t launch_sum_partials(const double *partials, int nblocks, double *out, hipStream_t s);
// Moment sums of AoS FP64 populations (nblocks_out partial rows).
hipError_t launch_moments_sum_aos(const double *f, int64_t n, double *partials, hipStream_t s,
                                  int *nblocks_out);
int partial_rows_needed(const Geom &g, const LaunchCfg &cfg);

// Stateless AoS kernels; dtype LBM_DT_F64 / LBM_DT_F32.
hipError_t launch_stream_aos(const void *in, void *out, int nx, int ny, int dtype, hipStream_t s);
hipError_t launch_equilibrium_aos(const void *rho, const void *u, void *feq, int64_t n, int dtype,
                                  hipStream_t s);
hipError_t launch_density_aos(const void *f, void *rho, int64_t n, int dtype, hipStream_t s);
hipError_t launch_velocity_aos(const void *rho, const void *f, void *u, int64_t n, int dtype,
                               hipStream_t s);
hipError_t launch_collide_aos(const void *f, const void *rho, const void *u, double omega, void *out,
                              int64_t n, int dtype, hipStream_t s);
hipError_t launch_bc_apply(const Geom &g, const BcView &bc, const double *f_pre, double *f_post,
                           const double *f_prev, hipStream_t s);
hipError_t launch_bc_pressure(const Geom &g, const BcView &bc, double *f_pre, const double *rho,
                              const double *u, hipStream_t s);

}  // namespace lbm
