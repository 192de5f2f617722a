// Kernel argument blocks and host-side launch wrappers (lbm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_internal.h"
#include "lbm_types.h"

namespace lbm {


// `align`: elements per 128-byte line of the plane type (1 for AoS-only uses).
Geom make_geom(uint32_t lx, uint32_t ly, uint32_t align = 1);


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
// Output rows per block that launch_sweep2 uses with this configuration.
int sweep2_rows(int storage, const Geom &g, const BcView &bc, const LaunchCfg &cfg);
// Row ranges of the two-step launch the GPU holds at once (0: unknown).
int sweep2_resident_ranges(int storage, const Geom &g, const BcView &bc, const LaunchCfg &cfg);
// planes[src] -> planes[src^1] and side[src] -> side[src^1], rows moved to
// (to_ilv) or from their slots in the range-interleaved order of `rows`-row
// ranges (lbm_rowlayout.h); launch_sweep2 with cfg.sweep_il and cfg.sweep_rows
// = rows then runs on buffers in that order.
hipError_t launch_row_layout(int storage, const Geom &g, const StateView &st, int src, uint32_t rows, bool to_ilv,
                             hipStream_t s);
// The sweep kernel needs no pressure fix, at most 128 boundary classes,
// lx >= 3, the state buffer (9 lx pitch values) addressable with 32-bit byte
// offsets, and its LDS (12 (cols + 1) storage values, 6 cols compute values,
// the class table, one byte per row and wave column; cols = ly or a column
// tile's width plus 2) within 160 KiB.
bool sweep_supported(int storage, const Geom &g, const BcView &bc, int tile_threads = 0, int tile_v = 0);
// K (3 or 4) fused steps in one launch: the row sweep with K stages, one
// column per thread over whole rows (ly <= 1024), FP32/FP16 storage.  The
// samples of step s (1-based) go to ring slot probe_slot + s - 1 (< 0: none).
// hipErrorInvalidValue if sweepk_supported() is false.
hipError_t launch_sweepk(int storage, int K, const Geom &g, const StateView &st, int src, const BcView &bc,
                         double omega, const ProbeView &pr, int64_t probe_slot, const LaunchCfg &cfg,
                         hipStream_t s);
// As sweep_supported(), and ly <= 1024, FP32 or FP16, its LDS (K - 1 stages of
// exchange and outlet buffers) within 160 KiB.
bool sweepk_supported(int storage, const Geom &g, const BcView &bc, int K);
// Rows per block of the K-step kernel: `rows`, raised until every block
// (the last one included) holds at least K - 1 rows.
int sweepk_rows(const Geom &g, int rows, int K);
// LDS bytes of the K-step kernel for a lattice of this ly, class count and rows per block.
size_t sweepk_lds_bytes(int storage, uint32_t ly, uint32_t nclasses, uint32_t rows, int K);
// K (= 3) steps of every job of a batch bucket (whole rows of at most 1024
// columns, FP32/FP16, state buffers below 4 GiB, rows per block from
// sweepk_rows); probe slots as launch_sweep2_batch, plus s - 1 for step s.
hipError_t launch_sweepk_batch(int storage, int K, const SweepJob *jobs, const uint32_t *block_job, unsigned nblocks,
                               unsigned threads, size_t lds, const int64_t *step, int64_t step_now, hipStream_t s);
hipError_t prepare_sweepk_batch(int storage, int K, size_t lds);
// The batched sweep takes whole rows only (no column tiles).
bool sweep_batch_supported(int storage, const Geom &g, const BcView &bc);
SweepShape sweep_shape(int storage, uint32_t ly, int tile_threads = 0, int tile_v = 0);
// Output rows per block of the sweep kernel: cfg.sweep_rows if set, else
// spread over whole waves of `slots` resident blocks.
int sweep_rows_for(const Geom &g, const LaunchCfg &cfg, int slots);
// Nodes per thread of the 16-byte kernel for a storage type.
int vec16_nodes(int storage);


// One fused step of every job of a batch (block b runs job block_job[b]),
// followed by a one-thread kernel that advances the batch step counter *step.
hipError_t launch_fused_batch(int storage, int nodes_per_thread, const BatchJob *jobs, const uint32_t *block_job,
                              unsigned nblocks, int64_t *step, unsigned *done, hipStream_t s);
// Blocks of one job in the batched one-step launch (its rows' waves, kBlock/64
// waves and nodes_per_thread nodes a thread).
int64_t batch_job_blocks(const Geom &g, int nodes_per_thread);


// Two steps of every job of a bucket (all jobs with `threads` threads per
// block and v columns per thread; lds = the largest layout among them).
// Probe slots count from *step, or from step_now where step is null.
// big: the bucket's state buffers are 4 GiB or more (sweep_big).
hipError_t launch_sweep2_batch(int storage, int v, bool big, const SweepJob *jobs, const uint32_t *block_job,
                               unsigned nblocks, unsigned threads, size_t lds, const int64_t *step, int64_t step_now,
                               hipStream_t s);
// Blocks of the batched sweep kernel of this shape one CU holds at once with
// `lds` bytes of LDS each (1 if the runtime cannot say).
int sweep2_batch_per_cu(int storage, int v, bool big, unsigned threads, size_t lds);
// Allow `lds` bytes of dynamic LDS for the batched sweep kernel (call before
// launching or capturing it).
hipError_t prepare_sweep2_batch(int storage, int v, bool big, size_t lds);
// The packed two-step launch (k_sweep2_packed): blocks of
// packed_block_waves() waves, each sweeping the subs of one PackedBlock side
// by side; one column per thread, state buffers below 4 GiB.  A sub of a
// lattice with this storage, ly, class count and rows takes
// packed_sub_lds(...) bytes of its block's LDS.
uint32_t packed_block_waves();
uint32_t packed_sub_lds(int storage, uint32_t ly, uint32_t nclasses, uint32_t rows);
// Block g of the launch sweeps packed blocks g, g + grid, ...: packed_grid
// gives one block per packed block, or with `persistent` the blocks the GPU
// holds at once (CUs x blocks per CU at this LDS size, at most nblocks).
hipError_t prepare_sweep2_packed(int storage, size_t lds);
hipError_t packed_grid(int storage, size_t lds, unsigned nblocks, bool persistent, unsigned *grid);
hipError_t launch_sweep2_packed(int storage, const SweepJob *jobs, const PackedSub *subs, const PackedBlock *blocks,
                                unsigned nblocks, unsigned grid, size_t lds, const int64_t *step, int64_t step_now,
                                hipStream_t s);
// Whether a lattice's state buffers need 64-bit row bases in the sweep
// kernels (4 GiB or more each).
bool sweep_big(int storage, const Geom &g);
// LDS bytes of the sweep kernel for a lattice of this ly, class count and rows per block.
size_t sweep_lds_bytes(int storage, uint32_t ly, uint32_t nclasses, uint32_t rows);
// The two-step kernel's row-sweep access pattern with no work (measurement
// only): FP32 buffers of 9 lx pitch values, `rows` rows per block, rows in
// the plain or the range-interleaved order of those ranges; with `pull` each
// plane is read from the row the pull reads it from (x - cx_i).
hipError_t launch_row_sweep_probe(const void *src, void *dst, uint32_t lx, uint32_t pitch, uint32_t rows,
                                  bool interleaved, bool pull, hipStream_t s);
// dst[i] = src[i] for n16 16-byte chunks (streaming copy, measurement only).
hipError_t launch_copy16(const void *src, void *dst, int64_t n16, hipStream_t s);
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
