/*
 * lbm_hip.h - C ABI of the MI355X (gfx950) D2Q9 lattice-Boltzmann library.
 *
 * The reference solver (SimonSchrodi/lattice_boltzmann_parallel_solver) is
 * numpy code with no native layer, so there is no C interface to mirror
 * one-for-one.  Each entry point below names the reference Python function
 * whose work it replaces (file:line in the reference tree); the Python
 * package lattice_boltzmann_parallel_solver_amd keeps those functions'
 * signatures and calls these entry points through ctypes.
 *
 * Conventions
 *  - Every function returns int: LBM_OK (0) or a negative LBM_ERR_* code;
 *    lbm_last_error() returns a thread-local message for the last failure.
 *  - Arrays are C order with x slowest, as in the reference: populations
 *    f[x][y][i] (nx*ny*9), density rho[x][y] (nx*ny), velocity u[x][y][k]
 *    (nx*ny*2).  Inside a context the populations are stored as nine
 *    y-contiguous planes instead; set/get_state convert.
 *  - Pointers flagged LBM_HOST are host memory; LBM_DEVICE pointers are HIP
 *    device memory on the context's device (e.g. a torch tensor's data_ptr).
 *    Host pointers are borrowed for the duration of the call only.
 *  - `stream` arguments are hipStream_t values (NULL = the null stream).
 *  - A context is not re-entrant; distinct contexts are independent.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define LBM_OK 0
#define LBM_ERR_ARG -1      /* invalid argument: shape, omega, descriptor */
#define LBM_ERR_HIP -2      /* a HIP runtime call failed */
#define LBM_ERR_STATE -3    /* call not valid in the context's current state */
#define LBM_ERR_NOMEM -4    /* host or device allocation failed */
#define LBM_ERR_UNSTABLE -5 /* non-finite moments detected */

/* ---- population storage inside a context ----------------------------- */
#define LBM_STORE_F64 0 /* f_i in FP64, FP64 arithmetic in the reference's order */
#define LBM_STORE_F32 1 /* deviations f_i - w_i in FP32, FP32 collide */
#define LBM_STORE_F16 2 /* deviations f_i - w_i in FP16, FP32 collide */

/* ---- element types of the stateless entry points --------------------- */
#define LBM_DT_F64 0
#define LBM_DT_F32 1

/* ---- pointer location ------------------------------------------------- */
#define LBM_HOST 0
#define LBM_DEVICE 1

/* ---- boundary descriptors --------------------------------------------- */
#define LBM_BC_RIGID_WALL 1        /* boundary_conditions.py:78-113  rigid_wall(boundary) */
#define LBM_BC_MOVING_WALL 2       /* boundary_conditions.py:170-214 moving_wall(boundary, u_w, avg_density) */
#define LBM_BC_RIGID_OBJECT 3      /* boundary_conditions.py:116-167 rigid_object(boundary) */
#define LBM_BC_INLET 4             /* boundary_conditions.py:217-254 inlet(shape, density_in, velocity_in) */
#define LBM_BC_OUTLET 5            /* boundary_conditions.py:257-283 outlet() */
#define LBM_BC_PRESSURE_PERIODIC 6 /* boundary_conditions.py:286-348 periodic_with_pressure_variations(boundary, p_in, p_out) */

typedef struct lbm_bc_desc {
    int32_t kind;         /* LBM_BC_* */
    int32_t reserved;     /* must be 0 */
    const uint8_t *mask;  /* host, nx*ny bytes, nonzero = node in the set; NULL for INLET/OUTLET */
    double u_w[2];        /* MOVING_WALL: wall velocity */
    double rho_w;         /* MOVING_WALL: wall density (the reference's avg_density) */
    double rho_in;        /* INLET: density_in */
    double u_in;          /* INLET: velocity_in (x component; y is 0) */
    double p_in;          /* PRESSURE_PERIODIC: inlet pressure */
    double p_out;         /* PRESSURE_PERIODIC: outlet pressure */
} lbm_bc_desc;

/* A boundary program: an ordered descriptor list compiled into per-node
 * classes (last writer wins per node and direction, as in the reference's
 * composites, boundary_utils.py:35-53, 87-108, 164-203).  Host-side object. */
typedef struct lbm_bc_prog lbm_bc_prog;

/* A simulation context: one lattice resident on one device. */
typedef struct lbm_ctx lbm_ctx;

/* ---- library ---------------------------------------------------------- */
const char *lbm_last_error(void);
int lbm_version(void);          /* returns the ABI version (>0) */
int lbm_device_count(int *n);   /* number of visible HIP devices */

/* ---- boundary programs (host only, no device needed) ------------------ */
/* Compile descriptors for an nx*ny lattice.  Replaces the per-step work of the
 * closures built by boundary_conditions.py:78-348 and boundary_utils.py:9-205. */
int lbm_bc_compile(lbm_bc_prog **out, int nx, int ny, const lbm_bc_desc *descs, int n);
int lbm_bc_destroy(lbm_bc_prog *prog);
/* Sizes of the compiled tables: number of node classes (class 0 = plain
 * fluid), number of nodes with a nonzero class, and summary flags (bit 0:
 * outlet present, bit 1: pressure fix present). */
int lbm_bc_query(const lbm_bc_prog *prog, int *nclasses, int64_t *nspecial, uint32_t *flags);
/* Copy the compiled tables out (for inspection and host-side tests):
 * codes[nx*ny] class per node; per class: src word (2 bits per direction:
 * 0 stream, 1 reflect f*[opp] - k, 2 constant k, 3 outlet copy), flags word,
 * k[9]; params = {rho_in, rho_out, D1 bits, D2 bits} of the pressure fix.
 * Any output pointer may be NULL. */
int lbm_bc_export(const lbm_bc_prog *prog, uint16_t *codes, uint32_t *class_src,
                  uint32_t *class_flags, double *class_k, double *params);

/* ---- contexts ----------------------------------------------------------- */
/* Allocate an nx*ny lattice with the given storage on `device`. */
int lbm_create(lbm_ctx **out, int nx, int ny, int storage, int device);
int lbm_destroy(lbm_ctx *ctx);
/* Enqueue the context's work on `hip_stream` (default: a stream the context
 * creates at default priority). */
int lbm_set_stream(lbm_ctx *ctx, void *hip_stream);
/* Attach a compiled boundary program (copied to the device; prog may be
 * destroyed afterwards).  NULL = fully periodic. */
int lbm_set_bc(lbm_ctx *ctx, const lbm_bc_prog *prog);
/* Load the step-input state (f, rho, u) of lattice_boltzmann_method.py:191.
 * rho and u are used as given by the first collision, as the reference does. */
int lbm_set_state(lbm_ctx *ctx, const double *f, const double *rho, const double *u, int where);
/* Same with f = equilibrium(rho, u) (the reference's initialisation,
 * e.g. experiments.py:686-687). */
int lbm_set_state_equilibrium(lbm_ctx *ctx, const double *rho, const double *u, int where);
/* Advance nsteps reference steps (lattice_boltzmann_method.py:191-228 with the
 * attached boundary program).  Asynchronous on the context's stream. */
int lbm_run(lbm_ctx *ctx, double omega, int64_t nsteps);
/* Read the current state (f, rho, u) as the reference's step returns it.
 * Synchronous.  Any output may be NULL. */
int lbm_get_state(lbm_ctx *ctx, double *f, double *rho, double *u, int where);
/* Record (rho, ux, uy) at node (px, py) after every step into a device ring
 * of `capacity` samples (the probe of experiments.py:688-704).  capacity 0
 * disables. Resets the ring. */
int lbm_set_probe(lbm_ctx *ctx, int px, int py, int64_t capacity);
/* Copy the probe samples recorded since lbm_set_probe, oldest first, at most
 * `max` rows of (rho, ux, uy); *count receives the number copied. */
int lbm_read_probe(lbm_ctx *ctx, double *out, int64_t max, int64_t *count);
/* FP64 sums of rho, rho*ux, rho*uy over the current state (mass and
 * momentum; tests/test_mass_preservation.py:33).  Synchronous. */
int lbm_moments_sum(lbm_ctx *ctx, double out[3]);
/* |u| = sqrt(ux*ux + uy*uy) of every node of the current state, nx*ny
 * doubles in C order (x slowest) at `out` (LBM_HOST or LBM_DEVICE memory).
 * The field the reference writes every 100 steps with save_mpiio
 * (experiments.py:638-640, parallelization_utils.py:203-253).  Synchronous. */
int lbm_speed(lbm_ctx *ctx, double *out, int where);
/* Checkpoint of the context's state as it is held on the device (running:
 * populations in the storage type and the outlet buffer; pending: the FP64
 * step input), with a 64-byte header (magic "LBMSTAT1", version 2, nx, ny,
 * storage, phase, step count, payload size).  Resuming from it continues bit
 * for bit as if the run had not stopped, in every storage.  The boundary
 * program and probe are not part of it: load into a context created with the
 * same shape, storage and boundary.  lbm_checkpoint_bytes gives the size
 * lbm_checkpoint_save needs; both are synchronous, host memory only.  No
 * counterpart in the reference (its experiments restart from the initial
 * state). */
int lbm_checkpoint_bytes(lbm_ctx *ctx, int64_t *bytes);
int lbm_checkpoint_save(lbm_ctx *ctx, void *host, int64_t capacity);
int lbm_checkpoint_load(lbm_ctx *ctx, const void *host, int64_t bytes);
int lbm_synchronize(lbm_ctx *ctx);
int lbm_steps_done(const lbm_ctx *ctx, int64_t *steps);
/* Launch tuning of the fused step: nodes per thread (1, 2 or 4; 0 = aligned
 * 16-byte groups, needs ny divisible by 2/4/8 for FP64/FP32/FP16), XCD-aware
 * block order (0/1), non-temporal stores (0/1).  Default {1, 1, 0}, the
 * fastest measured on MI355X for every storage. */
int lbm_set_launch(lbm_ctx *ctx, int cells_per_thread, int xcd_remap, int nt_stores);
/* Temporal blocking of lbm_run: steps_per_launch = 2 advances two steps per
 * kernel launch (a row sweep that keeps the intermediate state on chip and
 * reads and writes the populations once per two steps; results are identical
 * to single steps), 1 = one step per launch (default).  rows_per_block: rows
 * each block of the two-step kernel sweeps (0 = automatic).  Rows wider than
 * one block (ny > 2048 FP32/FP16, > 1024 FP64) run as column tiles with a
 * halo column each side.  Two steps need no pressure fix in the boundary
 * program, nx >= 3 and at most 128 boundary classes; otherwise LBM_ERR_ARG.
 * steps_per_launch = 3 or 4: that many steps per launch (FP32/FP16, ny <=
 * 1024, the LDS of K - 1 intermediate stages within 160 KiB: ny <= 1024 for
 * 3, about 700 for 4); a run's remainder goes two steps and one step at a
 * time.  No counterpart in the reference (a launch setting of this library). */
int lbm_set_temporal(lbm_ctx *ctx, int steps_per_launch, int rows_per_block);
/* Column tiles of the two-step kernel: blocks of at most `threads` threads (a
 * multiple of 64 up to 1024; 0 = whole rows up to 1024 threads, the default),
 * each sweeping a tile of columns plus one halo column each side, so that the
 * row barrier of a small block synchronises only its own waves.
 * cols_per_thread: 1 or 2 columns per thread in FP32/FP16 tiles (0 = 2; FP64
 * always 1).  Results are identical for every tile shape.  LBM_ERR_ARG for
 * other values, or when the tiles would be empty.  No counterpart in the
 * reference (a launch setting). */
int lbm_set_sweep_tile(lbm_ctx *ctx, int threads, int cols_per_thread);
/* Row order of long two-step runs.  interleave = 1 (default): an lbm_run of
 * at least 64 steps with two steps per launch first copies the state into an
 * order that stores offset t of every row range of the sweep next to each
 * other (so the rows that all blocks read and write at once are neighbours in
 * memory), sweeps there, and copies it back before returning -- for FP32 rows
 * of two columns per thread (ny > 1024) and FP64 column tiles (ny > 1024)
 * with 16 or more rows per block and at most two rounds of the row ranges
 * the GPU holds at once, where it measured faster; 0: the plain row
 * order throughout.  Results are identical; the state is in the plain order
 * whenever lbm_run has returned.  LBM_ERR_ARG for other values.  No
 * counterpart in the reference (a layout setting of this library). */
int lbm_set_row_layout(lbm_ctx *ctx, int interleave);

/* ---- batches: many lattices advanced by one launch per step ------------- */
/* A batch groups contexts (same device, same storage) so that one kernel
 * launch per step advances all of them; it replaces the reference's one
 * simulation after another (main.py:150-207) for many small lattices.  Each
 * context keeps its own boundary program, probe and state, and stays usable
 * on its own between batch runs; destroy the batch before its contexts. */
typedef struct lbm_batch lbm_batch;
/* omegas[i] is the relaxation of ctxs[i]. */
int lbm_batch_create(lbm_batch **out, lbm_ctx *const *ctxs, const double *omegas, int n);
int lbm_batch_destroy(lbm_batch *batch);
/* Nodes per thread (1 or 2) and whether step pairs replay from a captured
 * HIP graph (1) or are launched one by one (0).  Default {1, 1}. */
int lbm_batch_set_launch(lbm_batch *batch, int nodes_per_thread, int use_graph);
/* Advance every context of the batch nsteps reference steps (contexts still
 * holding a loaded state start with their first collision).  Synchronous. */
int lbm_batch_run(lbm_batch *batch, int64_t nsteps);
/* Two steps per launch for the batch (steps_per_launch = 2): every context
 * the two-step sweep kernel supports (no pressure fix, ny <= 2048 FP32/FP16 or
 * 1024 FP64) is advanced by it, grouped by block width, each group on a
 * side stream when there are several, `rows_per_block` rows per block; the
 * others keep the one-step batched kernel.  rows_per_block = 0 (automatic):
 * 8, at least 32 for lattices wide enough for blocks of 512 or more threads,
 * and in small batches fewer, so that the batch's rows still make about four
 * blocks per CU (two for the wide lattices); where the wide lattices hold
 * most of the batch, one row count for all of them that fills whole rounds
 * of the blocks the GPU holds at once.
 * steps_per_launch = 3: as 2, and the FP32 groups whose rows take 7-8 or
 * 11-16 wavefronts (385-512 or 641-1024 columns) take three steps per launch,
 * where that measured faster, on a side stream (a batch of one such group
 * only from 8 Mi nodes; with automatic rows only in batches whose rows make 8
 * per block at four blocks per CU).  Results are identical in every mode.
 * Default 1.  No counterpart in the reference. */
int lbm_batch_set_temporal(lbm_batch *batch, int steps_per_launch, int rows_per_block);
/* With two steps per launch, FP32/FP16 contexts of at most 1024 columns can
 * be advanced by one packed launch per step pair: their row ranges,
 * `rows_per_block` rows each, placed side by side in blocks of 16
 * wavefronts, so lattices of every width share the launch and the pair is
 * captured in a HIP graph on the batch's one stream.  packed = 1: one block
 * per packed block; 2: a persistent grid of the blocks the GPU holds at once,
 * each sweeping every grid-th packed block; 0 (default): one kernel per group
 * of equal row width, on side streams (the fastest measured:
 * 130.2 GLUPS on 256 uneven channels against 126.7 packed and 120.3
 * persistent).  Results are identical in every mode.  No counterpart in the
 * reference. */
int lbm_batch_set_packing(lbm_batch *batch, int packed);

/* ---- measurement ------------------------------------------------------ */
/* Streaming device copy of `bytes` (a multiple of 16) from src to dst
 * (LBM_DEVICE), 16 bytes per lane, asynchronous on `stream`: the practical
 * HBM ceiling (read + write) the step kernels are measured against.  No
 * counterpart in the reference. */
int lbm_copy_probe(const void *src, void *dst, int64_t bytes, void *stream);
/* The two-step kernel's memory access pattern with no work: FP32 buffers src
 * and dst of 9 * nx * pitch values (LBM_DEVICE), nine planes of `pitch`
 * values per row, each block sweeping `rows` consecutive rows, loading the
 * next row before it writes the current one, a barrier per row; asynchronous
 * on `stream`.  The bandwidth ceiling of a row sweep, beside the copy's.
 * pitch even, at most 4096.  No counterpart in the reference. */
int lbm_row_sweep_probe(const void *src, void *dst, int nx, int pitch, int rows, void *stream);
/* lbm_row_sweep_probe with options, `mode` a combination of:
 * LBM_PROBE_INTERLEAVED: rows in the range-interleaved order of long
 *   two-step runs (offset t of every `rows`-row range stored together,
 *   lbm_set_row_layout; rows <= 128, nx < 2^24), else the plain order;
 * LBM_PROBE_PULL: plane i of output row x read from row x - cx_i (periodic
 *   in x, the D2Q9 velocities), the rows the pull reads, else from row x.
 * dst row x plane i = the transformed src row (x or x - cx_i) plane i, each
 * at its row's place in the chosen order.  LBM_ERR_ARG for other modes.  No
 * counterpart in the reference. */
#define LBM_PROBE_INTERLEAVED 1
#define LBM_PROBE_PULL 2
int lbm_row_sweep_probe_ordered(const void *src, void *dst, int nx, int pitch, int rows, int mode, void *stream);

/* ---- reductions that allocate nothing ----------------------------------- */
/* Largest ncells lbm_moments_sum_noalloc accepts (174762 blocks of 256). */
#define LBM_NOALLOC_MAX_CELLS 44739072
/* FP64 sums of rho, rho*ux, rho*uy over `ncells` nodes of AoS FP64
 * populations f (LBM_DEVICE, ncells*9 doubles), written to out[0..2]
 * (LBM_DEVICE) asynchronously on `stream`.  The same kernels and reduction
 * order as lbm_moments_sum on a context holding f as its step input, so the
 * sums are identical.  No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops).
 * It calls no hipMalloc/hipFree: the per-block partial sums go to a fixed
 * 4 MiB device array that each GPU holds until the process ends, so ncells is
 * at most LBM_NOALLOC_MAX_CELLS (LBM_ERR_ARG beyond).  All calls on one GPU
 * share that array: calls that overlap on one GPU (other streams, other
 * threads) race and must be ordered by the caller.  ncells = 0 writes zeros. */
int lbm_moments_sum_noalloc(const double *f, int64_t ncells, double *out, void *stream);

/* ---- stateless drop-ins on device arrays (AoS, C order) ----------------- */
/* streaming(), lattice_boltzmann_method.py:140-159: periodic shift, bit-exact. */
int lbm_stream(const void *f_in, void *f_out, int nx, int ny, int dtype, void *stream);
/* equilibrium_distr_func(), lattice_boltzmann_method.py:162-188. */
int lbm_equilibrium(const void *rho, const void *u, void *feq, int64_t ncells, int dtype, void *stream);
/* compute_density(), lattice_boltzmann_method.py:93-105. */
int lbm_density(const void *f, void *rho, int64_t ncells, int dtype, void *stream);
/* compute_velocity_field(), lattice_boltzmann_method.py:108-137 (uses the given rho). */
int lbm_velocity(const void *rho, const void *f, void *u, int64_t ncells, int dtype, void *stream);
/* BGK collision f + (equilibrium(rho, u) - f)*omega, lattice_boltzmann_method.py:213-215. */
int lbm_collide(const void *f, const void *rho, const void *u, double omega, void *f_out,
                int64_t ncells, int dtype, void *stream);
/* Post-streaming boundary rules of a program applied in place to f_post
 * (the closures of boundary_conditions.py:93-111, 138-165, 192-212, 239-252,
 * 267-281).  f_prev is read only if the program has an outlet. FP64 only. */
int lbm_bc_apply(const lbm_bc_prog *prog, const double *f_pre, double *f_post,
                 const double *f_prev, void *stream);
/* Pre-streaming pressure fix of a program applied in place to f_pre
 * (boundary_conditions.py:320-346). FP64 only. */
int lbm_bc_pressure(const lbm_bc_prog *prog, double *f_pre, const double *rho,
                    const double *u, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
