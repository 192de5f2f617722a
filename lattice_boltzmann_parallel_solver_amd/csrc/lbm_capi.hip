// C ABI of liblbm_hip.so: contexts, state transfer, the step loop, probes,
// reductions and the stateless drop-ins (declared in include/lbm_hip.h).
//
// A context keeps one lattice resident on one device.  Its phase follows the
// reference's step (lattice_boltzmann_method.py:191-228):
//   PENDING  the step input (f, rho, u) sits in the AoS FP64 staging buffers;
//            the next step starts with a collision that uses rho, u as given.
//   RUNNING  the planes hold the post-collision populations of step n-1
//            (after the pre-streaming pressure fix); the reference's state f_n
//            is stream+boundary of them and is materialised only on request.
// A run of k steps is therefore 1 first-collision + (k-1) fused kernels
// from PENDING, or k fused kernels from RUNNING.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "lbm_batch_tables.h"
#include "lbm_rowlayout.h"
#include "lbm_launch.h"

namespace lbm {
namespace {
// The last error message of this thread, in a fixed buffer so that setting it
// allocates nothing (the no-allocation entry point reports errors through it).
thread_local char g_err[512];

void set_error_fmt(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
}  // namespace
void set_error(const std::string &msg) { set_error_fmt("%s", msg.c_str()); }
}  // namespace lbm

using namespace lbm;

#define LBM_HIP_CHECK(expr, what)                                                  \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            set_error(std::string(what) + ": " + hipGetErrorString(e_));           \
            return LBM_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

namespace {

enum Phase { kEmpty = 0, kPending = 1, kRunning = 2 };

size_t plane_elem(int storage) { return storage == LBM_STORE_F64 ? 8 : (storage == LBM_STORE_F32 ? 4 : 2); }
size_t side_elem(int storage) { return storage == LBM_STORE_F64 ? 8 : 4; }

// One device allocation holding a boundary program's tables.
struct BcUpload {
    void *mem = nullptr;
    BcView view;
};

hipError_t upload_bc(const lbm_bc_prog &p, BcUpload &up) {
    const size_t N = p.codes.size();
    const size_t o_codes = 0;
    const size_t o_chunk = (o_codes + N * sizeof(uint16_t) + 255) & ~size_t(255);
    const size_t o_wave = (o_chunk + p.chunk_flags.size() + 255) & ~size_t(255);
    const size_t o_cls = (o_wave + p.wave_flags.size() + 255) & ~size_t(255);
    const size_t o_spec = (o_cls + p.classes.size() * sizeof(ClassEntry) + 255) & ~size_t(255);
    const size_t total = o_spec + p.special.size() * sizeof(uint32_t) + 256;
    char *mem = nullptr;
    hipError_t e = hipMalloc(&mem, total);
    if (e != hipSuccess) return e;
    e = hipMemcpy(mem + o_codes, p.codes.data(), N * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(mem + o_chunk, p.chunk_flags.data(), p.chunk_flags.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(mem + o_wave, p.wave_flags.data(), p.wave_flags.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(mem + o_cls, p.classes.data(), p.classes.size() * sizeof(ClassEntry), hipMemcpyHostToDevice);
    if (e == hipSuccess && !p.special.empty())
        e = hipMemcpy(mem + o_spec, p.special.data(), p.special.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(mem);
        return e;
    }
    up.mem = mem;
    up.view.codes = reinterpret_cast<const uint16_t *>(mem + o_codes);
    up.view.chunk = reinterpret_cast<const uint8_t *>(mem + o_chunk);
    up.view.wave = reinterpret_cast<const uint8_t *>(mem + o_wave);
    up.view.classes = reinterpret_cast<const ClassEntry *>(mem + o_cls);
    up.view.nclasses = (uint32_t)p.classes.size();
    up.view.special = reinterpret_cast<const uint32_t *>(mem + o_spec);
    up.view.nspecial = (int64_t)p.special.size();
    up.view.rho_in = p.rho_in;
    up.view.rho_out = p.rho_out;
    up.view.d1 = p.d1;
    up.view.d2 = p.d2;
    up.view.flags = p.flags;
    return hipSuccess;
}

}  // namespace

struct lbm_ctx {
    int nx = 0, ny = 0, storage = 0, device = 0;
    Geom g;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    StateView st;
    double *stage_f = nullptr, *stage_rho = nullptr, *stage_u = nullptr;
    BcUpload bc;
    int cur = 0;
    int phase = kEmpty;
    int64_t steps = 0;
    ProbeView probe;
    int64_t probe_base = 0;
    double *partials = nullptr;
    double *sums = nullptr;
    LaunchCfg cfg;
};

namespace {

// Two-step runs of at least this many launches take the range-interleaved row
// order (moving the state there and back costs about two launches).
constexpr int64_t kInterleaveMinPairs = 32;
constexpr int kInterleaveMinRows = 16;

// Bytes of one state buffer: nine planes of lx rows, `pitch` elements a row.
size_t plane_bytes(const lbm_ctx *c) { return 9 * (size_t)c->g.lx * c->g.pitch * plane_elem(c->storage); }

void release(lbm_ctx *c) {
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < 2; k++) {
        (void)hipFree(c->st.plane[k]);
        (void)hipFree(c->st.side[k]);
    }
    (void)hipFree(c->st.scratch);
    (void)hipFree(c->stage_f);
    (void)hipFree(c->stage_rho);
    (void)hipFree(c->stage_u);
    (void)hipFree(c->bc.mem);
    (void)hipFree(c->probe.ring);
    (void)hipFree(c->partials);
    (void)hipFree(c->sums);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
}

int check_ctx(lbm_ctx *c, const char *fn) {
    if (!c) {
        set_error(std::string(fn) + ": context is NULL");
        return LBM_ERR_ARG;
    }
    LBM_HIP_CHECK(hipSetDevice(c->device), fn);
    return LBM_OK;
}

hipMemcpyKind kind_in(int where) { return where == LBM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice; }
hipMemcpyKind kind_out(int where) { return where == LBM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost; }

}  // namespace

extern "C" {

const char *lbm_last_error(void) { return g_err; }

int lbm_version(void) { return 1; }

int lbm_device_count(int *n) {
    if (!n) {
        set_error("lbm_device_count: n is NULL");
        return LBM_ERR_ARG;
    }
    *n = 0;
    LBM_HIP_CHECK(hipGetDeviceCount(n), "hipGetDeviceCount");
    return LBM_OK;
}

int lbm_bc_destroy(lbm_bc_prog *prog) {
    if (prog && prog->dev_mem) {
        (void)hipSetDevice(prog->dev_id);
        (void)hipFree(prog->dev_mem);
    }
    delete prog;
    return LBM_OK;
}

int lbm_create(lbm_ctx **out, int nx, int ny, int storage, int device) {
    if (!out) {
        set_error("lbm_create: out is NULL");
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    if (nx < 1 || ny < 1 || (int64_t)nx * ny > 0x7fffffffLL) {
        set_error("lbm_create: lattice must satisfy 1 <= nx*ny < 2^31");
        return LBM_ERR_ARG;
    }
    if (storage < LBM_STORE_F64 || storage > LBM_STORE_F16) {
        set_error("lbm_create: unknown storage " + std::to_string(storage));
        return LBM_ERR_ARG;
    }
    int ndev = 0;
    LBM_HIP_CHECK(hipGetDeviceCount(&ndev), "lbm_create: hipGetDeviceCount");
    if (device < 0 || device >= ndev) {
        set_error("lbm_create: device " + std::to_string(device) + " not present (" + std::to_string(ndev) +
                  " visible)");
        return LBM_ERR_ARG;
    }
    auto *c = new (std::nothrow) lbm_ctx();
    if (!c) {
        set_error("lbm_create: out of host memory");
        return LBM_ERR_NOMEM;
    }
    c->nx = nx;
    c->ny = ny;
    c->storage = storage;
    c->device = device;
    c->g = make_geom((uint32_t)nx, (uint32_t)ny, (uint32_t)(128 / plane_elem(storage)));
    // one node per thread with XCD-aware block order: the fastest shape
    // measured on MI355X for all three storages (8192x2048 channel)
    c->cfg = LaunchCfg{1, 1, 0};
    const size_t N = (size_t)nx * ny;
    hipError_t e = hipSetDevice(device);
    auto alloc = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    for (int k = 0; k < 2; k++) {
        alloc(&c->st.plane[k], plane_bytes(c));
        alloc(&c->st.side[k], 3 * (size_t)ny * side_elem(storage));
    }
    alloc(&c->st.scratch, kSweepScratchBytes);
    alloc(reinterpret_cast<void **>(&c->stage_f), 9 * N * sizeof(double));
    alloc(reinterpret_cast<void **>(&c->stage_rho), N * sizeof(double));
    alloc(reinterpret_cast<void **>(&c->stage_u), 2 * N * sizeof(double));
    alloc(reinterpret_cast<void **>(&c->partials), 3 * (size_t)partial_rows_needed(c->g, LaunchCfg{1, 1, 0}) *
                                                       sizeof(double));
    alloc(reinterpret_cast<void **>(&c->sums), 3 * sizeof(double));
    for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipMemset(c->st.side[k], 0, 3 * (size_t)ny * side_elem(storage));
    if (e == hipSuccess) {
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        c->own_stream = e == hipSuccess;
    }
    if (e != hipSuccess) {
        set_error(std::string("lbm_create: ") + hipGetErrorString(e));
        release(c);
        delete c;
        return e == hipErrorOutOfMemory ? LBM_ERR_NOMEM : LBM_ERR_HIP;
    }
    *out = c;
    return LBM_OK;
}

int lbm_destroy(lbm_ctx *c) {
    if (!c) return LBM_OK;
    release(c);
    delete c;
    return LBM_OK;
}

int lbm_set_stream(lbm_ctx *c, void *s) {
    int rc = check_ctx(c, "lbm_set_stream");
    if (rc) return rc;
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_set_stream");
    if (c->own_stream) LBM_HIP_CHECK(hipStreamDestroy(c->stream), "lbm_set_stream");
    c->stream = static_cast<hipStream_t>(s);
    c->own_stream = false;
    return LBM_OK;
}

int lbm_set_launch(lbm_ctx *c, int v, int xcd, int nt) {
    int rc = check_ctx(c, "lbm_set_launch");
    if (rc) return rc;
    if (v != 0 && v != 1 && v != 2 && v != 4) {
        set_error("lbm_set_launch: cells_per_thread must be 0 (16-byte groups), 1, 2 or 4");
        return LBM_ERR_ARG;
    }
    if (v == 0 && c->ny % vec16_nodes(c->storage) != 0) {
        set_error("lbm_set_launch: 16-byte groups need ny divisible by " + std::to_string(vec16_nodes(c->storage)));
        return LBM_ERR_ARG;
    }
    c->cfg.cells_per_thread = v;
    c->cfg.xcd_remap = xcd ? 1 : 0;
    c->cfg.nt_stores = nt ? 1 : 0;
    return LBM_OK;
}

int lbm_set_temporal(lbm_ctx *c, int steps_per_launch, int rows_per_block) {
    int rc = check_ctx(c, "lbm_set_temporal");
    if (rc) return rc;
    if (steps_per_launch < 1 || steps_per_launch > 4) {
        set_error("lbm_set_temporal: steps_per_launch must be 1, 2, 3 or 4");
        return LBM_ERR_ARG;
    }
    if (rows_per_block < 0) {
        set_error("lbm_set_temporal: rows_per_block must be >= 0 (0 = automatic)");
        return LBM_ERR_ARG;
    }
    if (steps_per_launch == 2 && !sweep_supported(c->storage, c->g, c->bc.view, c->cfg.sweep_tile, c->cfg.sweep_v)) {
        set_error("lbm_set_temporal: two steps per launch need no pressure fix, at most 128 boundary classes "
                  "and nx >= 3");
        return LBM_ERR_ARG;
    }
    if (steps_per_launch >= 3 && !sweepk_supported(c->storage, c->g, c->bc.view, steps_per_launch)) {
        set_error("lbm_set_temporal: " + std::to_string(steps_per_launch) +
                  " steps per launch need FP32 or FP16 storage, ny <= 1024, no pressure fix, at most 128 boundary "
                  "classes, nx >= 3 and their LDS within 160 KiB");
        return LBM_ERR_ARG;
    }
    c->cfg.steps_per_launch = steps_per_launch;
    c->cfg.sweep_rows = rows_per_block;
    return LBM_OK;
}

int lbm_set_sweep_tile(lbm_ctx *c, int threads, int cols_per_thread) {
    int rc = check_ctx(c, "lbm_set_sweep_tile");
    if (rc) return rc;
    if (threads < 0 || threads % 64 != 0 || threads > kSweepThreadsMax) {
        set_error("lbm_set_sweep_tile: threads must be 0 (whole rows) or a multiple of 64 up to 1024");
        return LBM_ERR_ARG;
    }
    if (cols_per_thread < 0 || cols_per_thread > 2) {
        set_error("lbm_set_sweep_tile: cols_per_thread must be 0 (automatic), 1 or 2");
        return LBM_ERR_ARG;
    }
    if (c->cfg.steps_per_launch == 2 && !sweep_supported(c->storage, c->g, c->bc.view, threads, cols_per_thread)) {
        set_error("lbm_set_sweep_tile: no column tiles of " + std::to_string(threads) + " threads for ny = " +
                  std::to_string(c->ny));
        return LBM_ERR_ARG;
    }
    c->cfg.sweep_tile = threads;
    c->cfg.sweep_v = cols_per_thread;
    return LBM_OK;
}

int lbm_set_row_layout(lbm_ctx *c, int interleave) {
    int rc = check_ctx(c, "lbm_set_row_layout");
    if (rc) return rc;
    if (interleave != 0 && interleave != 1) {
        set_error("lbm_set_row_layout: interleave must be 0 or 1");
        return LBM_ERR_ARG;
    }
    c->cfg.row_interleave = interleave;
    return LBM_OK;
}

// Materialise f_n, rho_n, u_n of a running context into (device) outputs.
static int finalize_into(lbm_ctx *c, double *f, double *rho, double *u) {
    LBM_HIP_CHECK(launch_finalize(c->storage, c->g, c->st, c->cur, c->bc.view, f, rho, u, nullptr, c->stream,
                                  nullptr),
                  "finalize kernel");
    return LBM_OK;
}

int lbm_set_bc(lbm_ctx *c, const lbm_bc_prog *prog) {
    int rc = check_ctx(c, "lbm_set_bc");
    if (rc) return rc;
    if (prog && (prog->nx != c->nx || prog->ny != c->ny)) {
        set_error("lbm_set_bc: program compiled for " + std::to_string(prog->nx) + "x" + std::to_string(prog->ny) +
                  ", context is " + std::to_string(c->nx) + "x" + std::to_string(c->ny));
        return LBM_ERR_ARG;
    }
    if (c->phase == kRunning) {
        // the boundary changes what the lazily-held f_n is; fix f_n first
        rc = finalize_into(c, c->stage_f, c->stage_rho, c->stage_u);
        if (rc) return rc;
        c->phase = kPending;
    }
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_set_bc");
    (void)hipFree(c->bc.mem);
    c->bc = BcUpload{};
    if (prog) LBM_HIP_CHECK(upload_bc(*prog, c->bc), "lbm_set_bc: upload");
    return LBM_OK;
}

int lbm_set_state(lbm_ctx *c, const double *f, const double *rho, const double *u, int where) {
    int rc = check_ctx(c, "lbm_set_state");
    if (rc) return rc;
    if (!f || !rho || !u) {
        set_error("lbm_set_state: f, rho and u are required");
        return LBM_ERR_ARG;
    }
    const size_t N = c->g.N;
    LBM_HIP_CHECK(hipMemcpyAsync(c->stage_f, f, 9 * N * sizeof(double), kind_in(where), c->stream), "lbm_set_state");
    LBM_HIP_CHECK(hipMemcpyAsync(c->stage_rho, rho, N * sizeof(double), kind_in(where), c->stream), "lbm_set_state");
    LBM_HIP_CHECK(hipMemcpyAsync(c->stage_u, u, 2 * N * sizeof(double), kind_in(where), c->stream), "lbm_set_state");
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_set_state");
    c->phase = kPending;
    c->steps = 0;
    c->probe_base = 0;
    return LBM_OK;
}

int lbm_set_state_equilibrium(lbm_ctx *c, const double *rho, const double *u, int where) {
    int rc = check_ctx(c, "lbm_set_state_equilibrium");
    if (rc) return rc;
    if (!rho || !u) {
        set_error("lbm_set_state_equilibrium: rho and u are required");
        return LBM_ERR_ARG;
    }
    const size_t N = c->g.N;
    LBM_HIP_CHECK(hipMemcpyAsync(c->stage_rho, rho, N * sizeof(double), kind_in(where), c->stream),
                  "lbm_set_state_equilibrium");
    LBM_HIP_CHECK(hipMemcpyAsync(c->stage_u, u, 2 * N * sizeof(double), kind_in(where), c->stream),
                  "lbm_set_state_equilibrium");
    LBM_HIP_CHECK(launch_equilibrium_aos(c->stage_rho, c->stage_u, c->stage_f, (int64_t)N, LBM_DT_F64, c->stream),
                  "equilibrium kernel");
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_set_state_equilibrium");
    c->phase = kPending;
    c->steps = 0;
    c->probe_base = 0;
    return LBM_OK;
}

int lbm_run(lbm_ctx *c, double omega, int64_t nsteps) {
    int rc = check_ctx(c, "lbm_run");
    if (rc) return rc;
    if (!(omega > 0.0 && omega < 2.0)) {  // lattice_boltzmann_method.py:212
        set_error("lbm_run: omega must satisfy 0 < omega < 2");
        return LBM_ERR_ARG;
    }
    if (nsteps < 0) {
        set_error("lbm_run: nsteps must be >= 0");
        return LBM_ERR_ARG;
    }
    if (c->phase == kEmpty) {
        set_error("lbm_run: no state loaded (call lbm_set_state first)");
        return LBM_ERR_STATE;
    }
    if (nsteps == 0) return LBM_OK;
    int64_t left = nsteps;
    if (c->phase == kPending) {
        LBM_HIP_CHECK(launch_collide_first(c->storage, c->g, c->stage_f, c->stage_rho, c->stage_u, c->st, c->cur,
                                           c->bc.view, omega, c->stream),
                      "first-collision kernel");
        c->phase = kRunning;
        c->steps++;
        left--;
    }
    const bool probing = c->probe.cell >= 0;
    const int K = c->cfg.steps_per_launch;
    if (K >= 3 && sweepk_supported(c->storage, c->g, c->bc.view, K)) {
        for (; left >= K; left -= K) {
            // materialises f_{steps} .. f_{steps+K-1}
            const int64_t slot = probing ? c->steps - c->probe_base - 1 : -1;
            LBM_HIP_CHECK(launch_sweepk(c->storage, K, c->g, c->st, c->cur, c->bc.view, omega, c->probe, slot,
                                        c->cfg, c->stream),
                          "K-step sweep kernel");
            c->cur ^= 1;
            c->steps += K;
        }
    }
    if (K >= 2 && sweep_supported(c->storage, c->g, c->bc.view, c->cfg.sweep_tile, c->cfg.sweep_v)) {
        // A long run moves the state into the range-interleaved row order of
        // the sweep's row ranges (the rows all blocks touch at once become
        // neighbours in memory) and back at its end: two copies of the state.
        // Measured where it pays: FP32 rows of two columns per thread and FP64
        // column tiles, with 16 or more rows per block and at most two rounds
        // of resident row ranges (8192 x 2048 FP32 137.3 -> 147.4 GLUPS, FP64
        // 67.7 -> 69.4, 4200 x 1800 FP32 125.5 -> 128.3, 65536 x 2048 FP32 in
        // two rounds 141.8 -> 144.1).  It lost on one column per thread (8192
        // x 1024 FP32 130.1 -> 121.5), at 8 rows per block (2048 x 2048 FP32
        // 125.7 -> 122.9), in FP16 (220.5 -> 214.4) and over seven rounds
        // (65536 x 2048 FP64 69.1 -> 65.1), which keep the plain order.
        LaunchCfg cfg = c->cfg;
        bool ilv = false;
        const SweepShape sh = sweep_shape(c->storage, c->g.ly, cfg.sweep_tile, cfg.sweep_v);
        if (cfg.row_interleave && left / 2 >= kInterleaveMinPairs && c->g.lx < kRowLayoutMaxLx &&
            c->storage != LBM_STORE_F16 && (sh.v == 2 || sh.ntiles > 1)) {
            const int rows = sweep2_rows(c->storage, c->g, c->bc.view, cfg);
            const int64_t ranges = rows > 0 ? ((int64_t)c->g.lx + rows - 1) / rows : 0;
            const int resident = sweep2_resident_ranges(c->storage, c->g, c->bc.view, cfg);
            if (rows >= kInterleaveMinRows && rows <= (int)kRowLayoutMaxRows && rows < (int)c->g.lx &&
                resident > 0 && ranges <= 2 * (int64_t)resident) {
                LBM_HIP_CHECK(launch_row_layout(c->storage, c->g, c->st, c->cur, (uint32_t)rows, true, c->stream),
                              "row order kernel");
                c->cur ^= 1;
                cfg.sweep_rows = rows;
                cfg.sweep_il = 1;
                ilv = true;
            }
        }
        hipError_t e = hipSuccess;
        for (; left >= 2; left -= 2) {
            // materialises f_{steps} and f_{steps+1}
            const int64_t slot = probing ? c->steps - c->probe_base - 1 : -1;
            e = launch_sweep2(c->storage, c->g, c->st, c->cur, c->bc.view, omega, c->probe, slot,
                              probing ? slot + 1 : -1, cfg, c->stream);
            if (e != hipSuccess) break;
            c->cur ^= 1;
            c->steps += 2;
        }
        // back to the plain order, also after a failed launch (the state then
        // holds the steps done so far)
        if (ilv) {
            const hipError_t eb = launch_row_layout(c->storage, c->g, c->st, c->cur, (uint32_t)cfg.sweep_rows, false,
                                                    c->stream);
            if (eb == hipSuccess) c->cur ^= 1;
            LBM_HIP_CHECK(eb, "row order kernel");
        }
        LBM_HIP_CHECK(e, "two-step sweep kernel");
    }
    for (; left > 0; left--) {
        // this launch materialises f_{steps} in registers: record it if probed
        const int64_t slot = probing ? c->steps - c->probe_base - 1 : -1;
        LBM_HIP_CHECK(launch_fused(c->storage, c->g, c->st, c->cur, c->bc.view, omega, c->probe, slot, nullptr,
                                   c->cfg, c->stream, nullptr),
                      "fused step kernel");
        c->cur ^= 1;
        c->steps++;
    }
    if (probing && c->steps > c->probe_base)
        LBM_HIP_CHECK(launch_probe_cell(c->storage, c->g, c->st, c->cur, c->bc.view, c->probe,
                                        c->steps - c->probe_base - 1, c->stream),
                      "probe kernel");
    return LBM_OK;
}

int lbm_get_state(lbm_ctx *c, double *f, double *rho, double *u, int where) {
    int rc = check_ctx(c, "lbm_get_state");
    if (rc) return rc;
    if (c->phase == kEmpty) {
        set_error("lbm_get_state: no state loaded");
        return LBM_ERR_STATE;
    }
    const size_t N = c->g.N;
    if (c->phase == kRunning) {
        if (where == LBM_DEVICE) {
            rc = finalize_into(c, f, rho, u);
            if (rc) return rc;
            LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_get_state");
            return LBM_OK;
        }
        rc = finalize_into(c, f ? c->stage_f : nullptr, rho ? c->stage_rho : nullptr, u ? c->stage_u : nullptr);
        if (rc) return rc;
    }
    // staging now holds f_n (PENDING: the loaded input)
    if (f) LBM_HIP_CHECK(hipMemcpyAsync(f, c->stage_f, 9 * N * sizeof(double), kind_out(where), c->stream), "lbm_get_state");
    if (rho) LBM_HIP_CHECK(hipMemcpyAsync(rho, c->stage_rho, N * sizeof(double), kind_out(where), c->stream), "lbm_get_state");
    if (u) LBM_HIP_CHECK(hipMemcpyAsync(u, c->stage_u, 2 * N * sizeof(double), kind_out(where), c->stream), "lbm_get_state");
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_get_state");
    return LBM_OK;
}

int lbm_set_probe(lbm_ctx *c, int px, int py, int64_t capacity) {
    int rc = check_ctx(c, "lbm_set_probe");
    if (rc) return rc;
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_set_probe");
    (void)hipFree(c->probe.ring);
    c->probe = ProbeView{};
    if (capacity <= 0) return LBM_OK;
    if (px < 0 || px >= c->nx || py < 0 || py >= c->ny) {
        set_error("lbm_set_probe: probe node outside the lattice");
        return LBM_ERR_ARG;
    }
    LBM_HIP_CHECK(hipMalloc(&c->probe.ring, 3 * (size_t)capacity * sizeof(double)), "lbm_set_probe");
    c->probe.capacity = capacity;
    c->probe.cell = (int64_t)px * c->ny + py;
    c->probe_base = c->steps;
    return LBM_OK;
}

int lbm_read_probe(lbm_ctx *c, double *out, int64_t max, int64_t *count) {
    int rc = check_ctx(c, "lbm_read_probe");
    if (rc) return rc;
    if (count) *count = 0;
    if (c->probe.cell < 0) {
        set_error("lbm_read_probe: no probe set");
        return LBM_ERR_STATE;
    }
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_read_probe");
    const int64_t recorded = c->steps - c->probe_base;
    const int64_t avail = recorded < c->probe.capacity ? recorded : c->probe.capacity;
    const int64_t n = avail < max ? avail : max;
    if (n <= 0 || !out) return LBM_OK;
    // oldest kept sample first; copy the tail of the window
    const int64_t first = recorded - n;
    for (int64_t k = 0; k < n;) {
        const int64_t slot = (first + k) % c->probe.capacity;
        const int64_t run = std::min<int64_t>(n - k, c->probe.capacity - slot);
        LBM_HIP_CHECK(hipMemcpy(out + 3 * k, c->probe.ring + 3 * slot, 3 * (size_t)run * sizeof(double),
                                hipMemcpyDeviceToHost),
                      "lbm_read_probe");
        k += run;
    }
    if (count) *count = n;
    return LBM_OK;
}

int lbm_moments_sum(lbm_ctx *c, double out[3]) {
    int rc = check_ctx(c, "lbm_moments_sum");
    if (rc) return rc;
    if (!out) {
        set_error("lbm_moments_sum: out is NULL");
        return LBM_ERR_ARG;
    }
    int nb = 0;
    if (c->phase == kRunning) {
        LBM_HIP_CHECK(launch_finalize(c->storage, c->g, c->st, c->cur, c->bc.view, nullptr, nullptr, nullptr,
                                      c->partials, c->stream, &nb),
                      "moment-sum kernel");
    } else if (c->phase == kPending) {
        LBM_HIP_CHECK(launch_moments_sum_aos(c->stage_f, (int64_t)c->g.N, c->partials, c->stream, &nb),
                      "moment-sum kernel");
    } else {
        set_error("lbm_moments_sum: no state loaded");
        return LBM_ERR_STATE;
    }
    LBM_HIP_CHECK(launch_sum_partials(c->partials, nb, c->sums, c->stream), "partial-sum kernel");
    LBM_HIP_CHECK(hipMemcpyAsync(out, c->sums, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                  "lbm_moments_sum");
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_moments_sum");
    if (!(std::isfinite(out[0]) && std::isfinite(out[1]) && std::isfinite(out[2]))) {
        set_error("lbm_moments_sum: non-finite moments (the run is unstable)");
        return LBM_ERR_UNSTABLE;
    }
    return LBM_OK;
}

int lbm_speed(lbm_ctx *c, double *out, int where) {
    int rc = check_ctx(c, "lbm_speed");
    if (rc) return rc;
    if (!out) {
        set_error("lbm_speed: out is NULL");
        return LBM_ERR_ARG;
    }
    if (c->phase == kEmpty) {
        set_error("lbm_speed: no state loaded");
        return LBM_ERR_STATE;
    }
    const size_t N = c->g.N;
    if (c->phase == kRunning) {
        // staging is scratch while running: velocities into stage_u
        rc = finalize_into(c, nullptr, nullptr, c->stage_u);
        if (rc) return rc;
    }
    double *dst = where == LBM_DEVICE ? out : c->stage_rho;
    if (where == LBM_HOST && c->phase == kPending) {
        // keep the pending input's density: compute into a temporary
        double *tmp = nullptr;
        LBM_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&tmp), N * sizeof(double)), "lbm_speed");
        hipError_t e = launch_speed(c->stage_u, tmp, (int64_t)N, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, N * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        (void)hipFree(tmp);
        LBM_HIP_CHECK(e, "lbm_speed");
        return LBM_OK;
    }
    LBM_HIP_CHECK(launch_speed(c->stage_u, dst, (int64_t)N, c->stream), "speed kernel");
    if (where == LBM_HOST)
        LBM_HIP_CHECK(hipMemcpyAsync(out, dst, N * sizeof(double), hipMemcpyDeviceToHost, c->stream), "lbm_speed");
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_speed");
    return LBM_OK;
}

// Checkpoint layout: a 64-byte header, then the state as the context holds it
// (running: populations of the storage type and the outlet buffer; pending:
// the FP64 step input f, rho, u).
namespace {
struct CkptHeader {
    char magic[8];
    int32_t version, nx, ny, storage, phase, reserved;
    int64_t steps, payload;
    int64_t pad[2];
};
static_assert(sizeof(CkptHeader) == 64, "checkpoint header is 64 bytes");
constexpr char kCkptMagic[8] = {'L', 'B', 'M', 'S', 'T', 'A', 'T', '1'};
// 2: plane rows padded to whole 128-byte lines
constexpr int32_t kCkptVersion = 2;

int64_t ckpt_payload(const lbm_ctx *c) {
    const int64_t N = c->g.N;
    if (c->phase == kRunning) return (int64_t)plane_bytes(c) + 3 * (int64_t)c->ny * (int64_t)side_elem(c->storage);
    return 12 * N * (int64_t)sizeof(double);
}
}  // namespace

int lbm_checkpoint_bytes(lbm_ctx *c, int64_t *bytes) {
    int rc = check_ctx(c, "lbm_checkpoint_bytes");
    if (rc) return rc;
    if (!bytes) {
        set_error("lbm_checkpoint_bytes: bytes is NULL");
        return LBM_ERR_ARG;
    }
    if (c->phase == kEmpty) {
        set_error("lbm_checkpoint_bytes: no state loaded");
        return LBM_ERR_STATE;
    }
    *bytes = (int64_t)sizeof(CkptHeader) + ckpt_payload(c);
    return LBM_OK;
}

int lbm_checkpoint_save(lbm_ctx *c, void *host, int64_t capacity) {
    int rc = check_ctx(c, "lbm_checkpoint_save");
    if (rc) return rc;
    if (c->phase == kEmpty) {
        set_error("lbm_checkpoint_save: no state loaded");
        return LBM_ERR_STATE;
    }
    const int64_t need = (int64_t)sizeof(CkptHeader) + ckpt_payload(c);
    if (!host || capacity < need) {
        set_error("lbm_checkpoint_save: buffer NULL or smaller than lbm_checkpoint_bytes");
        return LBM_ERR_ARG;
    }
    CkptHeader h{};
    std::memcpy(h.magic, kCkptMagic, 8);
    h.version = kCkptVersion;
    h.nx = c->nx;
    h.ny = c->ny;
    h.storage = c->storage;
    h.phase = c->phase;
    h.steps = c->steps;
    h.payload = ckpt_payload(c);
    std::memcpy(host, &h, sizeof h);
    char *dst = static_cast<char *>(host) + sizeof h;
    const size_t N = c->g.N;
    if (c->phase == kRunning) {
        const size_t pb = plane_bytes(c), sb = 3 * (size_t)c->ny * side_elem(c->storage);
        LBM_HIP_CHECK(hipMemcpyAsync(dst, c->st.plane[c->cur], pb, hipMemcpyDeviceToHost, c->stream), "lbm_checkpoint_save");
        LBM_HIP_CHECK(hipMemcpyAsync(dst + pb, c->st.side[c->cur], sb, hipMemcpyDeviceToHost, c->stream), "lbm_checkpoint_save");
    } else {
        LBM_HIP_CHECK(hipMemcpyAsync(dst, c->stage_f, 9 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream), "lbm_checkpoint_save");
        LBM_HIP_CHECK(hipMemcpyAsync(dst + 9 * N * sizeof(double), c->stage_rho, N * sizeof(double), hipMemcpyDeviceToHost, c->stream), "lbm_checkpoint_save");
        LBM_HIP_CHECK(hipMemcpyAsync(dst + 10 * N * sizeof(double), c->stage_u, 2 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream), "lbm_checkpoint_save");
    }
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_checkpoint_save");
    return LBM_OK;
}

int lbm_checkpoint_load(lbm_ctx *c, const void *host, int64_t bytes) {
    int rc = check_ctx(c, "lbm_checkpoint_load");
    if (rc) return rc;
    if (!host || bytes < (int64_t)sizeof(CkptHeader)) {
        set_error("lbm_checkpoint_load: buffer NULL or shorter than the header");
        return LBM_ERR_ARG;
    }
    CkptHeader h;
    std::memcpy(&h, host, sizeof h);
    if (std::memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != kCkptVersion) {
        set_error("lbm_checkpoint_load: not a checkpoint of this library (bad magic or version)");
        return LBM_ERR_ARG;
    }
    if (h.nx != c->nx || h.ny != c->ny || h.storage != c->storage) {
        set_error("lbm_checkpoint_load: checkpoint is " + std::to_string(h.nx) + "x" + std::to_string(h.ny) +
                  " storage " + std::to_string(h.storage) + ", context is " + std::to_string(c->nx) + "x" +
                  std::to_string(c->ny) + " storage " + std::to_string(c->storage));
        return LBM_ERR_ARG;
    }
    if ((h.phase != kRunning && h.phase != kPending) || h.steps < 0) {
        set_error("lbm_checkpoint_load: corrupt header (phase or step count)");
        return LBM_ERR_ARG;
    }
    const int saved_phase = c->phase;
    c->phase = h.phase;
    const int64_t payload = ckpt_payload(c);
    c->phase = saved_phase;
    if (h.payload != payload || bytes < (int64_t)sizeof h + payload) {
        set_error("lbm_checkpoint_load: payload size does not match the lattice");
        return LBM_ERR_ARG;
    }
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_checkpoint_load");
    const char *src = static_cast<const char *>(host) + sizeof h;
    const size_t N = c->g.N;
    if (h.phase == kRunning) {
        const size_t pb = plane_bytes(c), sb = 3 * (size_t)c->ny * side_elem(c->storage);
        LBM_HIP_CHECK(hipMemcpyAsync(c->st.plane[0], src, pb, hipMemcpyHostToDevice, c->stream), "lbm_checkpoint_load");
        LBM_HIP_CHECK(hipMemcpyAsync(c->st.side[0], src + pb, sb, hipMemcpyHostToDevice, c->stream), "lbm_checkpoint_load");
        c->cur = 0;
    } else {
        LBM_HIP_CHECK(hipMemcpyAsync(c->stage_f, src, 9 * N * sizeof(double), hipMemcpyHostToDevice, c->stream), "lbm_checkpoint_load");
        LBM_HIP_CHECK(hipMemcpyAsync(c->stage_rho, src + 9 * N * sizeof(double), N * sizeof(double), hipMemcpyHostToDevice, c->stream), "lbm_checkpoint_load");
        LBM_HIP_CHECK(hipMemcpyAsync(c->stage_u, src + 10 * N * sizeof(double), 2 * N * sizeof(double), hipMemcpyHostToDevice, c->stream), "lbm_checkpoint_load");
    }
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_checkpoint_load");
    c->phase = h.phase;
    c->steps = h.steps;
    c->probe_base = c->steps;
    return LBM_OK;
}

int lbm_synchronize(lbm_ctx *c) {
    int rc = check_ctx(c, "lbm_synchronize");
    if (rc) return rc;
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_synchronize");
    return LBM_OK;
}

int lbm_steps_done(const lbm_ctx *c, int64_t *steps) {
    if (!c || !steps) {
        set_error("lbm_steps_done: NULL argument");
        return LBM_ERR_ARG;
    }
    *steps = c->steps;
    return LBM_OK;
}

// ---- batches ---------------------------------------------------------------

}  // extern "C"

struct lbm_batch {
    std::vector<lbm_ctx *> ctxs;
    std::vector<double> omegas;
    int device = 0, storage = 0;
    hipStream_t stream = nullptr;
    int v = 1;
    bool use_graph = true;
    BatchJob *d_jobs[2] = {nullptr, nullptr};
    uint32_t *d_block_job = nullptr;
    size_t block_cap = 0;
    unsigned nblocks = 0;
    int64_t *d_step = nullptr;
    unsigned *d_done = nullptr;
    int64_t step_host = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned graph_blocks = 0;
    int graph_v = 0;
    // two steps per launch for the jobs the sweep kernel supports
    int temporal = 1;
    int sweep_rows = 8;
    bool rows_auto = true;  // rows per block not given: fewer rows for small batches
    // one packed launch for every job of one column per thread (FP32/FP16):
    // 0 off (per-width groups, the faster on MI355X), 1 one block per packed
    // block, 2 persistent blocks
    int packed = 0;
    void *d_sweep = nullptr;  // sweep job tables and block tables of the last run
    size_t sweep_cap = 0;
    // side streams (default priority) that run the groups of a large batch
    // side by side, with the events that fork them from and join them to
    // `stream`
    std::vector<hipStream_t> lanes;
    std::vector<hipEvent_t> lane_done;
    hipEvent_t fork = nullptr;
    // pinned host staging of the run's tables: the copies to the device read
    // it asynchronously, and it is rewritten only once the batch stream and
    // every side stream have drained (a run starts and ends that way)
    void *h_stage = nullptr;
    size_t h_stage_cap = 0;
    TableImage image;
    // set when this run has enqueued work on the contexts' state
    bool dirty = false;
};

namespace {

void drop_graph(lbm_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_blocks = 0;
}

// Wait for everything the batch has enqueued: its stream and every side stream.
hipError_t quiesce(lbm_batch *b) {
    hipError_t first = hipSuccess;
    if (b->stream) first = hipStreamSynchronize(b->stream);
    for (hipStream_t l : b->lanes) {
        const hipError_t e = hipStreamSynchronize(l);
        if (first == hipSuccess) first = e;
    }
    return first;
}

// Copy the packed tables into the batch's pinned staging buffer (grown as
// needed); the caller has drained the batch's streams, so no copy reads it.
hipError_t stage_tables(lbm_batch *b) {
    const size_t bytes = b->image.bytes.size();
    if (bytes > b->h_stage_cap) {
        if (b->h_stage) (void)hipHostFree(b->h_stage);
        b->h_stage = nullptr;
        b->h_stage_cap = 0;
        const hipError_t e = hipHostMalloc(&b->h_stage, bytes, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        b->h_stage_cap = bytes;
    }
    if (bytes) std::memcpy(b->h_stage, b->image.bytes.data(), bytes);
    return hipSuccess;
}

void release_batch(lbm_batch *b) {
    (void)hipSetDevice(b->device);
    (void)quiesce(b);
    drop_graph(b);
    (void)hipFree(b->d_jobs[0]);
    (void)hipFree(b->d_jobs[1]);
    (void)hipFree(b->d_block_job);
    (void)hipFree(b->d_step);
    (void)hipFree(b->d_done);
    (void)hipFree(b->d_sweep);
    for (hipStream_t l : b->lanes) {
        (void)hipStreamSynchronize(l);
        (void)hipStreamDestroy(l);
    }
    for (hipEvent_t e : b->lane_done) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
}

// Whether three steps per launch beat two for FP32 rows of `waves`
// wavefronts (measured on 8192-row channels: blocks of 7-8 and of 11 or more
// waves gain 10-36 %; 9-10 lose or tie, 6 and fewer lose or tie).
bool three_steps_pay(int storage, unsigned waves) {
    return storage == LBM_STORE_F32 && (waves == 7 || waves == 8 || waves >= 11);
}

// Side streams of a batch, and the batch size (nodes) from which a batch
// whose jobs all fall into one three-step group runs it on a side stream
// (several groups take side streams at any size).
constexpr size_t kBatchLanes = 4;
constexpr int64_t kBatchLaneNodes = 8 << 20;

// At least `k` side streams (and their join events) for the batch.
hipError_t ensure_lanes(lbm_batch *b, size_t k) {
    if (!b->fork) {
        hipError_t e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    while (b->lanes.size() < k) {
        hipStream_t l = nullptr;
        hipError_t e = hipStreamCreateWithFlags(&l, hipStreamNonBlocking);
        if (e != hipSuccess) return e;
        b->lanes.push_back(l);
        hipEvent_t ev = nullptr;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return e;
        b->lane_done.push_back(ev);
    }
    return hipSuccess;
}

// Steps of a batch with the two-step sweep kernel: jobs it supports are
// grouped by block width (one launch per group per step pair), the others
// take two one-step batched launches per pair.  A batch of several groups
// runs each group on a side stream; otherwise graph-captured pairs of pairs
// replay the whole run on the batch stream.  An odd last step goes job by
// job.
int batch_run_sweep(lbm_batch *b, int64_t nsteps) {
    const int n = (int)b->ctxs.size();
    struct Bucket {
        int v = 1;
        unsigned threads = 0;
        bool big = false;  // state buffers of 4 GiB or more
        int k = 2;         // 3: three steps per launch where the group runs on a side stream
        size_t lds = 0;    // of the two-step kernel
        size_t lds_k = 0;  // of the three-step kernel
        std::vector<int> members;
        unsigned nblocks = 0;  // blocks of the group's launch
    };
    std::vector<Bucket> buckets;
    std::vector<int> fjobs;  // one-step jobs
    std::vector<int> pjobs;  // jobs of the packed launch
    const bool can_pack = b->packed != 0 && b->storage != LBM_STORE_F64;
    // A small batch has few rows to share out: with the row counts chosen
    // for large batches its launches would hold fewer blocks than the GPU
    // takes at once.  With automatic rows, no group gets more rows per block
    // than leave, over the whole batch, about four blocks per CU for rows of
    // fewer than 512 threads and two for wider rows (channels of 0.77 M nodes
    // in all: 52.5 GLUPS at 8 rows, 86.6 at 2; of 2.1 M nodes with one
    // 514-column channel: 64.4 GLUPS at one block per CU for its rows, 86.1
    // at two), unless the wide groups hold most of the batch (below).  Three
    // steps per launch, whose halo rows weigh more in short blocks, wait for
    // 8 such rows (at 2 rows: 68.2 GLUPS against 86.6 with two steps per
    // launch).
    int64_t rows_total = 0;
    int cus = 0;
    bool three_fill = true;
    if (b->rows_auto) {
        for (int j = 0; j < n; j++) rows_total += b->ctxs[j]->g.lx;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        three_fill = batch_three_steps_fill(rows_total, cus);
    }
    for (int j = 0; j < n; j++) {
        lbm_ctx *c = b->ctxs[j];
        if (!c->st.scratch || !sweep_batch_supported(c->storage, c->g, c->bc.view)) {
            fjobs.push_back(j);
            continue;
        }
        const SweepShape sh = sweep_shape(c->storage, c->g.ly);  // whole rows here
        const int v = sh.v;
        const unsigned threads = sh.threads;
        const bool big = sweep_big(c->storage, c->g);
        if (can_pack && v == 1 && !big && (c->g.ly + 63) / 64 <= packed_block_waves()) {
            pjobs.push_back(j);
            continue;
        }
        const int k = b->temporal == 3 && three_fill && v == 1 && !big && three_steps_pay(c->storage, threads / 64) &&
                              sweepk_supported(c->storage, c->g, c->bc.view, 3)
                          ? 3
                          : 2;
        Bucket *bk = nullptr;
        for (Bucket &x : buckets)
            if (x.v == v && x.threads == threads && x.big == big && x.k == k) bk = &x;
        if (!bk) {
            buckets.emplace_back();
            bk = &buckets.back();
            bk->v = v;
            bk->threads = threads;
            bk->big = big;
            bk->k = k;
        }
        bk->members.push_back(j);
    }
    // Wide rows (blocks of 512 threads or more) leave one to three blocks per
    // CU, so their launches run in rounds of the blocks the GPU holds at
    // once, and a last round that is barely filled costs as much as a full
    // one.  With automatic rows the wide two-step groups share the row count
    // that minimises rounds x (rows + 3) over all their blocks together (a
    // block also reads and computes about three halo rows).  Measured, FP32
    // GLUPS at 32 rows / at the chosen rows: four 4200 x 1800 channels 108.3 /
    // 131.8 (66), one 2048 x 2048 lattice 50.7 / 115.2 (8), two 8192 x 2048
    // channels with a 2048 and a 1500 square 124.0 / 133.1 (79), one 8192 x
    // 2048 channel 133.0 (32 chosen; 34 gives 136.3).  Where the wide groups
    // hold less than half of the swept nodes their blocks share the GPU with
    // the narrow groups' and the rule for those applies to them too.
    int wide_rows = 0;
    int64_t wide_nodes = 0, swept_nodes = 0;
    for (const Bucket &bk : buckets)
        for (int j : bk.members) {
            swept_nodes += b->ctxs[j]->g.N;
            if (bk.k == 2 && bk.threads >= 512) wide_nodes += b->ctxs[j]->g.N;
        }
    if (b->rows_auto && 2 * wide_nodes >= swept_nodes) {
        std::vector<WideLattice> wide;
        for (const Bucket &bk : buckets) {
            if (bk.k != 2 || bk.threads < 512) continue;
            size_t lds64 = 0;
            for (int j : bk.members) {
                const lbm_ctx *c = b->ctxs[j];
                lds64 = std::max(lds64, sweep_lds_bytes(c->storage, c->g.ly, c->bc.view.chunk ? c->bc.view.nclasses : 0, 64));
            }
            const int per_cu = sweep2_batch_per_cu(b->storage, bk.v, bk.big, bk.threads, lds64);
            for (int j : bk.members) wide.push_back({b->ctxs[j]->g.lx, per_cu});
        }
        wide_rows = pick_wide_rows(wide, cus, 128, 3);
    }
    // tables: sweep jobs per bucket and pair parity with their block tables,
    // then the one-step jobs' records for the two steps of a pair
    std::vector<std::pair<int, uint32_t>> ilv;  // (job, rows) in the interleaved row order
    std::vector<SweepGroupTables> gt(buckets.size());
    for (size_t bi = 0; bi < buckets.size(); bi++) {
        Bucket &bk = buckets[bi];
        SweepGroupTables &t = gt[bi];
        for (size_t m = 0; m < bk.members.size(); m++) {
            lbm_ctx *c = b->ctxs[bk.members[m]];
            // short rows (blocks of a few waves) need many blocks in flight:
            // few rows per block; wide rows (512+ threads per block) take 32
            // three steps per launch read three halo rows on each side: four
            // times the rows per block (the 256-channel batch with its
            // three-step groups at 8, 16, 24, 32 and 48 rows: 132, 143, 147,
            // 149 and 148 GLUPS)
            LaunchCfg cfg = c->cfg;
            cfg.sweep_rows =
                batch_group_rows((int)bk.threads, bk.k, b->sweep_rows, b->rows_auto, rows_total, cus, wide_rows);
            uint32_t rows = (uint32_t)sweep_rows_for(c->g, cfg, 0);
            // the three-step kernel needs two rows in every block; the same
            // tables also serve its two-step launches
            if (bk.k == 3) rows = (uint32_t)sweepk_rows(c->g, (int)rows, 3);
            const uint32_t nb = (c->g.lx + rows - 1) / rows;
            const uint32_t block0 = (uint32_t)t.block_job.size();
            t.block_job.insert(t.block_job.end(), nb, (uint32_t)m);
            const uint32_t ncl = c->bc.view.chunk ? c->bc.view.nclasses : 0;
            bk.lds = std::max(bk.lds, sweep_lds_bytes(c->storage, c->g.ly, ncl, rows));
            if (bk.k == 3) bk.lds_k = std::max(bk.lds_k, sweepk_lds_bytes(c->storage, c->g.ly, ncl, rows, 3));
            // FP32 rows of two columns per thread in a long run take the
            // range-interleaved row order, as in lbm_run: moved there now (on
            // the batch stream, ahead of every launch of the run) and back
            // after the pairs
            RowLayout il;
            if (bk.k == 2 && bk.v == 2 && c->storage == LBM_STORE_F32 && c->cfg.row_interleave &&
                nsteps / 2 >= kInterleaveMinPairs && rows >= (uint32_t)kInterleaveMinRows && rows < c->g.lx &&
                c->g.lx < kRowLayoutMaxLx) {
                il = row_layout(c->g.lx, rows);
                LBM_HIP_CHECK(launch_row_layout(c->storage, c->g, c->st, c->cur, rows, true, b->stream),
                              "lbm_batch_run: row order kernel");
                c->cur ^= 1;
                ilv.push_back({bk.members[m], rows});
            }
            for (int pp = 0; pp < 2; pp++) {
                SweepJob d;
                d.il_rows = il.rows;
                d.il_ranges = il.ranges;
                d.il_whole = il.whole;
                d.il_magic = il.magic;
                const int src = c->cur ^ pp;
                d.src = c->st.plane[src];
                d.dst = c->st.plane[src ^ 1];
                d.side_prev = c->st.side[src];
                d.side_next = c->st.side[src ^ 1];
                d.scratch = c->st.scratch;
                d.g = c->g;
                d.bc = c->bc.view;
                d.omega = b->omegas[bk.members[m]];
                d.rows = rows;
                d.block0 = block0;
                if (c->probe.cell >= 0) {
                    d.probe_cell = c->probe.cell;
                    d.probe_ring = c->probe.ring;
                    d.probe_cap = c->probe.capacity;
                    d.probe_offset = c->steps - c->probe_base - 1 - b->step_host;
                }
                t.jobs[pp].push_back(d);
            }
        }
        bk.nblocks = (unsigned)t.block_job.size();
    }
    StepTables st;
    for (size_t m = 0; m < fjobs.size(); m++) {
        lbm_ctx *c = b->ctxs[fjobs[m]];
        const uint32_t nb = (uint32_t)batch_job_blocks(c->g, b->v);
        const uint32_t block0 = (uint32_t)st.block_job.size();
        st.block_job.insert(st.block_job.end(), nb, (uint32_t)m);
        for (int p = 0; p < 2; p++) {
            BatchJob d;
            const int src = c->cur ^ p;
            d.src = c->st.plane[src];
            d.dst = c->st.plane[src ^ 1];
            d.side_prev = c->st.side[src];
            d.side_next = c->st.side[src ^ 1];
            d.g = c->g;
            d.bc = c->bc.view;
            d.omega = b->omegas[fjobs[m]];
            d.block0 = block0;
            if (c->probe.cell >= 0) {
                d.probe_cell = c->probe.cell;
                d.probe_ring = c->probe.ring;
                d.probe_cap = c->probe.capacity;
                d.probe_offset = c->steps - c->probe_base - 1 - b->step_host;
            }
            st.jobs[p].push_back(d);
        }
    }
    const unsigned nfblocks = (unsigned)st.block_job.size();
    // the packed launch: row ranges of `rows` rows, placed side by side in
    // blocks of packed_block_waves() waves and at most half the CU's LDS (two
    // blocks per CU), or more where one range needs it
    PackedTables pt;
    size_t plds = 0;
    if (!pjobs.empty()) {
        std::vector<PackItem> items;
        uint32_t budget = 80u << 10;
        for (size_t m = 0; m < pjobs.size(); m++) {
            lbm_ctx *c = b->ctxs[pjobs[m]];
            const uint32_t rows = (uint32_t)std::max(1, b->sweep_rows);
            const uint32_t waves = (c->g.ly + 63) / 64;
            const uint32_t ncl = c->bc.view.chunk ? c->bc.view.nclasses : 0;
            for (uint32_t xa = 0; xa < c->g.lx; xa += rows) {
                PackItem it;
                it.job = (uint32_t)m;
                it.xa = xa;
                it.rows = std::min(rows, c->g.lx - xa);
                it.waves = waves;
                it.lds = packed_sub_lds(c->storage, c->g.ly, ncl, it.rows);
                budget = std::max(budget, it.lds);
                items.push_back(it);
            }
            for (int pp = 0; pp < 2; pp++) {
                SweepJob d;
                const int src = c->cur ^ pp;
                d.src = c->st.plane[src];
                d.dst = c->st.plane[src ^ 1];
                d.side_prev = c->st.side[src];
                d.side_next = c->st.side[src ^ 1];
                d.scratch = c->st.scratch;
                d.g = c->g;
                d.bc = c->bc.view;
                d.omega = b->omegas[pjobs[m]];
                d.rows = rows;
                if (c->probe.cell >= 0) {
                    d.probe_cell = c->probe.cell;
                    d.probe_ring = c->probe.ring;
                    d.probe_cap = c->probe.capacity;
                    d.probe_offset = c->steps - c->probe_base - 1 - b->step_host;
                }
                pt.jobs[pp].push_back(d);
            }
        }
        if (!plan_packed_blocks(items, packed_block_waves(), budget, pt)) {
            set_error("lbm_batch_run: packing the row ranges failed");
            return LBM_ERR_ARG;
        }
        for (const PackedBlock &pb : pt.blocks) plds = std::max(plds, (size_t)pb.lds);
        LBM_HIP_CHECK(prepare_sweep2_packed(b->storage, plds), "lbm_batch_run: packed kernel LDS");
    }
    const unsigned npblocks = (unsigned)pt.blocks.size();
    unsigned pgrid = 0;
    if (npblocks)
        LBM_HIP_CHECK(packed_grid(b->storage, plds, npblocks, b->packed == 2, &pgrid), "lbm_batch_run: packed grid");
    pack_batch_tables(gt, pt, st, b->image);
    const TableImage &img = b->image;
    for (const Bucket &bk : buckets) {
        LBM_HIP_CHECK(prepare_sweep2_batch(b->storage, bk.v, bk.big, bk.lds), "lbm_batch_run: sweep kernel LDS");
        if (bk.k == 3)
            LBM_HIP_CHECK(prepare_sweepk_batch(b->storage, 3, bk.lds_k), "lbm_batch_run: three-step kernel LDS");
    }
    drop_graph(b);  // the one-step graph of an earlier run would hold stale tables
    if (img.bytes.size() > b->sweep_cap) {
        (void)hipFree(b->d_sweep);
        b->d_sweep = nullptr;
        b->sweep_cap = 0;
        LBM_HIP_CHECK(hipMalloc(&b->d_sweep, img.bytes.size()), "lbm_batch_run: sweep tables");
        b->sweep_cap = img.bytes.size();
    }
    char *dev = static_cast<char *>(b->d_sweep);
    LBM_HIP_CHECK(stage_tables(b), "lbm_batch_run: pinned staging of the tables");
    LBM_HIP_CHECK(hipMemcpyAsync(dev, b->h_stage, img.bytes.size(), hipMemcpyHostToDevice, b->stream),
                  "lbm_batch_run: sweep tables");
    // A pair, in order on the batch stream: the groups (they read the pair
    // counter d_step[1]), then the one-step jobs' two launches (they advance
    // d_step[0] themselves), then the pair counter.
    auto pair = [&](int pp) -> hipError_t {
        hipError_t e = hipSuccess;
        if (npblocks)
            e = launch_sweep2_packed(b->storage, reinterpret_cast<const SweepJob *>(dev + img.packed_jobs[pp]),
                                     reinterpret_cast<const PackedSub *>(dev + img.packed_subs),
                                     reinterpret_cast<const PackedBlock *>(dev + img.packed_blocks), npblocks, pgrid,
                                     plds, b->d_step + 1, 0, b->stream);
        for (size_t i = 0; i < buckets.size() && e == hipSuccess; i++) {
            const Bucket &bk = buckets[i];
            e = launch_sweep2_batch(b->storage, bk.v, bk.big, reinterpret_cast<const SweepJob *>(dev + img.group_jobs[pp][i]),
                                    reinterpret_cast<const uint32_t *>(dev + img.group_blocks[i]), bk.nblocks,
                                    bk.threads, bk.lds, b->d_step + 1, 0, b->stream);
        }
        if (e == hipSuccess && !fjobs.empty())
            for (int p = 0; p < 2 && e == hipSuccess; p++)
                e = launch_fused_batch(b->storage, b->v, reinterpret_cast<const BatchJob *>(dev + img.step_jobs[p]),
                                       reinterpret_cast<const uint32_t *>(dev + img.step_blocks), nfblocks,
                                       b->d_step, b->d_done, b->stream);
        if (e == hipSuccess) e = launch_add_step(b->d_step + 1, 2, b->stream);
        if (e == hipSuccess && fjobs.empty()) e = launch_add_step(b->d_step, 2, b->stream);
        return e;
    };
    const int64_t pairs = nsteps / 2;
    int64_t done_pairs = 0;
    // groups of three steps per launch advance the pairs' 2 * pairs steps in
    // q launches and a remainder of r steps (side streams only)
    const int64_t q3 = 2 * pairs / 3, r3 = 2 * pairs - 3 * q3;
    bool three_ran = false;
    // A large batch of several groups: each group (and the one-step jobs) on
    // a side stream of its own, so that the tail of one group's launch
    // overlaps the next group instead of idling the GPU (one stream runs its
    // launches strictly one after another).  Groups touch disjoint lattices;
    // the groups' probe slots come from the pair index by value, and the
    // one-step jobs keep their counter on their own stream.
    int64_t nodes = 0;
    for (lbm_ctx *c : b->ctxs) nodes += c->g.N;
    // groups: the buckets, then the packed launch, then the one-step jobs
    const size_t gpacked = buckets.size(), gstep = buckets.size() + (npblocks ? 1 : 0);
    const size_t groups = gstep + (fjobs.empty() ? 0 : 1);
    // Several groups run on side streams at any batch size: graph replay of
    // the pairs on one stream measured no faster than plain launches on this
    // GPU, and one stream leaves each group's small launch alone on the GPU
    // (4 uneven channels: 13.5 GLUPS on one stream, 52.5 on side streams; 16:
    // 35.6 and 83.0).  Three-step groups run only on side streams (their
    // launch count differs from the pairs'), so a large batch holding one
    // takes them even with a single group.
    bool any_three = false;
    for (const Bucket &bk : buckets) any_three |= bk.k == 3;
    if ((groups > 1 || (any_three && nodes >= (int64_t)kBatchLaneNodes)) && pairs >= 1) {
        const size_t nl = std::min(groups, (size_t)kBatchLanes);
        LBM_HIP_CHECK(ensure_lanes(b, nl), "lbm_batch_run: side streams");
        // groups onto lanes, largest first onto the least loaded
        std::vector<std::pair<int64_t, size_t>> load;  // (nodes, group)
        for (size_t i = 0; i < buckets.size(); i++) {
            int64_t g = 0;
            for (int j : buckets[i].members) g += b->ctxs[j]->g.N;
            load.push_back({g, i});
        }
        if (npblocks) {
            int64_t g = 0;
            for (int j : pjobs) g += b->ctxs[j]->g.N;
            load.push_back({g, gpacked});
        }
        if (!fjobs.empty()) {
            int64_t g = 0;
            for (int j : fjobs) g += b->ctxs[j]->g.N;
            load.push_back({g, gstep});
        }
        std::sort(load.begin(), load.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
        std::vector<int64_t> lane_load(nl, 0);
        std::vector<size_t> lane_of(groups, 0);
        for (const auto &gl : load) {
            const size_t l = (size_t)(std::min_element(lane_load.begin(), lane_load.end()) - lane_load.begin());
            lane_of[gl.second] = l;
            lane_load[l] += gl.first;
        }
        LBM_HIP_CHECK(hipEventRecord(b->fork, b->stream), "lbm_batch_run: fork");
        for (size_t l = 0; l < nl; l++) LBM_HIP_CHECK(hipStreamWaitEvent(b->lanes[l], b->fork, 0), "lbm_batch_run: fork");
        for (size_t i = 0; i < buckets.size(); i++) {
            // a three-step group: its own sequence of launches on its lane
            const Bucket &bk = buckets[i];
            if (bk.k != 3) continue;
            hipStream_t lane = b->lanes[lane_of[i]];
            for (int64_t t = 0; t < q3; t++)
                LBM_HIP_CHECK(launch_sweepk_batch(b->storage, 3,
                                                  reinterpret_cast<const SweepJob *>(dev + img.group_jobs[t & 1][i]),
                                                  reinterpret_cast<const uint32_t *>(dev + img.group_blocks[i]),
                                                  bk.nblocks, bk.threads, bk.lds_k, nullptr, b->step_host + 3 * t, lane),
                              "batched three-step sweep");
            if (r3 == 2) {
                LBM_HIP_CHECK(launch_sweep2_batch(b->storage, bk.v, bk.big,
                                                  reinterpret_cast<const SweepJob *>(dev + img.group_jobs[q3 & 1][i]),
                                                  reinterpret_cast<const uint32_t *>(dev + img.group_blocks[i]),
                                                  bk.nblocks, bk.threads, bk.lds, nullptr, b->step_host + 3 * q3, lane),
                              "batched sweep after three-step launches");
            } else if (r3 == 1) {
                for (int j : bk.members) {
                    lbm_ctx *c = b->ctxs[j];
                    const int64_t slot = c->probe.cell >= 0 ? c->steps + 3 * q3 - c->probe_base - 1 : -1;
                    LBM_HIP_CHECK(launch_fused(c->storage, c->g, c->st, c->cur ^ (int)(q3 & 1), c->bc.view, b->omegas[j],
                                               c->probe, slot, nullptr, c->cfg, lane, nullptr),
                                  "step after three-step launches");
                }
            }
        }
        three_ran = true;
        for (int64_t k = 0; k < pairs; k++) {
            const int pp = (int)(k & 1);
            for (size_t i = 0; i < buckets.size(); i++) {
                const Bucket &bk = buckets[i];
                if (bk.k == 3) continue;
                LBM_HIP_CHECK(launch_sweep2_batch(b->storage, bk.v, bk.big,
                                                  reinterpret_cast<const SweepJob *>(dev + img.group_jobs[pp][i]),
                                                  reinterpret_cast<const uint32_t *>(dev + img.group_blocks[i]),
                                                  bk.nblocks, bk.threads, bk.lds, nullptr, b->step_host + 2 * k,
                                                  b->lanes[lane_of[i]]),
                              "batched sweep pair");
            }
            if (npblocks)
                LBM_HIP_CHECK(launch_sweep2_packed(b->storage,
                                                   reinterpret_cast<const SweepJob *>(dev + img.packed_jobs[pp]),
                                                   reinterpret_cast<const PackedSub *>(dev + img.packed_subs),
                                                   reinterpret_cast<const PackedBlock *>(dev + img.packed_blocks),
                                                   npblocks, pgrid, plds, nullptr, b->step_host + 2 * k,
                                                   b->lanes[lane_of[gpacked]]),
                              "packed sweep pair");
            if (!fjobs.empty())
                for (int p = 0; p < 2; p++)
                    LBM_HIP_CHECK(launch_fused_batch(b->storage, b->v,
                                                     reinterpret_cast<const BatchJob *>(dev + img.step_jobs[p]),
                                                     reinterpret_cast<const uint32_t *>(dev + img.step_blocks),
                                                     nfblocks, b->d_step, b->d_done, b->lanes[lane_of[gstep]]),
                                  "batched one-step pair");
        }
        for (size_t l = 0; l < nl; l++) {
            LBM_HIP_CHECK(hipEventRecord(b->lane_done[l], b->lanes[l]), "lbm_batch_run: join");
            LBM_HIP_CHECK(hipStreamWaitEvent(b->stream, b->lane_done[l], 0), "lbm_batch_run: join");
        }
        LBM_HIP_CHECK(launch_add_step(b->d_step + 1, 2 * pairs, b->stream), "lbm_batch_run: step counter");
        if (fjobs.empty()) LBM_HIP_CHECK(launch_add_step(b->d_step, 2 * pairs, b->stream), "lbm_batch_run: step counter");
        done_pairs = pairs;
    }
    if (b->use_graph && pairs - done_pairs >= 2) {
        LBM_HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "graph capture");
        hipError_t e0 = pair(0);
        hipError_t e1 = e0 == hipSuccess ? pair(1) : e0;
        hipError_t ec = hipStreamEndCapture(b->stream, &b->graph);
        if (e0 != hipSuccess || e1 != hipSuccess || ec != hipSuccess) {
            set_error("lbm_batch_run: graph capture of the sweep pairs failed");
            drop_graph(b);
            return LBM_ERR_HIP;
        }
        LBM_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
        // it refers to this run's tables only: no later run may replay it, and
        // it is destroyed after the final synchronize below, never in flight
        b->graph_blocks = 0;
        b->graph_v = -1;
        for (int64_t k = 0; k < pairs / 2; k++) LBM_HIP_CHECK(hipGraphLaunch(b->exec, b->stream), "graph launch");
        done_pairs = pairs / 2 * 2;
    }
    for (; done_pairs < pairs; done_pairs++) LBM_HIP_CHECK(pair((int)(done_pairs & 1)), "batched sweep pair");
    b->step_host += 2 * pairs;
    for (const Bucket &bk : buckets) {
        // launches that flipped each member's buffers: one per pair, or per
        // three-step launch plus the remainder's one
        const int64_t flips = bk.k == 3 && three_ran ? q3 + (r3 > 0 ? 1 : 0) : pairs;
        for (int j : bk.members) b->ctxs[j]->cur ^= (int)(flips & 1);
    }
    for (int j : pjobs) b->ctxs[j]->cur ^= (int)(pairs & 1);
    for (int j = 0; j < n; j++) b->ctxs[j]->steps += 2 * pairs;
    // back to the plain row order (after the join: on the batch stream)
    for (const auto &iv : ilv) {
        lbm_ctx *c = b->ctxs[iv.first];
        LBM_HIP_CHECK(launch_row_layout(c->storage, c->g, c->st, c->cur, iv.second, false, b->stream),
                      "lbm_batch_run: row order kernel");
        c->cur ^= 1;
    }
    if (nsteps & 1) {
        // the odd step, job by job
        for (int j = 0; j < n; j++) {
            lbm_ctx *c = b->ctxs[j];
            const int64_t slot = c->probe.cell >= 0 ? c->steps - c->probe_base - 1 : -1;
            LBM_HIP_CHECK(launch_fused(c->storage, c->g, c->st, c->cur, c->bc.view, b->omegas[j], c->probe, slot,
                                       nullptr, c->cfg, b->stream, nullptr),
                          "lbm_batch_run: odd step");
            c->cur ^= 1;
            c->steps++;
        }
        LBM_HIP_CHECK(launch_add_step(b->d_step, 1, b->stream), "lbm_batch_run: step counter");
        LBM_HIP_CHECK(launch_add_step(b->d_step + 1, 1, b->stream), "lbm_batch_run: step counter");
        b->step_host += 1;
    }
    for (int j = 0; j < n; j++) {
        lbm_ctx *c = b->ctxs[j];
        if (c->probe.cell >= 0 && c->steps > c->probe_base)
            LBM_HIP_CHECK(launch_probe_cell(c->storage, c->g, c->st, c->cur, c->bc.view, c->probe,
                                            c->steps - c->probe_base - 1, b->stream),
                          "probe kernel");
    }
    LBM_HIP_CHECK(hipStreamSynchronize(b->stream), "lbm_batch_run");
    drop_graph(b);
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_batch_create(lbm_batch **out, lbm_ctx *const *ctxs, const double *omegas, int n) {
    if (!out || n < 1 || !ctxs || !omegas) {
        set_error("lbm_batch_create: need out, n >= 1 contexts and their omegas");
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    for (int j = 0; j < n; j++) {
        if (!ctxs[j]) {
            set_error("lbm_batch_create: context " + std::to_string(j) + " is NULL");
            return LBM_ERR_ARG;
        }
        if (ctxs[j]->device != ctxs[0]->device || ctxs[j]->storage != ctxs[0]->storage) {
            set_error("lbm_batch_create: all contexts must share device and storage");
            return LBM_ERR_ARG;
        }
        if (!(omegas[j] > 0.0 && omegas[j] < 2.0)) {
            set_error("lbm_batch_create: omega must satisfy 0 < omega < 2");
            return LBM_ERR_ARG;
        }
    }
    auto *b = new (std::nothrow) lbm_batch();
    if (!b) {
        set_error("lbm_batch_create: out of host memory");
        return LBM_ERR_NOMEM;
    }
    b->ctxs.assign(ctxs, ctxs + n);
    b->omegas.assign(omegas, omegas + n);
    b->device = ctxs[0]->device;
    b->storage = ctxs[0]->storage;
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&b->d_jobs[0], sizeof(BatchJob) * n);
    if (e == hipSuccess) e = hipMalloc(&b->d_jobs[1], sizeof(BatchJob) * n);
    // d_step[0]: step counter of the one-step batched launches (advanced by
    // them); d_step[1]: the counter the sweep launches of a pair read, advanced
    // after the pair
    if (e == hipSuccess) e = hipMalloc(&b->d_step, 2 * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&b->d_done, sizeof(unsigned));
    if (e == hipSuccess) e = hipMemset(b->d_step, 0, 2 * sizeof(int64_t));

    if (e == hipSuccess) e = hipMemset(b->d_done, 0, sizeof(unsigned));
    if (e != hipSuccess) {
        set_error(std::string("lbm_batch_create: ") + hipGetErrorString(e));
        release_batch(b);
        delete b;
        return LBM_ERR_HIP;
    }
    *out = b;
    return LBM_OK;
}

int lbm_batch_destroy(lbm_batch *b) {
    if (!b) return LBM_OK;
    release_batch(b);
    delete b;
    return LBM_OK;
}

int lbm_batch_set_launch(lbm_batch *b, int v, int use_graph) {
    if (!b || (v != 1 && v != 2)) {
        set_error("lbm_batch_set_launch: nodes_per_thread must be 1 or 2");
        return LBM_ERR_ARG;
    }
    b->v = v;
    b->use_graph = use_graph != 0;
    return LBM_OK;
}

int lbm_batch_set_temporal(lbm_batch *b, int steps_per_launch, int rows_per_block) {
    if (!b || steps_per_launch < 1 || steps_per_launch > 3 || rows_per_block < 0) {
        set_error("lbm_batch_set_temporal: steps_per_launch must be 1, 2 or 3 and rows_per_block >= 0");
        return LBM_ERR_ARG;
    }
    b->temporal = steps_per_launch;
    b->sweep_rows = rows_per_block > 0 ? rows_per_block : 8;
    b->rows_auto = rows_per_block <= 0;
    drop_graph(b);
    return LBM_OK;
}

}  // extern "C"

namespace {

// The body of lbm_batch_run; it may return early on an error, with work of
// this run still in flight (lbm_batch_run drains it).
int batch_run_body(lbm_batch *b, int64_t nsteps) {
    const int n = (int)b->ctxs.size();
    for (int j = 0; j < n; j++) {
        lbm_ctx *c = b->ctxs[j];
        if (c->phase == kEmpty) {
            set_error("lbm_batch_run: context " + std::to_string(j) + " has no state loaded");
            return LBM_ERR_STATE;
        }
        LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "lbm_batch_run: context stream");
    }
    if (nsteps == 0) return LBM_OK;
    b->dirty = true;  // from here on, an error leaves the contexts' states part-way
    // A context holding a loaded step input starts with its first collision,
    // which is one step; if any does, every other context takes one step on
    // its own too, so that the batched launches advance all of them alike.
    bool any_pending = false;
    for (lbm_ctx *c : b->ctxs) any_pending |= c->phase == kPending;
    if (any_pending) {
        for (int j = 0; j < n; j++) {
            lbm_ctx *c = b->ctxs[j];
            if (c->phase == kPending) {
                LBM_HIP_CHECK(launch_collide_first(c->storage, c->g, c->stage_f, c->stage_rho, c->stage_u, c->st,
                                                   c->cur, c->bc.view, b->omegas[j], b->stream),
                              "lbm_batch_run: first collision");
                c->phase = kRunning;
            } else {
                const int64_t slot = c->probe.cell >= 0 ? c->steps - c->probe_base - 1 : -1;
                LBM_HIP_CHECK(launch_fused(c->storage, c->g, c->st, c->cur, c->bc.view, b->omegas[j], c->probe, slot,
                                           nullptr, c->cfg, b->stream, nullptr),
                              "lbm_batch_run: aligning step");
                c->cur ^= 1;
            }
            c->steps++;
        }
        nsteps--;
    }
    if (b->temporal >= 2) {
        // The jobs two steps per launch cannot take (the pressure fix) run
        // their one-step launches on a side stream of their own beside the
        // sweep groups, which pays at any share of them (8192 x 2048 channels
        // with 2048 x 2048 Poiseuille boxes, FP32, GLUPS with the sweep
        // groups / with one-step launches for every job: 20 % of the nodes in
        // boxes 123.0 / 79.4, 33 % 108.7 / 80.1, 50 % 99.2 / 80.2).  Only a
        // batch of such jobs alone keeps the one-step launches.
        bool any_sweep = false;
        for (lbm_ctx *c : b->ctxs)
            any_sweep |= c->st.scratch && sweep_batch_supported(c->storage, c->g, c->bc.view);
        if (any_sweep) return batch_run_sweep(b, nsteps);
    }

    // job tables for the two step parities of this run
    StepTables st;
    for (int p = 0; p < 2; p++) st.jobs[p].resize(n);
    for (int j = 0; j < n; j++) {
        lbm_ctx *c = b->ctxs[j];
        const uint32_t nb = (uint32_t)batch_job_blocks(c->g, b->v);
        const uint32_t block0 = (uint32_t)st.block_job.size();
        st.block_job.insert(st.block_job.end(), nb, (uint32_t)j);
        for (int p = 0; p < 2; p++) {
            BatchJob &d = st.jobs[p][j];
            const int src = c->cur ^ p;
            d.src = c->st.plane[src];
            d.dst = c->st.plane[src ^ 1];
            d.side_prev = c->st.side[src];
            d.side_next = c->st.side[src ^ 1];
            d.g = c->g;
            d.bc = c->bc.view;
            d.omega = b->omegas[j];
            d.block0 = block0;
            if (c->probe.cell >= 0) {
                d.probe_cell = c->probe.cell;
                d.probe_ring = c->probe.ring;
                d.probe_cap = c->probe.capacity;
                d.probe_offset = c->steps - c->probe_base - 1 - b->step_host;
            }
        }
    }
    const unsigned nblocks = (unsigned)st.block_job.size();
    pack_batch_tables({}, PackedTables{}, st, b->image);
    if (nblocks > b->block_cap) {
        drop_graph(b);
        (void)hipFree(b->d_block_job);
        b->d_block_job = nullptr;
        LBM_HIP_CHECK(hipMalloc(&b->d_block_job, sizeof(uint32_t) * nblocks), "lbm_batch_run: block table");
        b->block_cap = nblocks;
    }
    // through pinned staging into the fixed job and block buffers the
    // captured step pair reads
    LBM_HIP_CHECK(stage_tables(b), "lbm_batch_run: pinned staging of the tables");
    const char *h = static_cast<const char *>(b->h_stage);
    LBM_HIP_CHECK(hipMemcpyAsync(b->d_block_job, h + b->image.step_blocks, sizeof(uint32_t) * nblocks,
                                 hipMemcpyHostToDevice, b->stream), "lbm_batch_run: block table");
    for (int p = 0; p < 2; p++)
        LBM_HIP_CHECK(hipMemcpyAsync(b->d_jobs[p], h + b->image.step_jobs[p], sizeof(BatchJob) * n,
                                     hipMemcpyHostToDevice, b->stream), "lbm_batch_run: job table");
    b->nblocks = nblocks;

    // step pairs replayed from the graph, then the odd one
    int64_t left = nsteps;
    const int64_t pairs = left / 2;
    if (b->use_graph && pairs > 0) {
        if (!b->exec || b->graph_blocks != nblocks || b->graph_v != b->v) {
            drop_graph(b);
            LBM_HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "graph capture");
            hipError_t e0 = launch_fused_batch(b->storage, b->v, b->d_jobs[0], b->d_block_job, nblocks, b->d_step,
                                               b->d_done, b->stream);
            hipError_t e1 = launch_fused_batch(b->storage, b->v, b->d_jobs[1], b->d_block_job, nblocks, b->d_step,
                                               b->d_done, b->stream);
            hipError_t ec = hipStreamEndCapture(b->stream, &b->graph);
            if (e0 != hipSuccess || e1 != hipSuccess || ec != hipSuccess) {
                set_error("lbm_batch_run: graph capture failed");
                drop_graph(b);
                return LBM_ERR_HIP;
            }
            LBM_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
            b->graph_blocks = nblocks;
            b->graph_v = b->v;
        }
        for (int64_t k = 0; k < pairs; k++) LBM_HIP_CHECK(hipGraphLaunch(b->exec, b->stream), "graph launch");
        left -= 2 * pairs;
    }
    for (; left > 0; left--) {
        const int p = (int)((nsteps - left) & 1);
        LBM_HIP_CHECK(launch_fused_batch(b->storage, b->v, b->d_jobs[p], b->d_block_job, nblocks, b->d_step,
                                         b->d_done, b->stream),
                      "batched step kernel");
    }
    // keep the counter the two-step launches read in step with this one
    LBM_HIP_CHECK(launch_add_step(b->d_step + 1, nsteps, b->stream), "lbm_batch_run: step counter");
    b->step_host += nsteps;
    for (int j = 0; j < n; j++) {
        lbm_ctx *c = b->ctxs[j];
        c->cur ^= (int)(nsteps & 1);
        c->steps += nsteps;
        if (c->probe.cell >= 0 && c->steps > c->probe_base)
            LBM_HIP_CHECK(launch_probe_cell(c->storage, c->g, c->st, c->cur, c->bc.view, c->probe,
                                            c->steps - c->probe_base - 1, b->stream),
                          "probe kernel");
    }
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_batch_set_packing(lbm_batch *b, int packed) {
    if (!b || packed < 0 || packed > 2) {
        set_error("lbm_batch_set_packing: packed must be 0, 1 or 2");
        return LBM_ERR_ARG;
    }
    b->packed = packed;
    return LBM_OK;
}

int lbm_copy_probe(const void *src, void *dst, int64_t bytes, void *stream) {
    if (bytes < 0 || bytes % 16 || (bytes > 0 && (!src || !dst))) {
        set_error("lbm_copy_probe: need device buffers and bytes >= 0, a multiple of 16");
        return LBM_ERR_ARG;
    }
    LBM_HIP_CHECK(launch_copy16(src, dst, bytes / 16, static_cast<hipStream_t>(stream)), "lbm_copy_probe");
    return LBM_OK;
}

int lbm_row_sweep_probe_ordered(const void *src, void *dst, int nx, int pitch, int rows, int mode, void *stream) {
    if (nx <= 0 || rows <= 0 || pitch <= 0 || pitch % 2 != 0 || pitch > 4096 || !src || !dst || mode < 0 ||
        mode > 3) {
        set_error("lbm_row_sweep_probe: need device buffers, nx > 0, rows > 0, an even pitch up to 4096 and "
                  "mode 0 to 3");
        return LBM_ERR_ARG;
    }
    const bool interleaved = (mode & LBM_PROBE_INTERLEAVED) != 0;
    if (interleaved && (rows > (int)kRowLayoutMaxRows || (uint32_t)nx >= kRowLayoutMaxLx)) {
        set_error("lbm_row_sweep_probe: the interleaved order needs rows <= 128 and nx < 2^24");
        return LBM_ERR_ARG;
    }
    LBM_HIP_CHECK(launch_row_sweep_probe(src, dst, (uint32_t)nx, (uint32_t)pitch, (uint32_t)rows, interleaved,
                                         (mode & LBM_PROBE_PULL) != 0, static_cast<hipStream_t>(stream)),
                  "lbm_row_sweep_probe");
    return LBM_OK;
}

int lbm_row_sweep_probe(const void *src, void *dst, int nx, int pitch, int rows, void *stream) {
    return lbm_row_sweep_probe_ordered(src, dst, nx, pitch, rows, 0, stream);
}

int lbm_batch_run(lbm_batch *b, int64_t nsteps) {
    if (!b) {
        set_error("lbm_batch_run: batch is NULL");
        return LBM_ERR_ARG;
    }
    if (nsteps < 0) {
        set_error("lbm_batch_run: nsteps must be >= 0");
        return LBM_ERR_ARG;
    }
    LBM_HIP_CHECK(hipSetDevice(b->device), "lbm_batch_run");
    // Nothing of an earlier run is in flight when its tables are rebuilt, and
    // nothing of this run is when it returns, on success or on an error.
    LBM_HIP_CHECK(quiesce(b), "lbm_batch_run: earlier work");
    b->dirty = false;
    int rc = batch_run_body(b, nsteps);
    const hipError_t e = quiesce(b);
    // a one-step run keeps its captured pair for the next run (it reads the
    // batch's fixed job buffers); after an error nothing is kept
    if (rc != LBM_OK || e != hipSuccess) drop_graph(b);
    if (rc == LBM_OK && e != hipSuccess) {
        set_error(std::string("lbm_batch_run: ") + hipGetErrorString(e));
        rc = LBM_ERR_HIP;
    }
    if (rc != LBM_OK && b->dirty) {
        // part of the run was enqueued: the contexts' states are unknown
        for (lbm_ctx *c : b->ctxs) c->phase = kEmpty;
        set_error(std::string(lbm_last_error()) + " (the batch's contexts lost their state; load it again)");
    }
    return rc;
}

}  // extern "C"

// ---- the reduction that allocates nothing --------------------------------
// Block partial sums of lbm_moments_sum_noalloc: three doubles per 256-node
// block of k_moments_sum_aos.  Each GPU holds its own copy from the first
// load of this code object on it until the process ends.  In host code the
// name stands for no device address: kernels get the per-device address that
// hipGetSymbolAddress returns (noalloc_partials()).
static __device__ __align__(256) double scratch[(4u << 20) / sizeof(double)];

namespace {
constexpr int kNoallocMaxDevices = 64;
constexpr int64_t kNoallocBlocks = (int64_t)(sizeof(scratch) / (3 * sizeof(double)));
static_assert(kNoallocBlocks * 256 == LBM_NOALLOC_MAX_CELLS, "LBM_NOALLOC_MAX_CELLS is the scratch's block count * 256");
std::once_flag g_noalloc_once[kNoallocMaxDevices];
void *g_noalloc_ptr[kNoallocMaxDevices];
hipError_t g_noalloc_err[kNoallocMaxDevices];

// The device address of `scratch` on device `dev` (current), looked up once
// per device; null with the lookup's error in *err if it failed.
double *noalloc_partials(int dev, hipError_t *err) {
    std::call_once(g_noalloc_once[dev], [dev] {
        void *p = nullptr;
        g_noalloc_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_noalloc_ptr[dev] = g_noalloc_err[dev] == hipSuccess ? p : nullptr;
    });
    *err = g_noalloc_err[dev];
    return static_cast<double *>(g_noalloc_ptr[dev]);
}
}  // namespace

extern "C" {

int lbm_moments_sum_noalloc(const double *f, int64_t ncells, double *out, void *stream) {
    if (!out || ncells < 0 || (ncells > 0 && !f)) {
        set_error_fmt("lbm_moments_sum_noalloc: need out != NULL, ncells >= 0 and f != NULL when ncells > 0");
        return LBM_ERR_ARG;
    }
    if (ncells > LBM_NOALLOC_MAX_CELLS) {
        set_error_fmt("lbm_moments_sum_noalloc: ncells = %lld exceeds LBM_NOALLOC_MAX_CELLS = %d (the fixed "
                      "scratch holds %lld block sums)", (long long)ncells, LBM_NOALLOC_MAX_CELLS,
                      (long long)kNoallocBlocks);
        return LBM_ERR_ARG;
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error_fmt("lbm_moments_sum_noalloc: hipGetDevice: %s", hipGetErrorString(e));
        return LBM_ERR_HIP;
    }
    if (dev < 0 || dev >= kNoallocMaxDevices) {
        set_error_fmt("lbm_moments_sum_noalloc: device %d is beyond the %d this library keeps scratch for", dev,
                      kNoallocMaxDevices);
        return LBM_ERR_ARG;
    }
    double *partials = noalloc_partials(dev, &e);
    if (!partials) {
        set_error_fmt("lbm_moments_sum_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        return LBM_ERR_HIP;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    int nb = 0;
    // every one of the nb partials is written before k_sum_partials reads it,
    // so nothing left in the array by an earlier call is read
    if (ncells > 0) e = launch_moments_sum_aos(f, ncells, partials, s, &nb);
    if (e == hipSuccess) e = launch_sum_partials(partials, nb, out, s);
    if (e != hipSuccess) {
        set_error_fmt("lbm_moments_sum_noalloc: kernel launch: %s", hipGetErrorString(e));
        return LBM_ERR_HIP;
    }
    return LBM_OK;
}

// ---- stateless drop-ins ---------------------------------------------------

static int check_dtype(int dtype, const char *fn) {
    if (dtype != LBM_DT_F64 && dtype != LBM_DT_F32) {
        set_error(std::string(fn) + ": dtype must be LBM_DT_F64 or LBM_DT_F32");
        return LBM_ERR_ARG;
    }
    return LBM_OK;
}

int lbm_stream(const void *f_in, void *f_out, int nx, int ny, int dtype, void *stream) {
    if (int rc = check_dtype(dtype, "lbm_stream")) return rc;
    if (nx < 1 || ny < 1 || !f_in || !f_out || f_in == f_out) {
        set_error("lbm_stream: need nx, ny >= 1 and distinct non-NULL buffers");
        return LBM_ERR_ARG;
    }
    LBM_HIP_CHECK(launch_stream_aos(f_in, f_out, nx, ny, dtype, static_cast<hipStream_t>(stream)), "lbm_stream");
    return LBM_OK;
}

int lbm_equilibrium(const void *rho, const void *u, void *feq, int64_t n, int dtype, void *stream) {
    if (int rc = check_dtype(dtype, "lbm_equilibrium")) return rc;
    if (n < 0 || (n > 0 && (!rho || !u || !feq))) {
        set_error("lbm_equilibrium: bad arguments");
        return LBM_ERR_ARG;
    }
    if (n == 0) return LBM_OK;
    LBM_HIP_CHECK(launch_equilibrium_aos(rho, u, feq, n, dtype, static_cast<hipStream_t>(stream)), "lbm_equilibrium");
    return LBM_OK;
}

int lbm_density(const void *f, void *rho, int64_t n, int dtype, void *stream) {
    if (int rc = check_dtype(dtype, "lbm_density")) return rc;
    if (n < 0 || (n > 0 && (!f || !rho))) {
        set_error("lbm_density: bad arguments");
        return LBM_ERR_ARG;
    }
    if (n == 0) return LBM_OK;
    LBM_HIP_CHECK(launch_density_aos(f, rho, n, dtype, static_cast<hipStream_t>(stream)), "lbm_density");
    return LBM_OK;
}

int lbm_velocity(const void *rho, const void *f, void *u, int64_t n, int dtype, void *stream) {
    if (int rc = check_dtype(dtype, "lbm_velocity")) return rc;
    if (n < 0 || (n > 0 && (!f || !rho || !u))) {
        set_error("lbm_velocity: bad arguments");
        return LBM_ERR_ARG;
    }
    if (n == 0) return LBM_OK;
    LBM_HIP_CHECK(launch_velocity_aos(rho, f, u, n, dtype, static_cast<hipStream_t>(stream)), "lbm_velocity");
    return LBM_OK;
}

int lbm_collide(const void *f, const void *rho, const void *u, double omega, void *f_out, int64_t n, int dtype,
                void *stream) {
    if (int rc = check_dtype(dtype, "lbm_collide")) return rc;
    if (!(omega > 0.0 && omega < 2.0)) {
        set_error("lbm_collide: omega must satisfy 0 < omega < 2");
        return LBM_ERR_ARG;
    }
    if (n < 0 || (n > 0 && (!f || !rho || !u || !f_out))) {
        set_error("lbm_collide: bad arguments");
        return LBM_ERR_ARG;
    }
    if (n == 0) return LBM_OK;
    LBM_HIP_CHECK(launch_collide_aos(f, rho, u, omega, f_out, n, dtype, static_cast<hipStream_t>(stream)),
                  "lbm_collide");
    return LBM_OK;
}

static int prog_on_device(const lbm_bc_prog *prog, BcView &view, const char *fn) {
    int dev = 0;
    LBM_HIP_CHECK(hipGetDevice(&dev), fn);
    auto *p = const_cast<lbm_bc_prog *>(prog);
    if (p->dev_mem && p->dev_id != dev) {
        (void)hipSetDevice(p->dev_id);
        (void)hipFree(p->dev_mem);
        (void)hipSetDevice(dev);
        p->dev_mem = nullptr;
    }
    if (!p->dev_mem) {
        BcUpload up;
        LBM_HIP_CHECK(upload_bc(*p, up), fn);
        p->dev_mem = up.mem;
        p->dev_id = dev;
        p->dev_codes = up.view.codes;
        p->dev_chunk = up.view.chunk;
        p->dev_wave = up.view.wave;
        p->dev_classes = up.view.classes;
        p->dev_special = up.view.special;
    }
    view.codes = p->dev_codes;
    view.chunk = p->dev_chunk;
    view.wave = p->dev_wave;
    view.classes = p->dev_classes;
    view.nclasses = (uint32_t)p->classes.size();
    view.special = p->dev_special;
    view.nspecial = (int64_t)p->special.size();
    view.rho_in = p->rho_in;
    view.rho_out = p->rho_out;
    view.d1 = p->d1;
    view.d2 = p->d2;
    view.flags = p->flags;
    return LBM_OK;
}

int lbm_bc_apply(const lbm_bc_prog *prog, const double *f_pre, double *f_post, const double *f_prev, void *stream) {
    if (!prog || !f_pre || !f_post) {
        set_error("lbm_bc_apply: prog, f_pre and f_post are required");
        return LBM_ERR_ARG;
    }
    if ((prog->flags & kProgOutlet) && !f_prev) {
        set_error("lbm_bc_apply: the program has an outlet, f_prev is required");
        return LBM_ERR_ARG;
    }
    BcView view;
    if (int rc = prog_on_device(prog, view, "lbm_bc_apply")) return rc;
    const Geom g = make_geom((uint32_t)prog->nx, (uint32_t)prog->ny);
    LBM_HIP_CHECK(launch_bc_apply(g, view, f_pre, f_post, f_prev, static_cast<hipStream_t>(stream)), "lbm_bc_apply");
    return LBM_OK;
}

int lbm_bc_pressure(const lbm_bc_prog *prog, double *f_pre, const double *rho, const double *u, void *stream) {
    if (!prog || !f_pre || !rho || !u) {
        set_error("lbm_bc_pressure: prog, f_pre, rho and u are required");
        return LBM_ERR_ARG;
    }
    if (!(prog->flags & kProgPressure)) {
        set_error("lbm_bc_pressure: the program has no pressure-periodic descriptor");
        return LBM_ERR_ARG;
    }
    BcView view;
    if (int rc = prog_on_device(prog, view, "lbm_bc_pressure")) return rc;
    const Geom g = make_geom((uint32_t)prog->nx, (uint32_t)prog->ny);
    LBM_HIP_CHECK(launch_bc_pressure(g, view, f_pre, rho, u, static_cast<hipStream_t>(stream)), "lbm_bc_pressure");
    return LBM_OK;
}

}  // extern "C"
