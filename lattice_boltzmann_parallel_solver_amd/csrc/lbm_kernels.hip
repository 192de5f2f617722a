// D2Q9 lattice-Boltzmann kernels for gfx950 (MI355X).
//
// State layout in HBM: row-interleaved planes, population i of node (x, y) at
// ((x*9 + i)*pitch + y): every row stores its nine y-contiguous planes back to
// back, so the three source rows of a wavefront are three contiguous blocks
// (measured faster than nine whole-lattice planes, which stream from nine
// regions at once).  pitch is ly rounded up to whole 128-byte lines, so every
// plane row starts on a line (a row length off the line costs the one-step
// kernel about 5 %: each wave then touches one more line per plane); the pad
// elements are never read or written.  Node n = x*ly + y (x slowest, as the
// reference's f[x, y, i]); two buffers (A/B) alternate every step.  The hot kernel is a
// pull-scheme fused step: each thread gathers the nine populations streaming
// into its node, applies the node's boundary class, computes the moments,
// the equilibrium and the BGK relaxation in registers, and writes the nine
// post-collision populations.  It is bandwidth bound (18 plane accesses per
// node per step, no reuse inside a step), so the design goal is one
// coalesced, aligned store and one coalesced load per plane per node.
//
// Storage policies
//   double : f_i in FP64, arithmetic in the reference's exact operation order
//            (lattice_boltzmann_method.py:93-215) with FMA contraction off.
//   float  : deviations d_i = f_i - w_i in FP32, FP32 arithmetic.
//   __half : deviations in FP16, FP32 arithmetic.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <algorithm>
#include <type_traits>

#include "lbm_launch.h"
#include "lbm_rowlayout.h"

namespace lbm {

Geom make_geom(uint32_t lx, uint32_t ly, uint32_t align) {
    Geom g;
    g.lx = lx;
    g.ly = ly;
    g.N = lx * ly;
    g.pitch = (ly + align - 1) / align * align;
    g.wpr = (ly + 63) / 64;
    g.ly_magic = ly == 1 ? 0 : (~0ull) / ly + 1;
    g.wpr_magic = g.wpr == 1 ? 0 : (~0ull) / g.wpr + 1;
    return g;
}

namespace {

constexpr int kBlock = 256;
constexpr double kWd[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                           1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};

// ---------------------------------------------------------------------------
// storage I/O
// ---------------------------------------------------------------------------
template <class S>
struct Io;

// Pointers into global memory typed as such: loads and stores through them are
// global accesses.  Through a generic pointer (one read from a job record
// rather than passed as a kernel argument) the compiler emits flat accesses,
// which count against the LDS wait counter too, so that every wait for LDS
// data would also wait for the global loads in flight.
#define LBM_GLOBAL __attribute__((address_space(1)))

// The accessors take generic or global pointers (P).
template <>
struct Io<double> {
    using C = double;
    static constexpr bool kDev = false;
    template <class P>
    __device__ static C ld(P p) { return *p; }
    template <class P>
    __device__ static void st(P p, C v, bool nt) {
        if (nt) __builtin_nontemporal_store(v, p);
        else *p = v;
    }
};

template <>
struct Io<float> {
    using C = float;
    static constexpr bool kDev = true;
    template <class P>
    __device__ static C ld(P p) { return *p; }
    template <class P>
    __device__ static void st(P p, C v, bool nt) {
        if (nt) __builtin_nontemporal_store(v, p);
        else *p = v;
    }
};

// FP32 -> FP16 with round-to-nearest-even, always from the FP32-rounded value.
// The empty asm hides the value's origin from the instruction combiner, which
// would otherwise fold a preceding fmaf into v_fma_mix (one rounding straight
// to FP16 instead of FP32 then FP16) in some kernels and not in others.
__device__ __forceinline__ __half f32_to_f16(float v) {
    asm("" : "+v"(v));
    return __float2half_rn(v);
}

template <>
struct Io<__half> {
    using C = float;
    static constexpr bool kDev = true;
    __device__ static C ld(const __half *p) { return __half2float(*p); }
    __device__ static C ld(const LBM_GLOBAL __half *p) {
        return __half2float(__ushort_as_half(*(const LBM_GLOBAL unsigned short *)p));
    }
    __device__ static void st(__half *p, C v, bool nt) {
        unsigned short bits = __half_as_ushort(f32_to_f16(v));
        if (nt) __builtin_nontemporal_store(bits, reinterpret_cast<unsigned short *>(p));
        else *reinterpret_cast<unsigned short *>(p) = bits;
    }
    __device__ static void st(LBM_GLOBAL __half *p, C v, bool nt) {
        unsigned short bits = __half_as_ushort(f32_to_f16(v));
        using U = LBM_GLOBAL unsigned short;
        if (nt) __builtin_nontemporal_store(bits, (U *)p);
        else *(U *)p = bits;
    }
};

// Representation of a float64 population value f_i in the compute type of S
// (deviation f - w_i for the reduced-precision policies).
template <class S>
__device__ __forceinline__ typename Io<S>::C rep(double f, int i) {
    if constexpr (Io<S>::kDev) return (typename Io<S>::C)(f - kWd[i]);
    else return f;
}

// ---------------------------------------------------------------------------
// node index helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t row_of(const Geom &g, uint32_t n) {
    return g.ly == 1 ? n : (uint32_t)__umul64hi(g.ly_magic, (uint64_t)n);
}

// Spread consecutive block ids over the 8 XCDs so that each XCD works on one
// contiguous range of nodes; the y-shifted loads that straddle two blocks
// then hit the same L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t G) {
    const uint32_t q = G >> 3, r = G & 7, xcd = b & 7, local = b >> 3;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// Offset of population i of node (x, y) in a state buffer.
__device__ __forceinline__ size_t pop(const Geom &g, uint32_t x, int i, uint32_t y) {
    return ((size_t)x * 9 + i) * g.pitch + y;
}

// Source row of direction i for destination row x (periodic in x).
__device__ __forceinline__ uint32_t src_row(const Geom &g, uint32_t x, int cx) {
    if (cx == 1) return x == 0 ? g.lx - 1 : x - 1;
    if (cx == -1) return x == g.lx - 1 ? 0 : x + 1;
    return x;
}

__device__ __forceinline__ uint32_t src_col(const Geom &g, uint32_t y, int cy) {
    if (cy == 1) return y == 0 ? g.ly - 1 : y - 1;
    if (cy == -1) return y == g.ly - 1 ? 0 : y + 1;
    return y;
}

// Periodic pull: population i of node (x, y) comes from node (x - cx_i, y - cy_i).
template <class S, class P>
__device__ __forceinline__ void pull(P src, const Geom &g, uint32_t x, uint32_t y, typename Io<S>::C out[9]) {
#pragma unroll
    for (int i = 0; i < 9; i++)
        out[i] = Io<S>::ld(src + pop(g, src_row(g, x, kCX[i]), i, src_col(g, y, kCY[i])));
}

// Element `off` past `base` through a 32-bit byte offset (lets the compiler
// use the uniform-base + 32-bit-offset addressing form).  Valid while
// off * sizeof(S) < 2^32: for offsets from the start of a state buffer, while
// 9 * lx * pitch * sizeof(S) < 2^32.
template <class S>
__device__ __forceinline__ const S *at32(const S *base, uint32_t off) {
    return reinterpret_cast<const S *>(reinterpret_cast<const char *>(base) + (uint32_t)(off * (uint32_t)sizeof(S)));
}
template <class S>
__device__ __forceinline__ S *at32(S *base, uint32_t off) {
    return reinterpret_cast<S *>(reinterpret_cast<char *>(base) + (uint32_t)(off * (uint32_t)sizeof(S)));
}

template <class S>
__device__ __forceinline__ const LBM_GLOBAL S *at32(const LBM_GLOBAL S *base, uint32_t off) {
    return (const LBM_GLOBAL S *)((const LBM_GLOBAL char *)base + (uint32_t)(off * (uint32_t)sizeof(S)));
}
template <class S>
__device__ __forceinline__ LBM_GLOBAL S *at32(LBM_GLOBAL S *base, uint32_t off) {
    return (LBM_GLOBAL S *)((LBM_GLOBAL char *)base + (uint32_t)(off * (uint32_t)sizeof(S)));
}

// Plane i of row r of a state buffer, from a 64-bit base (wave-uniform where
// r is); lanes add their column with at32().  For buffers of 4 GiB and more.
template <class S>
__device__ __forceinline__ const S *plane_row(const S *base, uint32_t pitch, uint32_t r, int i) {
    return base + ((size_t)r * 9 + i) * pitch;
}
template <class S>
__device__ __forceinline__ S *plane_row(S *base, uint32_t pitch, uint32_t r, int i) {
    return base + ((size_t)r * 9 + i) * pitch;
}

// Population i of node (r, c): a 32-bit offset from the buffer start, or with
// kBig a 64-bit row base plus a 32-bit column offset.  The 32-bit form keeps
// fewer registers live in the row sweep (its offsets are not hoisted).
template <bool kBig, class S>
__device__ __forceinline__ S *pop32(S *base, uint32_t pitch, uint32_t r, int i, uint32_t c) {
    if constexpr (kBig) return at32(plane_row(base, pitch, r, i), c);
    else return at32(base, (r * 9 + i) * pitch + c);
}

// pull() with 32-bit offsets: the three source rows are given.
template <class S, bool kBig = false, class P>
__device__ __forceinline__ void pull32(P src, uint32_t pitch, uint32_t xm, uint32_t x, uint32_t xp, uint32_t y,
                                       uint32_t ym, uint32_t yp, typename Io<S>::C out[9]) {
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const uint32_t r = kCX[i] == 1 ? xm : (kCX[i] == -1 ? xp : x);
        const uint32_t c = kCY[i] == 1 ? ym : (kCY[i] == -1 ? yp : y);
        out[i] = Io<S>::ld(pop32<kBig>(src, pitch, r, i, c));
    }
}

// Apply the node's boundary class to the pulled populations; returns the
// class flags (0 for plain fluid).
template <class S, class P, class Q>
__device__ __forceinline__ uint32_t apply_class(const BcView &bc, P src, Q side_prev, const Geom &g, uint32_t n,
                                                uint32_t x, uint32_t y, typename Io<S>::C v[9]) {
    using C = typename Io<S>::C;
    const uint32_t code = bc.codes[n];
    if (code == 0) return 0;
    const ClassEntry &ce = bc.classes[code];
    const uint32_t w = ce.src;
#pragma unroll
    for (int d = 0; d < 9; d++) {
        const uint32_t kind = (w >> (2 * d)) & 3u;
        if (kind == kSrcReflect) {
            v[d] = Io<S>::ld(src + pop(g, x, kOpp[d], y)) - (C)ce.k[d];
        } else if (kind == kSrcConst) {
            if constexpr (Io<S>::kDev) v[d] = (C)(ce.k[d] - kWd[d]);
            else v[d] = ce.k[d];
        } else if (kind == kSrcOutlet) {
            v[d] = side_prev[(size_t)y * 3 + outlet_slot(d)];
        }
    }
    return ce.flags;
}

// ---------------------------------------------------------------------------
// FP64 arithmetic in the reference's operation order (no contraction)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void moments_f64(const double f[9], double &rho, double &jx, double &jy) {
#pragma clang fp contract(off)
    rho = (((f[0] + f[1]) + (f[2] + f[3])) + ((f[4] + f[5]) + (f[6] + f[7]))) + f[8];
    jx = ((f[1] + f[5]) + f[8]) - ((f[3] + f[6]) + f[7]);
    jy = ((f[2] + f[5]) + f[6]) - ((f[4] + f[7]) + f[8]);
}

__device__ __forceinline__ void velocity_f64(double rho, double jx, double jy, double &ux, double &uy) {
    if (rho != 0.0) {
        ux = jx / rho;
        uy = jy / rho;
    } else {
        ux = 0.0;
        uy = 0.0;
    }
}

__device__ __forceinline__ void feq_f64(double rho, double ux, double uy, double e[9]) {
#pragma clang fp contract(off)
    const double mag = __builtin_sqrt(ux * ux + uy * uy);
    const double usq = mag * mag;
    // c_i . u with exact products and one rounding, as numpy's u @ c.T
    const double cu[9] = {0.0, ux, uy, -ux, -uy, ux + uy, uy - ux, -(ux + uy), ux - uy};
#pragma unroll
    for (int i = 0; i < 9; i++)
        e[i] = (kWd[i] * rho) * (((1.0 + 3.0 * cu[i]) + 4.5 * (cu[i] * cu[i])) - 1.5 * usq);
}

// ---------------------------------------------------------------------------
// Collision of one node: moments of the post-boundary populations v, the
// equilibrium, BGK relaxation.  Outputs the relaxed populations, the
// equilibrium (needed by the pressure fix) and the moments.
// ---------------------------------------------------------------------------
template <class S>
struct NodeOut {
    using C = typename Io<S>::C;
    C s[9];
    C e[9];
    double rho, jx, jy;  // for probes and reductions
    C ux, uy;
};

template <class S>
__device__ __forceinline__ void collide_node(const typename Io<S>::C v[9], double omega, NodeOut<S> &o) {
    if constexpr (!Io<S>::kDev) {
#pragma clang fp contract(off)
        double rho, jx, jy, ux, uy;
        moments_f64(v, rho, jx, jy);
        velocity_f64(rho, jx, jy, ux, uy);
        feq_f64(rho, ux, uy, o.e);
#pragma unroll
        for (int i = 0; i < 9; i++) o.s[i] = v[i] + (o.e[i] - v[i]) * omega;
        o.rho = rho;
        o.jx = jx;
        o.jy = jy;
        o.ux = ux;
        o.uy = uy;
    } else {
        // Contraction off and the fused multiply-adds written out, so that every
        // kernel rounds the reduced-precision collision the same way.
        //   p_i = drho + rho (3 cu_i + 4.5 cu_i^2 - 1.5 |u|^2), f_eq,i - w_i = w_i p_i,
        //   s_i = v_i + omega (w_i p_i - v_i) = (1 - omega) v_i + (omega w_i) p_i.
        // Opposite directions share p's even part (c_opp = -c_i): per pair
        // even = B + 4.5 rho cu^2 and odd = 3 rho cu give p_i = even + odd and
        // p_opp = even - odd.
        // 1/rho: the hardware reciprocal refined by one Newton step (within an
        // ulp, like the correctly rounded division it replaces at a third of the
        // instructions; the raw reciprocal alone doubled the FP32 error against
        // FP64).
#pragma clang fp contract(off)
        const float om = (float)omega;
        const float om1 = 1.0f - om;
        const float drho = (((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]))) + v[8];
        const float rho = 1.0f + drho;
        const float a57 = v[5] - v[7], b86 = v[8] - v[6];
        const float jx = ((v[1] - v[3]) + a57) + b86;
        const float jy = ((v[2] - v[4]) + a57) - b86;
        float inv = 0.0f;
        if (rho != 0.0f) {
            const float r0 = __builtin_amdgcn_rcpf(rho);
            inv = __builtin_fmaf(__builtin_fmaf(-rho, r0, 1.0f), r0, r0);
        }
        const float ux = jx * inv, uy = jy * inv;
        const float usq = __builtin_fmaf(ux, ux, uy * uy);
        const float B = __builtin_fmaf(-1.5f * rho, usq, drho);
        const float r3 = 3.0f * rho, r45 = 4.5f * rho;
        float pp[9];
        pp[0] = B;
        const float cu[4] = {ux, uy, ux + uy, uy - ux};  // directions 1, 2, 5, 6
        constexpr int kPos[4] = {1, 2, 5, 6}, kNeg[4] = {3, 4, 7, 8};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float even = __builtin_fmaf(r45 * cu[k], cu[k], B);
            const float odd = r3 * cu[k];
            pp[kPos[k]] = even + odd;
            pp[kNeg[k]] = even - odd;
        }
#pragma unroll
        for (int i = 0; i < 9; i++) {
            const float wi = (float)kWd[i];
            o.s[i] = __builtin_fmaf(om * wi, pp[i], om1 * v[i]);
            o.e[i] = wi * pp[i];  // used by the pressure fix only
        }
        o.rho = 1.0 + (double)drho;
        o.jx = jx;
        o.jy = jy;
        o.ux = ux;
        o.uy = uy;
    }
}

// Equilibrium at (rho_b, u) of the node, in the policy's representation.
template <class S>
__device__ __forceinline__ void feq_rep(double rho_b, const NodeOut<S> &o, typename Io<S>::C e[9]) {
    if constexpr (!Io<S>::kDev) {
        feq_f64(rho_b, o.ux, o.uy, e);
    } else {
#pragma clang fp contract(off)
        const float ux = o.ux, uy = o.uy;
        const float r = (float)rho_b, dr = (float)(rho_b - 1.0);
        const float base = -1.5f * __builtin_fmaf(ux, ux, uy * uy);
        const float cu[9] = {0.0f, ux, uy, -ux, -uy, ux + uy, uy - ux, -(ux + uy), ux - uy};
#pragma unroll
        for (int i = 0; i < 9; i++)
            e[i] = (float)kWd[i] * __builtin_fmaf(r, __builtin_fmaf(cu[i], __builtin_fmaf(4.5f, cu[i], 3.0f), base), dr);
    }
}

// Store the relaxed node plus the boundary side effects of its class:
// outlet buffer (x = lx-2) and the pre-streaming pressure fix
// (boundary_conditions.py:337-344): the node at x = lx-2 writes
// E(rho_in, u) + (s - feq) into column 0 for directions D1, the node at x = 1
// writes E(rho_out, u) + (s - feq) into column lx-1 for directions D2.
template <class S, class P, class Q>
__device__ __forceinline__ void store_node(P dst, Q side_next, const Geom &g, const BcView &bc, uint32_t x,
                                           uint32_t y, uint32_t flags, const typename Io<S>::C v[9],
                                           const NodeOut<S> &o, bool nt) {
    using C = typename Io<S>::C;
    const uint32_t skip = flags >> kSkipShift;
#pragma unroll
    for (int i = 0; i < 9; i++)
        if (!((skip >> i) & 1u)) Io<S>::st(dst + pop(g, x, i, y), o.s[i], nt);
    if (flags & kFlagOutletSide) {
        side_next[(size_t)y * 3 + 0] = v[3];
        side_next[(size_t)y * 3 + 1] = v[6];
        side_next[(size_t)y * 3 + 2] = v[7];
    }
    if (flags & (kFlagPressureIn | kFlagPressureOut)) {
#pragma clang fp contract(off)
        C e[9];
        if (flags & kFlagPressureIn) {
            feq_rep<S>(bc.rho_in, o, e);
#pragma unroll
            for (int i = 0; i < 9; i++)
                if ((bc.d1 >> i) & 1u) Io<S>::st(dst + pop(g, 0, i, y), e[i] + (o.s[i] - o.e[i]), false);
        }
        if (flags & kFlagPressureOut) {
            feq_rep<S>(bc.rho_out, o, e);
#pragma unroll
            for (int i = 0; i < 9; i++)
                if ((bc.d2 >> i) & 1u)
                    Io<S>::st(dst + pop(g, g.lx - 1, i, y), e[i] + (o.s[i] - o.e[i]), false);
        }
    }
}

// Sum three doubles over the block in a fixed order; thread 0 writes out[0..2].
__device__ __forceinline__ void block_sum3(double a, double b, double c, double *out) {
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    __shared__ double sh[kBlock / 64][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sh[wave][0] = a;
        sh[wave][1] = b;
        sh[wave][2] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0, s2 = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            s0 += sh[w][0];
            s1 += sh[w][1];
            s2 += sh[w][2];
        }
        out[0] = s0;
        out[1] = s1;
        out[2] = s2;
    }
}

// ---------------------------------------------------------------------------
// the fused step
// ---------------------------------------------------------------------------
template <class S>
struct FusedParams {
    const S *src;
    S *dst;
    const typename Io<S>::C *side_prev;
    typename Io<S>::C *side_next;
    Geom g;
    BcView bc;
    double omega;
    int64_t probe_cell;
    double *probe_out;  // 3 doubles, or null
    double *partials;   // 3 doubles per block, or null
    int xcd;
    int nt;
};

// Waves of 64 threads tile each row from y = 0 (the last wave of a row is
// partly idle), so that with the line-aligned pitch every plane access of a
// wave starts on a 64-element boundary, and the lanes past the row end write
// its padding: no 128-byte line is left partly written (rows of 1800 nodes
// ran 5-7 % slower before).  Wave w of the launch is row w / wpr, columns
// 64 (w mod wpr) ... + 63.
template <class S, int V, bool kProbe, bool kReduce>
__global__ __launch_bounds__(kBlock) void k_fused(FusedParams<S> p) {
    using C = typename Io<S>::C;
    const uint32_t blk = p.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const Geom &g = p.g;
    const uint32_t lane = threadIdx.x & 63;
    C v[V][9];
    uint32_t xs[V], ys[V], ws[V];
    bool live[V];
#pragma unroll
    for (int k = 0; k < V; k++) {
        const uint32_t w = __builtin_amdgcn_readfirstlane((blk * V + k) * (kBlock / 64) + (threadIdx.x >> 6));
        const uint32_t x = g.wpr == 1 ? w : (uint32_t)__umul64hi(g.wpr_magic, (uint64_t)w);
        ws[k] = w;
        xs[k] = x;
        ys[k] = (w - x * g.wpr) * 64 + lane;
        live[k] = x < g.lx && ys[k] < g.ly;
        if (live[k]) pull<S>(p.src, g, xs[k], ys[k], v[k]);
    }
    double acc0 = 0, acc1 = 0, acc2 = 0;
#pragma unroll
    for (int k = 0; k < V; k++) {
        if (live[k]) {
            const uint32_t n = xs[k] * g.ly + ys[k];
            uint32_t flags = 0;
            if (p.bc.wave && p.bc.wave[ws[k]])  // wave w is row x, columns 64 (w mod wpr) on: its flag is entry w
                flags = apply_class<S>(p.bc, p.src, p.side_prev, g, n, xs[k], ys[k], v[k]);
            NodeOut<S> o;
            collide_node<S>(v[k], p.omega, o);
            store_node<S>(p.dst, p.side_next, g, p.bc, xs[k], ys[k], flags, v[k], o, p.nt != 0);
            if constexpr (kProbe) {
                if ((int64_t)n == p.probe_cell) {
                    p.probe_out[0] = o.rho;
                    p.probe_out[1] = (double)o.ux;
                    p.probe_out[2] = (double)o.uy;
                }
            }
            if constexpr (kReduce) {
                acc0 += o.rho;
                acc1 += o.jx;
                acc2 += o.jy;
            }
        } else if (xs[k] < g.lx && ys[k] < g.pitch) {
            // lanes over the row padding write zeros there (never read), right
            // after the wave's own stores, so that the last line of each plane
            // row is written whole: a partly written line that leaves the
            // cache costs the memory a read-modify-write
#pragma unroll
            for (int i = 0; i < 9; i++) Io<S>::st(p.dst + pop(g, xs[k], i, ys[k]), (C)0, p.nt != 0);
        }
    }
    if constexpr (kReduce) block_sum3(acc0, acc1, acc2, p.partials + 3 * (size_t)blockIdx.x);
}

// ---------------------------------------------------------------------------
// the fused step with 16-byte accesses
//
// Each thread owns one aligned 16-byte group of every plane: V = 16/sizeof(S)
// consecutive nodes along y (2 FP64, 4 FP32, 8 FP16), so every load and store
// is an aligned dwordx4.  A plane shifted in y (cy = +-1) needs one element
// outside the group; it comes from the neighbouring lane's group by a lane
// shuffle, or, for the first/last lane of a wave and at the row ends, from a
// predicated single-element load.  Requires ly % V == 0.
// ---------------------------------------------------------------------------
using raw4 = unsigned int __attribute__((ext_vector_type(4)));

template <class S>
struct Pack;

template <>
struct Pack<double> {
    static constexpr int V = 2;
    __device__ static double get(const raw4 &r, int k) {
        return __hiloint2double((int)r[2 * k + 1], (int)r[2 * k]);
    }
    __device__ static void set(raw4 &r, int k, double v) {
        r[2 * k] = (unsigned)__double2loint(v);
        r[2 * k + 1] = (unsigned)__double2hiint(v);
    }
    __device__ static double from_prev(const raw4 &r) {  // element V-1 of lane-1
        const unsigned lo = __shfl_up(r[2], 1, 64), hi = __shfl_up(r[3], 1, 64);
        return __hiloint2double((int)hi, (int)lo);
    }
    __device__ static double from_next(const raw4 &r) {  // element 0 of lane+1
        const unsigned lo = __shfl_down(r[0], 1, 64), hi = __shfl_down(r[1], 1, 64);
        return __hiloint2double((int)hi, (int)lo);
    }
};

template <>
struct Pack<float> {
    static constexpr int V = 4;
    __device__ static float get(const raw4 &r, int k) { return __uint_as_float(r[k]); }
    __device__ static void set(raw4 &r, int k, float v) { r[k] = __float_as_uint(v); }
    __device__ static float from_prev(const raw4 &r) { return __uint_as_float(__shfl_up(r[3], 1, 64)); }
    __device__ static float from_next(const raw4 &r) { return __uint_as_float(__shfl_down(r[0], 1, 64)); }
};

template <>
struct Pack<__half> {
    static constexpr int V = 8;
    __device__ static float half_bits(unsigned b) { return __half2float(__ushort_as_half((unsigned short)b)); }
    __device__ static float get(const raw4 &r, int k) { return half_bits((r[k >> 1] >> (16 * (k & 1))) & 0xffffu); }
    __device__ static void set(raw4 &r, int k, float v) {
        const unsigned b = __half_as_ushort(f32_to_f16(v));
        const int sh = 16 * (k & 1);
        r[k >> 1] = (r[k >> 1] & ~(0xffffu << sh)) | (b << sh);
    }
    __device__ static float from_prev(const raw4 &r) { return half_bits(__shfl_up(r[3], 1, 64) >> 16); }
    __device__ static float from_next(const raw4 &r) { return half_bits(__shfl_down(r[0], 1, 64) & 0xffffu); }
};

template <class S, bool kProbe, bool kReduce>
__global__ __launch_bounds__(kBlock) void k_fused16(FusedParams<S> p) {
    using C = typename Io<S>::C;
    using P = Pack<S>;
    constexpr int V = P::V;
    const Geom &g = p.g;
    const uint32_t blk = p.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t t = blk * kBlock + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    const bool live = (uint64_t)t * V < g.N;
    const uint32_t n0 = live ? t * V : 0;  // dead lanes load node 0 and store nothing
    const uint32_t x = row_of(g, n0), y0 = n0 - x * g.ly;

    // 1. one aligned 16-byte load per plane
    raw4 r[9];
    const S *rows[9];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        rows[i] = p.src + pop(g, src_row(g, x, kCX[i]), i, 0);
        r[i] = *reinterpret_cast<const raw4 *>(rows[i] + y0);
    }
    // 2. the element each y-shifted plane needs from outside the group
    C edge[9];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        if (kCY[i] == 1) edge[i] = P::from_prev(r[i]);
        else if (kCY[i] == -1) edge[i] = P::from_next(r[i]);
    }
    const bool fix_prev = lane == 0 || y0 == 0;
    const bool fix_next = lane == 63 || y0 + V == g.ly;
    if (fix_prev || fix_next) {
        const uint32_t yp = y0 == 0 ? g.ly - 1 : y0 - 1;
        const uint32_t yn = y0 + V == g.ly ? 0 : y0 + V;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            if (kCY[i] == 1 && fix_prev) edge[i] = Io<S>::ld(rows[i] + yp);
            if (kCY[i] == -1 && fix_next) edge[i] = Io<S>::ld(rows[i] + yn);
        }
    }
    // 3. boundary classes are looked up only in waves that hold a special node
    bool special = false;
    if (p.bc.chunk) {
        const uint32_t wave_first = (t - lane) * V;
        if (wave_first < g.N) {
            const uint32_t c0 = __builtin_amdgcn_readfirstlane(wave_first >> 6);
            uint32_t any = 0;
#pragma unroll
            for (int j = 0; j < V; j++) any |= p.bc.chunk[c0 + j];
            special = any != 0;
        }
    }
    // 4. collide node by node
    raw4 o[9];
    uint32_t skip_any = 0, skip_node[V];
    double acc0 = 0, acc1 = 0, acc2 = 0;
#pragma unroll
    for (int k = 0; k < V; k++) {
        C v[9];
#pragma unroll
        for (int i = 0; i < 9; i++) {
            if (kCY[i] == 0) v[i] = P::get(r[i], k);
            else if (kCY[i] == 1) v[i] = k == 0 ? edge[i] : P::get(r[i], k - 1);
            else v[i] = k == V - 1 ? edge[i] : P::get(r[i], k + 1);
        }
        const uint32_t n = n0 + k, y = y0 + k;
        uint32_t flags = 0;
        if (special && live) flags = apply_class<S>(p.bc, p.src, p.side_prev, g, n, x, y, v);
        NodeOut<S> out;
        collide_node<S>(v, p.omega, out);
#pragma unroll
        for (int i = 0; i < 9; i++) P::set(o[i], k, out.s[i]);
        skip_node[k] = flags >> kSkipShift;
        skip_any |= skip_node[k];
        if (flags & ~(0x1ffu << kSkipShift)) {
            // outlet buffer and pressure fix (store_node without the plane stores)
            store_node<S>(p.dst, p.side_next, g, p.bc, x, y, flags | (0x1ffu << kSkipShift), v, out, false);
        }
        if constexpr (kProbe) {
            if (live && (int64_t)n == p.probe_cell) {
                p.probe_out[0] = out.rho;
                p.probe_out[1] = (double)out.ux;
                p.probe_out[2] = (double)out.uy;
            }
        }
        if constexpr (kReduce) {
            if (live) {
                acc0 += out.rho;
                acc1 += out.jx;
                acc2 += out.jy;
            }
        }
    }
    // 5. one aligned 16-byte store per plane (per node where a pressure fix owns a direction)
    if (live) {
#pragma unroll
        for (int i = 0; i < 9; i++) {
            S *dst = p.dst + pop(g, x, i, y0);
            if (!((skip_any >> i) & 1u)) {
                if (p.nt) __builtin_nontemporal_store(o[i], reinterpret_cast<raw4 *>(dst));
                else *reinterpret_cast<raw4 *>(dst) = o[i];
            } else {
#pragma unroll
                for (int k = 0; k < V; k++)
                    if (!((skip_node[k] >> i) & 1u)) Io<S>::st(dst + k, P::get(o[i], k), false);
            }
        }
    }
    if constexpr (kReduce) block_sum3(acc0, acc1, acc2, p.partials + 3 * (size_t)blockIdx.x);
}

// ---------------------------------------------------------------------------
// the fused step over a batch of independent lattices
//
// Many small lattices (tens to hundreds of thousands of nodes) would each
// leave most of the 256 CUs idle and pay a launch per step; here one launch
// advances all of them.  Block b belongs to job block_job[b]; the job's
// descriptor gives its buffers, geometry, boundary tables and omega.  A
// device step counter, advanced by a one-thread kernel after each step, lets
// the probes of a graph-replayed launch sequence find their ring slot.
// ---------------------------------------------------------------------------
template <class S, int V>
__global__ __launch_bounds__(kBlock) void k_fused_batch(const BatchJob *__restrict__ jobs,
                                                        const uint32_t *__restrict__ block_job,
                                                        const int64_t *__restrict__ step) {
    using C = typename Io<S>::C;
    const BatchJob &d = jobs[block_job[blockIdx.x]];
    const Geom &g = d.g;
    // global pointers (read from the job record the compiler cannot tell)
    const LBM_GLOBAL S *src = (const LBM_GLOBAL S *)d.src;
    LBM_GLOBAL S *dst = (LBM_GLOBAL S *)d.dst;
    const LBM_GLOBAL C *side_prev = (const LBM_GLOBAL C *)d.side_prev;
    LBM_GLOBAL C *side_next = (LBM_GLOBAL C *)d.side_next;
    const int64_t s_now = *step;
    const uint32_t lb = blockIdx.x - d.block0;
#pragma unroll
    for (int k = 0; k < V; k++) {
        // waves tile each row from y = 0, as in k_fused
        const uint32_t w = __builtin_amdgcn_readfirstlane((lb * V + k) * (kBlock / 64) + (threadIdx.x >> 6));
        const uint32_t x = g.wpr == 1 ? w : (uint32_t)__umul64hi(g.wpr_magic, (uint64_t)w);
        const uint32_t y = (w - x * g.wpr) * 64 + (threadIdx.x & 63);
        if (x < g.lx && y < g.ly) {
            const uint32_t n = x * g.ly + y;
            C v[9];
            pull<S>(src, g, x, y, v);
            uint32_t flags = 0;
            if (d.bc.wave && d.bc.wave[w]) flags = apply_class<S>(d.bc, src, side_prev, g, n, x, y, v);
            NodeOut<S> o;
            collide_node<S>(v, d.omega, o);
            store_node<S>(dst, side_next, g, d.bc, x, y, flags, v, o, false);
            if ((int64_t)n == d.probe_cell) {
                const int64_t slot = s_now + d.probe_offset;
                if (slot >= 0) {
                    LBM_GLOBAL double *out = (LBM_GLOBAL double *)d.probe_ring + 3 * (slot % d.probe_cap);
                    out[0] = o.rho;
                    out[1] = (double)o.ux;
                    out[2] = (double)o.uy;
                }
            }
        } else if (x < g.lx && y < g.pitch) {
            // complete the last line of each plane row (the pad is never read)
#pragma unroll
            for (int i = 0; i < 9; i++) Io<S>::st(dst + pop(g, x, i, y), (C)0, false);
        }
    }
}

__global__ void k_next_step(int64_t *step) { *step += 1; }

// ---------------------------------------------------------------------------
// two steps per launch: temporal blocking over rows
//
// A block owns rows [xa, xb) of the output and all ly nodes of each row, or,
// for rows wider than a block, one column tile of them plus a halo column on
// each side (thread t handles columns t, t + T, ...).  It sweeps x upwards; iteration
// it = 0 .. R+1 (R = xb - xa)
//   1. runs step 1 for row x = xa - 1 + it (periodic): pull from the input
//      buffer, boundary class, collision; the relaxed populations s1 stay on
//      chip, rounded to the storage type exactly as a store would round them;
//   2. publishes the six y-moving populations of s1 in LDS, and after one
//      barrier reads them back shifted by one node in y (periodic): together
//      with the three populations that do not move in y this is row x of s1
//      streamed in y;
//   3. from it = 2 on, runs step 2 for row x - 1: directions with cx = +1
//      come from the y-streamed row x - 2, cx = 0 from row x - 1, cx = -1
//      from row x (the first two are carried in registers), boundary class
//      (reflections read the node's own s1, the outlet reads the step-1
//      outlet values kept in LDS), collision, store to the output buffer.
// Every value passes through the same arithmetic as in two single steps, so
// the result equals two k_fused launches bit for bit.  The input is read once
// (plus the two rows on each side that neighbouring blocks also read) and the
// output written once per two steps.  The step-1 loads of the next row are
// issued before the current row is processed (in FP32, half of them after the
// row's barrier).  Not used with the pressure fix,
// whose producers write rows owned by other blocks.
// ---------------------------------------------------------------------------
constexpr int kSweepThreads = kSweepThreadsMax;
// y-moving directions: the first three arrive from y - 1, the last three from y + 1.
constexpr int kYM[6] = {2, 5, 6, 4, 7, 8};
constexpr int kYSlot[9] = {-1, -1, 0, -1, 3, 1, 2, 4, 5};

template <class S>
struct SweepParams {
    const S *src;
    S *dst;
    const typename Io<S>::C *side_prev;
    typename Io<S>::C *side_next;
    Geom g;
    BcView bc;
    double omega;
    int64_t probe_cell;
    double *probe1;  // sample of step 1 (3 doubles) or null
    double *probe2;  // sample of step 2
    uint32_t rows;   // output rows per block
    uint32_t ntiles; // column tiles per row range (1: whole rows)
    uint32_t tw;     // output columns per tile (ly with whole rows)
    uint32_t lw;     // LDS columns of a tile: tw + 2 with tiles (halo column each side), else ly
    int nt;
    int xcd;         // XCD-aware block order (neighbouring row ranges share an L2)
    S *scratch;      // 9 V blockDim.x values, written and never read
    RowLayout il;    // row order of src and dst (rows = 0: plain)
};

template <class S>
__device__ __forceinline__ typename Io<S>::C round_to_storage(typename Io<S>::C v) {
    if constexpr (sizeof(S) == 2) return __half2float(f32_to_f16(v));
    else return v;
}

template <class S>
__device__ __forceinline__ void lds_put(S *p, typename Io<S>::C v) {
    if constexpr (sizeof(S) == 2) *p = f32_to_f16(v);
    else *p = v;
}

template <class S>
__device__ __forceinline__ typename Io<S>::C lds_get(const S *p) {
    if constexpr (sizeof(S) == 2) return __half2float(*p);
    else return *p;
}

constexpr uint32_t kSweepMaxClasses = 128;
constexpr uint32_t kSweepMaxRows = 128;

__host__ __device__ constexpr size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }

// Dynamic LDS of the sweep kernel for `cols` columns (ly, or a column tile
// with its two halo columns): the y-exchange rows (two buffers of six planes
// of cols + 1 values, the last a spare column for lanes past the last column),
// the step-1 outlet values, the incoming outlet values (the input side
// buffer), the class table and one "special" byte per (row, wave column) of
// the block's rows.
struct SweepLds {
    size_t side, side_in, cls, flag, total;
};

template <class S>
__host__ __device__ inline SweepLds sweep_lds_layout(uint32_t cols, uint32_t nclasses, uint32_t rows) {
    using C = typename Io<S>::C;
    SweepLds L;
    L.side = align16(12 * ((size_t)cols + 1) * sizeof(S));  // two buffers of six planes of cols + 1
    L.side_in = align16(L.side + 3 * (size_t)cols * sizeof(C));
    L.cls = align16(L.side_in + 3 * (size_t)cols * sizeof(C));
    L.flag = align16(L.cls + (size_t)nclasses * sizeof(ClassEntry));
    L.total = L.flag + (size_t)(rows + 2) * ((cols + 63) / 64);
    return L;
}

// pull32() for a node with a boundary class: directions the class reflects
// load the node's own opposite population (the reflection's source) in place
// of the upstream one, so that no load is needed later.
template <class S, bool kBig, class P>
__device__ __forceinline__ void pull32_bc(P src, uint32_t pitch, uint32_t xm, uint32_t x, uint32_t xp, uint32_t y,
                                          uint32_t ym, uint32_t yp, uint32_t w, typename Io<S>::C out[9]) {
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const uint32_t r = kCX[i] == 1 ? xm : (kCX[i] == -1 ? xp : x);
        const uint32_t c = kCY[i] == 1 ? ym : (kCY[i] == -1 ? yp : y);
        const bool refl = ((w >> (2 * i)) & 3u) == kSrcReflect;
        if constexpr (kBig) {
            out[i] = Io<S>::ld(refl ? pop32<true>(src, pitch, x, kOpp[i], y) : pop32<true>(src, pitch, r, i, c));
        } else {
            const uint32_t off = refl ? (x * 9 + kOpp[i]) * pitch + y : (r * 9 + i) * pitch + c;
            out[i] = Io<S>::ld(at32(src, off));
        }
    }
}

template <bool B>
struct Flag {
    static constexpr bool value = B;
};

// The threads that sweep one row range: the whole block, or in a packed block
// one run of its waves (tid counts from the run's first thread; T threads;
// its LDS area from smem), and the row iterations every wave of the block
// runs (packed blocks: the longest range's rows + 2).
struct SweepLanes {
    uint32_t tid, T;
    unsigned char *smem;
    uint32_t r_block;
};

// The rows [xa, xa + p.rows) of one lattice (clipped at lx), two steps.
// With kTiled the block owns only output columns [cy0, cy0 + W) of them, cy0 =
// tile * tw: step 1 also runs on one halo column each side (periodic in y),
// whose values step 2 needs; the halo columns' step-2 results are not kept.
// LDS and lane columns are then local: local column l is global column
// cy0 - 1 + l (mod ly).  kBig: state buffers of 4 GiB and more (64-bit row
// bases).  kPacked: the lanes share their block with other row ranges, so
// every barrier is block-wide and runs the same number of times on every
// wave: the set-up barrier always, then ln.r_block + 2 row iterations (those
// past this range's last row do nothing but the barrier).
template <class S, int V, bool kProbe, bool kTiled, bool kBig, bool kPacked = false>
__device__ __forceinline__ void sweep_block(const SweepParams<S> &p, uint32_t xa, uint32_t tile, SweepLanes ln) {
    using C = typename Io<S>::C;
    unsigned char *smem = ln.smem;
    const Geom &g = p.g;
    const uint32_t ly = g.ly, lx = g.lx, pitch = g.pitch;
    const uint32_t T = ln.T, tid = ln.tid;
    // the buffers as global pointers (the batched kernels read them from job
    // records, where the compiler cannot tell their address space)
    using GS = LBM_GLOBAL S;
    const GS *gsrc = (const GS *)p.src;
    GS *gdst = (GS *)p.dst;
    GS *gscr = (GS *)p.scratch;
    LBM_GLOBAL C *gside = (LBM_GLOBAL C *)p.side_next;
    const LBM_GLOBAL uint16_t *gcodes = (const LBM_GLOBAL uint16_t *)p.bc.codes;
    const uint32_t R = min(p.rows, lx - xa);
    const uint32_t cy0 = kTiled ? tile * p.tw : 0;  // first output column
    const uint32_t W = kTiled ? min(p.tw, ly - cy0) : ly;  // output columns
    const uint32_t lw = kTiled ? W + 2 : ly;              // computed columns
    const uint32_t cw = kTiled ? p.lw : ly;               // LDS columns
    const uint32_t cs = cw + 1;                           // LDS plane stride (a spare column)
    const uint32_t nw = (lw + 63) / 64;
    const bool has_bc = p.bc.chunk != nullptr;
    const SweepLds L = sweep_lds_layout<S>(cw, has_bc ? p.bc.nclasses : 0, p.rows);
    // global column of local column l < lw (tiles only; cy0 - 1 + l lies in [-1, ly])
    auto gcol = [&](uint32_t l) {
        const uint32_t t = cy0 + l + ly - 1;
        return t >= ly ? (t - ly >= ly ? t - 2 * ly : t - ly) : t;
    };
    S *ex = reinterpret_cast<S *>(smem);
    C *side = reinterpret_cast<C *>(smem + L.side);
    C *side_in = reinterpret_cast<C *>(smem + L.side_in);
    ClassEntry *cls = reinterpret_cast<ClassEntry *>(smem + L.cls);
    uint8_t *wflag = smem + L.flag;
    // row of iteration it (step-1 row xa - 1 + it, periodic)
    auto row_at = [&](uint32_t it) {
        const uint32_t r = xa + it + lx - 1;
        return r >= lx ? (r - lx >= lx ? r - 2 * lx : r - lx) : r;
    };
    if (has_bc) {
        const uint32_t words = p.bc.nclasses * (uint32_t)(sizeof(ClassEntry) / 4);
        for (uint32_t i = tid; i < words; i += T)
            reinterpret_cast<uint32_t *>(cls)[i] = ((const LBM_GLOBAL uint32_t *)p.bc.classes)[i];
        // step 1 reads the outlet values of the previous launch from LDS: only
        // blocks whose step-1 rows include lx-1 (the last block, and the first
        // through the periodic row xa-1) need them
        if ((p.bc.flags & kProgOutlet) && (xa == 0 || xa + R + 1 >= lx)) {
            for (uint32_t i = tid; i < 3 * lw; i += T) {
                if constexpr (kTiled) {
                    const uint32_t l = i / 3;
                    side_in[i] = ((const LBM_GLOBAL C *)p.side_prev)[gcol(l) * 3 + (i - 3 * l)];
                } else {
                    side_in[i] = ((const LBM_GLOBAL C *)p.side_prev)[i];
                }
            }
        }
        for (uint32_t i = tid; i < (R + 2) * nw; i += T) {
            const uint32_t rr = i / nw, j = i - rr * nw;
            const LBM_GLOBAL uint8_t *wr = (const LBM_GLOBAL uint8_t *)p.bc.wave + (size_t)row_at(rr) * g.wpr;
            if constexpr (kTiled) {
                // local wave j spans global columns a .. b (unwrapped, a >= -1,
                // b <= ly): the global waves of both ends, and of both row
                // ends where the span wraps
                const int a = (int)cy0 - 1 + 64 * (int)j;
                const int b = (int)cy0 - 1 + (int)min(64 * j + 63, lw - 1);
                const uint32_t aw = (uint32_t)(a < 0 ? a + (int)ly : (a >= (int)ly ? a - (int)ly : a));
                const uint32_t bw = (uint32_t)(b < 0 ? b + (int)ly : (b >= (int)ly ? b - (int)ly : b));
                uint8_t f = wr[aw / 64] | wr[bw / 64];
                if (a < 0 || b >= (int)ly) f |= wr[0] | wr[(ly - 1) / 64];
                wflag[i] = f;
            } else {
                wflag[i] = wr[j];
            }
        }
        if constexpr (!kPacked) __syncthreads();
    }
    if constexpr (kPacked) __syncthreads();
    // Every lane is active: lanes past the last column repeat column ly-1 with
    // the same inputs, boundary class and results (their stores write the same
    // values to the same places).  With no lane-dependent branch around memory
    // operations the compiler can count the stores issued after the prefetch
    // and does not wait for them at the top of the loop.
    // ys/ym/yp: global column and its y neighbours (HBM pulls, codes, outlet
    // buffer); with tiles ls/lm/lp are the local ones (LDS)
    uint32_t ys[V], ym[V], yp[V], jw[V], ls[V], lm[V], lp[V];
#pragma unroll
    for (int k = 0; k < V; k++) {
        const uint32_t l = min(tid + k * T, lw - 1);
        ys[k] = kTiled ? gcol(l) : l;
        ym[k] = ys[k] == 0 ? ly - 1 : ys[k] - 1;
        yp[k] = ys[k] + 1 >= ly ? 0 : ys[k] + 1;
        jw[k] = min((tid >> 6) + k * (T >> 6), nw - 1);
        ls[k] = l;
        lm[k] = l == 0 ? 0 : l - 1;
        lp[k] = min(l + 1, lw - 1);
    }
    auto lcol = [&](int k) { return kTiled ? ls[k] : ys[k]; };  // LDS column of the lane
    auto lmin = [&](int k) { return kTiled ? lm[k] : ym[k]; };  // LDS column of its y - 1
    auto lplu = [&](int k) { return kTiled ? lp[k] : yp[k]; };  // LDS column of its y + 1
    // the lane's column is an output column of this block (tiles: not a halo
    // or clamped lane)
    auto owned = [&](int k) {
        const uint32_t lu = tid + k * T;
        return !kTiled || (lu >= 1 && lu <= W);
    };
    // the wave holding column k has boundary nodes in row iteration rr (wave-uniform)
    auto special = [&](uint32_t rr, int k) -> bool {
        if (!has_bc) return false;
        return __builtin_amdgcn_readfirstlane(wflag[rr * nw + jw[k]]) != 0;
    };
    auto load_code = [&](uint32_t rr, int k) -> uint32_t {
        if (!special(rr, k)) return 0;
        return gcodes[row_at(rr) * ly + ys[k]];
    };
    // the pulls of row iteration rr for the lane's columns k0 <= k < k1
    // (rows are addressed through their slots in the buffers' row order)
    auto pull_row = [&](uint32_t rr, const uint32_t code[V], C out[V][9], auto k0_tag, auto k1_tag) {
        constexpr int k0 = decltype(k0_tag)::value, k1 = decltype(k1_tag)::value;
        const uint32_t xr = row_at(rr), xm = xr == 0 ? lx - 1 : xr - 1, xp = xr + 1 == lx ? 0 : xr + 1;
        const uint32_t sr = row_slot(p.il, xr), sm = row_slot(p.il, xm), sp = row_slot(p.il, xp);
#pragma unroll
        for (int k = k0; k < k1; k++) {
            if (special(rr, k)) {
                const uint32_t w = code[k] ? cls[code[k]].src : 0u;
                pull32_bc<S, kBig>(gsrc, pitch, sm, sr, sp, ys[k], ym[k], yp[k], w, out[k]);
            } else {
                pull32<S, kBig>(gsrc, pitch, sm, sr, sp, ys[k], ym[k], yp[k], out[k]);
            }
        }
    };
    // codes of the rows of step 2 (cm), step 1 (c0), the prefetched row (c1)
    // and the row after it (c2, loaded one iteration ahead)
    uint32_t cm[V], c0[V], c1[V], c2[V];
#pragma unroll
    for (int k = 0; k < V; k++) {
        cm[k] = 0;
        c0[k] = load_code(0, k);
        c1[k] = load_code(1, k);
        c2[k] = 0;
    }
    C A[V][3], B[V][3], Cc[V][3], D[V];  // y-streamed s1: A = row x-2 (1,5,8), B = row x-1 (0,2,4), Cc = row x-1 (1,5,8), D = row x-1 (3)
#pragma unroll
    for (int k = 0; k < V; k++) {
        D[k] = 0;
#pragma unroll
        for (int j = 0; j < 3; j++) A[k][j] = B[k][j] = Cc[k][j] = 0;
    }
    using K0 = std::integral_constant<int, 0>;
    using KV = std::integral_constant<int, V>;
    // FP32 with two columns per thread pulls the next row's first column at
    // the top of an iteration and its second after the barrier: the reads
    // spread over the iteration instead of arriving at HBM in one burst per
    // row (+3.5 % at 8192x2048; no gain in FP16, a loss in FP64 split by
    // direction).
    constexpr int kH = sizeof(S) == 4 && V == 2 ? 1 : V;
    using KH = std::integral_constant<int, kH>;
    C pf[V][9];
    pull_row(0, c0, pf, K0{}, KV{});

    // One iteration: step 1 on row x = row_at(it) (prefetching the next row
    // first if kPf, its second column after the barrier where kH < V), then,
    // if kS2, step 2 on row x - 1.
    auto body = [&](uint32_t it, auto pf_flag, auto s2_flag) {
        constexpr bool kPf = decltype(pf_flag)::value;
        constexpr bool kS2 = decltype(s2_flag)::value;
        const uint32_t x = row_at(it);
        C v[V][9];
#pragma unroll
        for (int k = 0; k < V; k++)
#pragma unroll
            for (int i = 0; i < 9; i++) v[k][i] = pf[k][i];
        if constexpr (kPf) {
            pull_row(it + 1, c1, pf, K0{}, KH{});
#pragma unroll
            for (int k = 0; k < V; k++) c2[k] = it + 2 <= R + 1 ? load_code(it + 2, k) : 0u;
        }
        S *exb = ex + (it & 1) * 6 * (size_t)cs;
        const S *exp_ = ex + ((it + 1) & 1) * 6 * (size_t)cs;
        C s013[V][3];
        // step 1 on row x
#pragma unroll
        for (int k = 0; k < V; k++) {
            const uint32_t y = ys[k];
            uint32_t flags = 0;
            if (c0[k]) {
                const ClassEntry &ce = cls[c0[k]];
                const uint32_t w = ce.src;
                flags = ce.flags;
#pragma unroll
                for (int d = 0; d < 9; d++) {
                    const uint32_t kind = (w >> (2 * d)) & 3u;
                    if (kind == kSrcReflect) {
                        v[k][d] = v[k][d] - (C)ce.k[d];  // v holds the node's own opposite population
                    } else if (kind == kSrcConst) {
                        if constexpr (Io<S>::kDev) v[k][d] = (C)(ce.k[d] - kWd[d]);
                        else v[k][d] = ce.k[d];
                    } else if (kind == kSrcOutlet) {
                        v[k][d] = side_in[(size_t)lcol(k) * 3 + outlet_slot(d)];
                    }
                }
            }
            NodeOut<S> o;
            collide_node<S>(v[k], p.omega, o);
            if (flags & kFlagOutletSide) {
                side[(size_t)lcol(k) * 3 + 0] = v[k][3];
                side[(size_t)lcol(k) * 3 + 1] = v[k][6];
                side[(size_t)lcol(k) * 3 + 2] = v[k][7];
            }
            if constexpr (kProbe) {
                if (p.probe1 && it >= 1 && it <= R && owned(k) && (int64_t)(x * ly + y) == p.probe_cell) {
                    LBM_GLOBAL double *pr = (LBM_GLOBAL double *)p.probe1;
                    pr[0] = o.rho;
                    pr[1] = (double)o.ux;
                    pr[2] = (double)o.uy;
                }
            }
            // Only the column's own lane publishes it.  A lane past the last
            // column repeats it, possibly in another wave, which may already be
            // one iteration ahead: its write of this buffer would overwrite the
            // values the owner still reads as its reflection sources.  Such
            // lanes write the spare column at the end of each plane (or, in a
            // narrower last tile, columns past its lw that nobody reads) instead.
            {
                const uint32_t col = min(tid + k * T, cw);
#pragma unroll
                for (int j = 0; j < 6; j++) lds_put<S>(exb + j * cs + col, o.s[kYM[j]]);
            }
            s013[k][0] = round_to_storage<S>(o.s[0]);
            s013[k][1] = round_to_storage<S>(o.s[1]);
            s013[k][2] = round_to_storage<S>(o.s[3]);
        }
        __syncthreads();
        if constexpr (kPf && kH < V) pull_row(it + 1, c1, pf, KH{}, KV{});
#pragma unroll
        for (int k = 0; k < V; k++) {
            const uint32_t y = ys[k];
            C Y[9];
            Y[0] = s013[k][0];
            Y[1] = s013[k][1];
            Y[3] = s013[k][2];
            Y[2] = lds_get<S>(exb + 0 * cs + lmin(k));
            Y[5] = lds_get<S>(exb + 1 * cs + lmin(k));
            Y[6] = lds_get<S>(exb + 2 * cs + lmin(k));
            Y[4] = lds_get<S>(exb + 3 * cs + lplu(k));
            Y[7] = lds_get<S>(exb + 4 * cs + lplu(k));
            Y[8] = lds_get<S>(exb + 5 * cs + lplu(k));
            if constexpr (kS2) {
                // step 2 on row r = x - 1 (stored at slot sr2)
                const uint32_t r = row_at(it - 1);
                const uint32_t sr2 = row_slot(p.il, r);
                C v2[9];
                v2[0] = B[k][0];
                v2[2] = B[k][1];
                v2[4] = B[k][2];
                v2[1] = A[k][0];
                v2[5] = A[k][1];
                v2[8] = A[k][2];
                v2[3] = Y[3];
                v2[6] = Y[6];
                v2[7] = Y[7];
                uint32_t flags2 = 0;
                if (cm[k]) {
                    // reflections read the node's own s1: directions 0, 1, 3 from
                    // registers, the y-moving ones from the LDS row this thread
                    // published one iteration earlier
                    const ClassEntry &ce = cls[cm[k]];
                    const uint32_t w = ce.src;
                    flags2 = ce.flags;
#pragma unroll
                    for (int d = 0; d < 9; d++) {
                        const uint32_t kind = (w >> (2 * d)) & 3u;
                        if (kind == kSrcReflect) {
                            const int s = kOpp[d];
                            const C own = s == 0 ? B[k][0] : s == 1 ? Cc[k][0] : s == 3 ? D[k]
                                                                                   : lds_get<S>(exp_ + kYSlot[s] * cs + lcol(k));
                            v2[d] = own - (C)ce.k[d];
                        } else if (kind == kSrcConst) {
                            if constexpr (Io<S>::kDev) v2[d] = (C)(ce.k[d] - kWd[d]);
                            else v2[d] = ce.k[d];
                        } else if (kind == kSrcOutlet) {
                            v2[d] = side[(size_t)lcol(k) * 3 + outlet_slot(d)];
                        }
                    }
                }
                NodeOut<S> o2;
                collide_node<S>(v2, p.omega, o2);
                // no pressure fix here, so every direction is stored and the
                // equilibrium is not kept for a fix-up; lanes past column ly-1
                // store its values into the row padding (never read) so that
                // every 128-byte line is written whole
                if constexpr (kTiled) {
                    // output and padding columns go to the output; halo lanes,
                    // and clamped lanes over other tiles' columns, to scratch
                    const uint32_t lu = tid + k * T;
                    const uint32_t vc = cy0 + lu - 1;  // lu = 0 only with the left halo
                    const bool keep = lu >= 1 && (lu <= W || (vc >= ly && vc < pitch));
                    GS *base = keep ? gdst + (size_t)sr2 * 9 * pitch + vc : gscr + (size_t)k * 9 * T + tid;
                    const uint32_t stride = keep ? pitch : T;
#pragma unroll
                    for (int q = 0; q < 9; q++) Io<S>::st(base + (size_t)q * stride, o2.s[q], p.nt != 0);
                } else {
                    const uint32_t yst = min(tid + k * T, pitch - 1);
#pragma unroll
                    for (int q = 0; q < 9; q++)
                        Io<S>::st(pop32<kBig>(gdst, pitch, sr2, q, yst), o2.s[q], p.nt != 0);
                }
                if ((flags2 & kFlagOutletSide) && owned(k)) {
                    gside[(size_t)y * 3 + 0] = v2[3];
                    gside[(size_t)y * 3 + 1] = v2[6];
                    gside[(size_t)y * 3 + 2] = v2[7];
                }
                if constexpr (kProbe) {
                    if (p.probe2 && owned(k) && (int64_t)(r * ly + y) == p.probe_cell) {
                        LBM_GLOBAL double *pr = (LBM_GLOBAL double *)p.probe2;
                        pr[0] = o2.rho;
                        pr[1] = (double)o2.ux;
                        pr[2] = (double)o2.uy;
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < 3; j++) A[k][j] = Cc[k][j];
            B[k][0] = Y[0];
            B[k][1] = Y[2];
            B[k][2] = Y[4];
            Cc[k][0] = Y[1];
            Cc[k][1] = Y[5];
            Cc[k][2] = Y[8];
            D[k] = Y[3];
            cm[k] = c0[k];
            c0[k] = c1[k];
            c1[k] = c2[k];
        }
    };
    // prologue: step-1 rows xa-1 and xa, prefetching rows xa and xa+1
    // (R >= 1, so both exist)
    for (uint32_t it = 0; it < 2; it++) body(it, Flag<true>{}, Flag<false>{});
    // As many stores on this path as on the loop's back edge, so that the
    // wait for the prefetched row at the loop top excludes the latest stores.
    // They go to a small scratch area shared by all blocks (it stays in L2).
#pragma unroll
    for (int k = 0; k < V; k++)
#pragma unroll
        for (int q = 0; q < 9; q++) Io<S>::st(at32(gscr, (k * 9 + q) * T + tid), (C)0, false);
    // steady state: prefetch, step 1, step 2
    for (uint32_t it = 2; it <= R; it++) body(it, Flag<true>{}, Flag<true>{});
    // epilogue: the last step-1 row (no prefetch) and the last output row
    body(R + 1, Flag<false>{}, Flag<true>{});
    if constexpr (kPacked)
        for (uint32_t it = R + 2; it < ln.r_block + 2; it++) __syncthreads();
}

template <class S, int V, bool kProbe, bool kTiled, bool kBig>
__global__ __launch_bounds__(kSweepThreads) void k_sweep2(SweepParams<S> p) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t blk = p.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const SweepLanes ln{threadIdx.x, blockDim.x, smem, 0};
    if constexpr (kTiled) {
        // the tiles of one row range on consecutive blocks (one XCD's L2)
        const uint32_t rb = blk / p.ntiles;
        sweep_block<S, V, kProbe, true, kBig>(p, rb * p.rows, blk - rb * p.ntiles, ln);
    } else {
        sweep_block<S, V, kProbe, false, kBig>(p, blk * p.rows, 0, ln);
    }
}

// Two steps of every job of a batch bucket: block b sweeps rows of job
// block_job[b]; probe samples go to ring slots *step + offset (+1).  kBig:
// the bucket's state buffers are 4 GiB or more (64-bit row bases).
// One column per thread with 32-bit offsets fits 64 registers in FP32 and
// FP16 without spilling: eight waves per SIMD, so that more of a batch's small
// blocks share a CU (+0.5 to +2.4 % on the 256-channel batch with three-step
// groups beside them; FP64 would spill about 115 registers).
template <class S, int V, bool kBig>
__global__ __launch_bounds__(kSweepThreads, sizeof(S) < 8 && V == 1 && !kBig ? 8 : 4) void k_sweep2_batch(const SweepJob *__restrict__ jobs,
                                                                const uint32_t *__restrict__ block_job,
                                                                const int64_t *__restrict__ step, int64_t step_now) {
    using C = typename Io<S>::C;
    extern __shared__ __align__(16) unsigned char smem[];
    // consecutive blocks (neighbouring row ranges of a job) on one XCD, so the
    // halo rows they both read can hit that XCD's L2
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const SweepJob &j = jobs[block_job[blk]];
    SweepParams<S> p;
    p.src = static_cast<const S *>(j.src);
    p.dst = static_cast<S *>(j.dst);
    p.side_prev = static_cast<const C *>(j.side_prev);
    p.side_next = static_cast<C *>(j.side_next);
    p.g = j.g;
    p.bc = j.bc;
    p.omega = j.omega;
    p.probe_cell = j.probe_cell;
    p.probe1 = p.probe2 = nullptr;
    if (j.probe_cell >= 0) {
        const int64_t slot = (step ? *step : step_now) + j.probe_offset;
        if (slot >= 0) p.probe1 = j.probe_ring + 3 * (slot % j.probe_cap);
        if (slot + 1 >= 0) p.probe2 = j.probe_ring + 3 * ((slot + 1) % j.probe_cap);
    }
    p.rows = j.rows;
    p.nt = 0;
    p.xcd = 1;
    p.scratch = static_cast<S *>(j.scratch);
    p.ntiles = 1;
    p.tw = p.lw = j.g.ly;
    // only two columns per thread take the interleaved row order (the
    // one-column instances keep the plain one at no register cost)
    if constexpr (V == 2) {
        p.il.rows = j.il_rows;
        p.il.ranges = j.il_ranges;
        p.il.whole = j.il_whole;
        p.il.magic = j.il_magic;
    }
    sweep_block<S, V, true, false, kBig>(p, (blk - j.block0) * j.rows, 0, SweepLanes{threadIdx.x, blockDim.x, smem, 0});
}

// Two steps of many row ranges in one persistent launch: a grid of resident
// blocks, block b sweeping the subs of packed blocks b, b + G, b + 2G, ...,
// each packed block's subs side by side on their own waves and LDS areas (one
// column per thread, state buffers below 4 GiB).  Waves that belong to no sub
// of a packed block take part in its barriers only.
constexpr int kPackedWaves = 16;

template <class S>
__device__ __forceinline__ void packed_item(const SweepJob *__restrict__ jobs, const PackedSub *__restrict__ subs,
                                            const PackedBlock &pb, const int64_t *__restrict__ step,
                                            int64_t step_now, unsigned char *smem) {
    using C = typename Io<S>::C;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t mine = pb.nsub;
    for (uint32_t k = 0; k < pb.nsub; k++) {
        const PackedSub &q = subs[pb.sub0 + k];
        if (wave >= q.wave0 && wave < q.wave0 + q.waves) mine = k;
    }
    mine = __builtin_amdgcn_readfirstlane(mine);
    if (mine == pb.nsub) {
        // the set-up barrier and one per row iteration, as the swept waves
        for (uint32_t i = 0; i < pb.r_block + 3; i++) __syncthreads();
        return;
    }
    const PackedSub q = subs[pb.sub0 + mine];
    const SweepJob &j = jobs[q.job];
    SweepParams<S> p;
    p.src = static_cast<const S *>(j.src);
    p.dst = static_cast<S *>(j.dst);
    p.side_prev = static_cast<const C *>(j.side_prev);
    p.side_next = static_cast<C *>(j.side_next);
    p.g = j.g;
    p.bc = j.bc;
    p.omega = j.omega;
    p.probe_cell = j.probe_cell;
    p.probe1 = p.probe2 = nullptr;
    if (j.probe_cell >= 0) {
        const int64_t slot = (step ? *step : step_now) + j.probe_offset;
        if (slot >= 0) p.probe1 = j.probe_ring + 3 * (slot % j.probe_cap);
        if (slot + 1 >= 0) p.probe2 = j.probe_ring + 3 * ((slot + 1) % j.probe_cap);
    }
    p.rows = q.rows;
    p.nt = 0;
    p.xcd = 1;
    p.scratch = static_cast<S *>(j.scratch);
    p.ntiles = 1;
    p.tw = p.lw = j.g.ly;
    const SweepLanes ln{threadIdx.x - q.wave0 * 64, q.waves * 64, smem + q.lds_off, pb.r_block};
    sweep_block<S, 1, true, false, false, true>(p, q.xa, 0, ln);
}

template <class S>
__global__ __launch_bounds__(kPackedWaves * 64, 8) void k_sweep2_packed(const SweepJob *__restrict__ jobs,
                                                                         const PackedSub *__restrict__ subs,
                                                                         const PackedBlock *__restrict__ blocks,
                                                                         uint32_t nblocks,
                                                                         const int64_t *__restrict__ step,
                                                                         int64_t step_now) {
    extern __shared__ __align__(16) unsigned char smem[];
    for (uint32_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
        // the previous packed block's last LDS reads come after its last
        // barrier; this one's set-up writes the same LDS
        if (b != blockIdx.x) __syncthreads();
        packed_item<S>(jobs, subs, blocks[b], step, step_now, smem);
    }
}

// ---------------------------------------------------------------------------
// K steps per launch (K >= 3), one column per thread over whole rows
// (ly <= 1024), FP32/FP16 storage.  The two-step sweep's pipeline with K
// stages: at row iteration it, stage 1 collides row x = xa - (K-1) + it and
// publishes its y-moving populations in LDS; after the one barrier of the
// iteration, stage s >= 2 collides row x - (2s - 3) from the y-streamed output
// of stage s - 1 (stage 2 reads what stage 1 published before this barrier,
// stage s >= 3 what stage s - 1 published after the previous one), the last
// stage stores to HBM and the others publish in LDS for the next stage.
// Stages run from the last to the second, so that a stage reads its own node's
// previous-stage values (its reflection sources) before that stage overwrites
// them.  Stage s covers rows [xa - (K - s), xa + R + K - s), so the input is
// read once plus K rows on each side, the output written once per K steps:
// 72 / K B per node and step in FP32 (24 B with K = 3).  Every value goes through the same
// arithmetic as in single steps (identical results).  Needs R >= K - 1 rows in
// every block; no pressure fix, at most kSweepMaxClasses classes.
// ---------------------------------------------------------------------------
constexpr int kSweepMaxK = 4;

template <class S>
struct SweepKParams {
    const S *src;
    S *dst;
    const typename Io<S>::C *side_prev;
    typename Io<S>::C *side_next;
    Geom g;
    BcView bc;
    double omega;
    int64_t probe_cell;   // -1: no probe
    double *probe_ring;   // samples of step s (1-based) go to slot probe_slot + s - 1
    int64_t probe_cap;
    int64_t probe_slot;
    uint32_t rows;        // output rows per block
    int nt;               // non-temporal stores
    S *scratch;           // 9 blockDim.x values, written and never read
};

// Dynamic LDS of the K-step kernel: per stage s < K two exchange buffers of six
// planes of ly + 1 values and the stage's outlet values (3 per column), then
// the incoming outlet values, the class table and one "special" byte per
// (stage-1 row, wave).
struct SweepKLds {
    size_t ex_stride, side, side_in, cls, flag, total;
};

template <class S>
__host__ __device__ inline SweepKLds sweepk_lds_layout(uint32_t ly, uint32_t nclasses, uint32_t rows, int K) {
    using C = typename Io<S>::C;
    SweepKLds L;
    L.ex_stride = align16(12 * ((size_t)ly + 1) * sizeof(S));
    L.side = L.ex_stride * (K - 1);
    L.side_in = align16(L.side + (size_t)(K - 1) * 3 * ly * sizeof(C));
    L.cls = align16(L.side_in + 3 * (size_t)ly * sizeof(C));
    L.flag = align16(L.cls + (size_t)nclasses * sizeof(ClassEntry));
    L.total = L.flag + (size_t)(rows + 2 * (K - 1)) * ((ly + 63) / 64);
    return L;
}

template <int B>
struct IntC {
    static constexpr int value = B;
};

template <class S, int K, bool kBig>
__device__ __forceinline__ void sweepk_block(const SweepKParams<S> &p, uint32_t xa, unsigned char *smem) {
    using C = typename Io<S>::C;
    static_assert(K >= 3 && K <= kSweepMaxK, "K-step sweep: 3 <= K <= kSweepMaxK");
    const Geom &g = p.g;
    const uint32_t ly = g.ly, lx = g.lx, pitch = g.pitch;
    const uint32_t T = blockDim.x, tid = threadIdx.x;
    using GS = LBM_GLOBAL S;
    const GS *gsrc = (const GS *)p.src;
    GS *gdst = (GS *)p.dst;
    GS *gscr = (GS *)p.scratch;
    LBM_GLOBAL C *gside = (LBM_GLOBAL C *)p.side_next;
    const LBM_GLOBAL uint16_t *gcodes = (const LBM_GLOBAL uint16_t *)p.bc.codes;
    const uint32_t R = min(p.rows, lx - xa);
    const uint32_t nw = (ly + 63) / 64;
    const uint32_t cs = ly + 1;  // LDS plane stride (a spare column)
    const uint32_t nrows1 = R + 2 * (K - 1);  // stage-1 rows
    const bool has_bc = p.bc.chunk != nullptr;
    const SweepKLds L = sweepk_lds_layout<S>(ly, has_bc ? p.bc.nclasses : 0, p.rows, K);
    C *side_in = reinterpret_cast<C *>(smem + L.side_in);
    ClassEntry *cls = reinterpret_cast<ClassEntry *>(smem + L.cls);
    uint8_t *wflag = smem + L.flag;
    // exchange buffer b of stage s (1 <= s < K), and the stage's outlet values
    auto exb = [&](int s, uint32_t b) { return reinterpret_cast<S *>(smem + (size_t)(s - 1) * L.ex_stride) + b * 6 * (size_t)cs; };
    auto sideb = [&](int s) { return reinterpret_cast<C *>(smem + L.side) + (size_t)(s - 1) * 3 * ly; };
    // stage-1 row of iteration it (periodic)
    auto row_at = [&](uint32_t it) {
        const uint32_t r = xa + it + lx - (K - 1);
        return r >= lx ? (r - lx >= lx ? r - 2 * lx : r - lx) : r;
    };
    if (has_bc) {
        const uint32_t words = p.bc.nclasses * (uint32_t)(sizeof(ClassEntry) / 4);
        for (uint32_t i = tid; i < words; i += T)
            reinterpret_cast<uint32_t *>(cls)[i] = ((const LBM_GLOBAL uint32_t *)p.bc.classes)[i];
        // stage 1 reads the previous launch's outlet values from LDS: blocks
        // whose stage-1 rows include lx - 1 need them
        if ((p.bc.flags & kProgOutlet) && (xa < (uint32_t)(K - 1) || xa + R + K - 1 >= lx)) {
            for (uint32_t i = tid; i < 3 * ly; i += T) side_in[i] = ((const LBM_GLOBAL C *)p.side_prev)[i];
        }
        for (uint32_t i = tid; i < nrows1 * nw; i += T) {
            const uint32_t rr = i / nw, j = i - rr * nw;
            wflag[i] = ((const LBM_GLOBAL uint8_t *)p.bc.wave)[(size_t)row_at(rr) * g.wpr + j];
        }
    }
    __syncthreads();
    const uint32_t y = min(tid, ly - 1);
    const uint32_t ym = y == 0 ? ly - 1 : y - 1, yp = y + 1 >= ly ? 0 : y + 1;
    const uint32_t jw = min(tid >> 6, nw - 1);
    const uint32_t col = min(tid, ly);  // LDS column written: lanes past the last column use the spare one
    auto special = [&](uint32_t rr) -> bool {
        if (!has_bc) return false;
        return __builtin_amdgcn_readfirstlane(wflag[rr * nw + jw]) != 0;
    };
    auto load_code = [&](uint32_t rr) -> uint32_t {
        if (!special(rr)) return 0;
        return gcodes[row_at(rr) * ly + y];
    };
    auto pull_row = [&](uint32_t rr, uint32_t code, C out[9]) {
        const uint32_t xr = row_at(rr), xm = xr == 0 ? lx - 1 : xr - 1, xp = xr + 1 == lx ? 0 : xr + 1;
        if (special(rr)) {
            const uint32_t w = code ? cls[code].src : 0u;
            pull32_bc<S, kBig>(gsrc, pitch, xm, xr, xp, y, ym, yp, w, out);
        } else {
            pull32<S, kBig>(gsrc, pitch, xm, xr, xp, y, ym, yp, out);
        }
    };
    auto probe_at = [&](int s) -> LBM_GLOBAL double * {
        const int64_t slot = p.probe_slot + s - 1;
        return slot >= 0 ? (LBM_GLOBAL double *)p.probe_ring + 3 * (slot % p.probe_cap) : nullptr;
    };
    // codes: c0 = stage-1 row of this iteration, c1 = the prefetched row, c2 =
    // the row after it; hist[j] = code of the stage-1 row of iteration it-1-j
    uint32_t c0 = load_code(0), c1 = load_code(1), c2 = 0;
    uint32_t hist[2 * K - 3];
#pragma unroll
    for (int j = 0; j < 2 * K - 3; j++) hist[j] = 0;
    // per stage s >= 2 (index s - 2): y-streamed output of stage s - 1 for
    // rows r - 1 (A: directions 1, 5, 8) and r (B: 0, 2, 4; Cc: 1, 5, 8; D: 3)
    C A[K - 1][3], B[K - 1][3], Cc[K - 1][3], D[K - 1];
    // directions 0, 1, 3 of the output of stage s (s < K) published in the
    // previous iteration (index s - 1); stage s + 1 reads them with its LDS row
    C P013[K - 1][3];
#pragma unroll
    for (int s = 0; s < K - 1; s++) {
        D[s] = 0;
#pragma unroll
        for (int j = 0; j < 3; j++) A[s][j] = B[s][j] = Cc[s][j] = P013[s][j] = 0;
    }
    C pf[9];
    pull_row(0, c0, pf);

    // one iteration; kPf: prefetch the next stage-1 row; kS1: stage 1 runs;
    // stage s >= 2 runs where bit s of kMask is set
    auto body = [&](uint32_t it, auto pf_flag, auto s1_flag, auto mask_tag) {
        constexpr bool kPf = decltype(pf_flag)::value;
        constexpr bool kS1 = decltype(s1_flag)::value;
        constexpr int kMask = decltype(mask_tag)::value;
        C v[9];
#pragma unroll
        for (int i = 0; i < 9; i++) v[i] = pf[i];
        if constexpr (kPf) {
            pull_row(it + 1, c1, pf);
            c2 = it + 2 < nrows1 ? load_code(it + 2) : 0u;
        }
        C s013[3] = {0, 0, 0};
        if constexpr (kS1) {
            const uint32_t x = row_at(it);
            uint32_t flags = 0;
            if (c0) {
                const ClassEntry &ce = cls[c0];
                const uint32_t w = ce.src;
                flags = ce.flags;
#pragma unroll
                for (int d = 0; d < 9; d++) {
                    const uint32_t kind = (w >> (2 * d)) & 3u;
                    if (kind == kSrcReflect) {
                        v[d] = v[d] - (C)ce.k[d];  // v holds the node's own opposite population
                    } else if (kind == kSrcConst) {
                        if constexpr (Io<S>::kDev) v[d] = (C)(ce.k[d] - kWd[d]);
                        else v[d] = ce.k[d];
                    } else if (kind == kSrcOutlet) {
                        v[d] = side_in[(size_t)y * 3 + outlet_slot(d)];
                    }
                }
            }
            NodeOut<S> o;
            collide_node<S>(v, p.omega, o);
            if (flags & kFlagOutletSide) {
                C *sd = sideb(1);
                sd[(size_t)y * 3 + 0] = v[3];
                sd[(size_t)y * 3 + 1] = v[6];
                sd[(size_t)y * 3 + 2] = v[7];
            }
            if (p.probe_cell >= 0 && it >= K - 1 && it < R + K - 1 && tid < ly &&
                (int64_t)(x * ly + y) == p.probe_cell) {
                if (LBM_GLOBAL double *pr = probe_at(1)) {
                    pr[0] = o.rho;
                    pr[1] = (double)o.ux;
                    pr[2] = (double)o.uy;
                }
            }
            S *e1 = exb(1, it & 1);
#pragma unroll
            for (int j = 0; j < 6; j++) lds_put<S>(e1 + j * cs + col, o.s[kYM[j]]);
            s013[0] = round_to_storage<S>(o.s[0]);
            s013[1] = round_to_storage<S>(o.s[1]);
            s013[2] = round_to_storage<S>(o.s[3]);
        }
        __syncthreads();
        // y-streamed input of every stage s >= 2, read before any stage publishes
        C Y[K - 1][9];
#pragma unroll
        for (int s = 2; s <= K; s++) {
            const S *e = s == 2 ? exb(1, it & 1) : exb(s - 1, (it + 1) & 1);
            C *Ys = Y[s - 2];
            if (s == 2) {
                Ys[0] = s013[0];
                Ys[1] = s013[1];
                Ys[3] = s013[2];
            } else {
                Ys[0] = P013[s - 2][0];
                Ys[1] = P013[s - 2][1];
                Ys[3] = P013[s - 2][2];
            }
            Ys[2] = lds_get<S>(e + 0 * cs + ym);
            Ys[5] = lds_get<S>(e + 1 * cs + ym);
            Ys[6] = lds_get<S>(e + 2 * cs + ym);
            Ys[4] = lds_get<S>(e + 3 * cs + yp);
            Ys[7] = lds_get<S>(e + 4 * cs + yp);
            Ys[8] = lds_get<S>(e + 5 * cs + yp);
        }
#pragma unroll
        for (int s = K; s >= 2; s--) {
            const int q = s - 2;
            if ((kMask >> s) & 1) {
                const uint32_t r = row_at(it - (2 * s - 3));
                const uint32_t code = hist[2 * s - 4];
                C w2[9];
                w2[0] = B[q][0];
                w2[2] = B[q][1];
                w2[4] = B[q][2];
                w2[1] = A[q][0];
                w2[5] = A[q][1];
                w2[8] = A[q][2];
                w2[3] = Y[q][3];
                w2[6] = Y[q][6];
                w2[7] = Y[q][7];
                uint32_t flags2 = 0;
                if (code) {
                    // reflections read the node's own stage s-1 output: directions
                    // 0, 1, 3 from registers, the y-moving ones from the LDS row
                    // stage s-1 published for this row (stage 1: one iteration
                    // ago; stage s-1 >= 2: two iterations ago, same buffer parity)
                    const S *own = s == 2 ? exb(1, (it + 1) & 1) : exb(s - 1, it & 1);
                    const ClassEntry &ce = cls[code];
                    const uint32_t w = ce.src;
                    flags2 = ce.flags;
#pragma unroll
                    for (int d = 0; d < 9; d++) {
                        const uint32_t kind = (w >> (2 * d)) & 3u;
                        if (kind == kSrcReflect) {
                            const int so = kOpp[d];
                            const C ov = so == 0 ? B[q][0] : so == 1 ? Cc[q][0] : so == 3 ? D[q]
                                                                                    : lds_get<S>(own + kYSlot[so] * cs + y);
                            w2[d] = ov - (C)ce.k[d];
                        } else if (kind == kSrcConst) {
                            if constexpr (Io<S>::kDev) w2[d] = (C)(ce.k[d] - kWd[d]);
                            else w2[d] = ce.k[d];
                        } else if (kind == kSrcOutlet) {
                            w2[d] = sideb(s - 1)[(size_t)y * 3 + outlet_slot(d)];
                        }
                    }
                }
                NodeOut<S> o2;
                collide_node<S>(w2, p.omega, o2);
                const bool own_row = s == K || (it >= (uint32_t)(2 * s - 3 + K - 1) && it < R + 2 * s - 3 + K - 1);
                if (p.probe_cell >= 0 && own_row && tid < ly && (int64_t)(r * ly + y) == p.probe_cell) {
                    if (LBM_GLOBAL double *pr = probe_at(s)) {
                        pr[0] = o2.rho;
                        pr[1] = (double)o2.ux;
                        pr[2] = (double)o2.uy;
                    }
                }
                if (s == K) {
                    // every direction is stored; lanes past column ly-1 store its
                    // values into the row padding (never read)
                    const uint32_t yst = min(tid, pitch - 1);
#pragma unroll
                    for (int d = 0; d < 9; d++) Io<S>::st(pop32<kBig>(gdst, pitch, r, d, yst), o2.s[d], p.nt != 0);
                    if ((flags2 & kFlagOutletSide) && tid < ly) {
                        gside[(size_t)y * 3 + 0] = w2[3];
                        gside[(size_t)y * 3 + 1] = w2[6];
                        gside[(size_t)y * 3 + 2] = w2[7];
                    }
                } else {
                    if (flags2 & kFlagOutletSide) {
                        C *sd = sideb(s);
                        sd[(size_t)y * 3 + 0] = w2[3];
                        sd[(size_t)y * 3 + 1] = w2[6];
                        sd[(size_t)y * 3 + 2] = w2[7];
                    }
                    S *es = exb(s, it & 1);
#pragma unroll
                    for (int j = 0; j < 6; j++) lds_put<S>(es + j * cs + col, o2.s[kYM[j]]);
                    P013[s - 1][0] = round_to_storage<S>(o2.s[0]);
                    P013[s - 1][1] = round_to_storage<S>(o2.s[1]);
                    P013[s - 1][2] = round_to_storage<S>(o2.s[3]);
                }
            }
        }
        // shift the carried rows of every stage
#pragma unroll
        for (int q = 0; q < K - 1; q++) {
#pragma unroll
            for (int j = 0; j < 3; j++) A[q][j] = Cc[q][j];
            B[q][0] = Y[q][0];
            B[q][1] = Y[q][2];
            B[q][2] = Y[q][4];
            Cc[q][0] = Y[q][1];
            Cc[q][1] = Y[q][5];
            Cc[q][2] = Y[q][8];
            D[q] = Y[q][3];
        }
#pragma unroll
        for (int j = 2 * K - 4; j > 0; j--) hist[j] = hist[j - 1];
        hist[0] = c0;
        c0 = c1;
        c1 = c2;
    };
    // prologue: iterations 0 .. 3K-5 (stage s >= 2 from iteration 3s - 4 on),
    // prefetching; R >= K - 1 keeps all of them before the epilogue
    auto prologue = [&](auto it_tag) {
        constexpr int it = decltype(it_tag)::value;
        constexpr int mask = (it >= 2 ? 4 : 0) | (K >= 3 && it >= 5 ? 8 : 0) | (K >= 4 && it >= 8 ? 16 : 0);
        body((uint32_t)it, Flag<true>{}, Flag<true>{}, IntC<mask>{});
    };
    prologue(IntC<0>{});
    prologue(IntC<1>{});
    prologue(IntC<2>{});
    prologue(IntC<3>{});
    prologue(IntC<4>{});
    if constexpr (K >= 4) {
        prologue(IntC<5>{});
        prologue(IntC<6>{});
        prologue(IntC<7>{});
    }
    // as many stores on this path as on the loop's back edge (see sweep_block)
#pragma unroll
    for (int q = 0; q < 9; q++) Io<S>::st(at32(gscr, q * T + tid), (C)0, false);
    constexpr int kAll = (1 << (K + 1)) - 4;  // bits 2 .. K
    for (uint32_t it = 3 * K - 4; it + 3 <= R + 2 * K - 1; it++) body(it, Flag<true>{}, Flag<true>{}, IntC<kAll>{});
    // epilogue: iteration R + 2K - 3 + e; stage 1 only at e = 0, stage s while e <= s - 2
    const uint32_t e0 = R + 2 * K - 3;
    body(e0, Flag<false>{}, Flag<true>{}, IntC<kAll>{});
    body(e0 + 1, Flag<false>{}, Flag<false>{}, IntC<kAll & ~4>{});
    if constexpr (K >= 4) body(e0 + 2, Flag<false>{}, Flag<false>{}, IntC<kAll & ~12>{});
}

// Three steps with 32-bit offsets fit 96 registers without spilling: five
// waves per SIMD, so that two blocks of up to ten waves share a CU.
template <class S, int K, bool kBig>
__global__ __launch_bounds__(kSweepThreads, K == 3 && !kBig ? 5 : 4) void k_sweepk(SweepKParams<S> p) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    sweepk_block<S, K, kBig>(p, blk * p.rows, smem);
}

// K steps of every job of a batch bucket (state buffers below 4 GiB): block b
// sweeps rows of job block_job[b]; the samples of step s go to ring slot
// (*step or step_now) + probe_offset + s - 1.
template <class S, int K>
__global__ __launch_bounds__(kSweepThreads, K == 3 ? 5 : 4) void k_sweepk_batch(const SweepJob *__restrict__ jobs,
                                                                               const uint32_t *__restrict__ block_job,
                                                                               const int64_t *__restrict__ step,
                                                                               int64_t step_now) {
    using C = typename Io<S>::C;
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const SweepJob &j = jobs[block_job[blk]];
    SweepKParams<S> p;
    p.src = static_cast<const S *>(j.src);
    p.dst = static_cast<S *>(j.dst);
    p.side_prev = static_cast<const C *>(j.side_prev);
    p.side_next = static_cast<C *>(j.side_next);
    p.g = j.g;
    p.bc = j.bc;
    p.omega = j.omega;
    p.probe_cell = j.probe_cell;
    p.probe_ring = j.probe_ring;
    p.probe_cap = j.probe_cap;
    p.probe_slot = (step ? *step : step_now) + j.probe_offset;
    p.rows = j.rows;
    p.nt = 0;
    p.scratch = static_cast<S *>(j.scratch);
    sweepk_block<S, K, false>(p, (blk - j.block0) * j.rows, smem);
}

__global__ void k_add_step(int64_t *step, int64_t k) { *step += k; }

// ---------------------------------------------------------------------------
// first collision (from the AoS FP64 step input)
// ---------------------------------------------------------------------------
template <class S>
__global__ __launch_bounds__(kBlock) void k_collide_first(const double *__restrict__ f,
                                                          const double *__restrict__ rho,
                                                          const double *__restrict__ u, S *__restrict__ dst,
                                                          typename Io<S>::C *__restrict__ side, Geom g,
                                                          BcView bc, double omega) {
#pragma clang fp contract(off)
    const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= g.N) return;
    const uint32_t x = row_of(g, n), y = n - x * g.ly;
    double fi[9], e[9], s[9];
#pragma unroll
    for (int i = 0; i < 9; i++) fi[i] = f[(size_t)n * 9 + i];
    const double r = rho[n], ux = u[2 * (size_t)n], uy = u[2 * (size_t)n + 1];
    feq_f64(r, ux, uy, e);
#pragma unroll
    for (int i = 0; i < 9; i++) s[i] = fi[i] + (e[i] - fi[i]) * omega;
    const uint32_t flags = bc.codes ? bc.classes[bc.codes[n]].flags : 0;
    const uint32_t skip = flags >> kSkipShift;
#pragma unroll
    for (int i = 0; i < 9; i++)
        if (!((skip >> i) & 1u)) Io<S>::st(dst + pop(g, x, i, y), rep<S>(s[i], i), false);
    if (flags & kFlagOutletSide) {
        side[(size_t)y * 3 + 0] = rep<S>(fi[3], 3);
        side[(size_t)y * 3 + 1] = rep<S>(fi[6], 6);
        side[(size_t)y * 3 + 2] = rep<S>(fi[7], 7);
    }
    if (flags & kFlagPressureIn) {
        double ein[9];
        feq_f64(bc.rho_in, ux, uy, ein);
        for (int i = 0; i < 9; i++)
            if ((bc.d1 >> i) & 1u) Io<S>::st(dst + pop(g, 0, i, y), rep<S>(ein[i] + (s[i] - e[i]), i), false);
    }
    if (flags & kFlagPressureOut) {
        double eout[9];
        feq_f64(bc.rho_out, ux, uy, eout);
        for (int i = 0; i < 9; i++)
            if ((bc.d2 >> i) & 1u)
                Io<S>::st(dst + pop(g, g.lx - 1, i, y), rep<S>(eout[i] + (s[i] - e[i]), i), false);
    }
}

// ---------------------------------------------------------------------------
// stream + boundary + moments into the reference's AoS FP64 layout
// ---------------------------------------------------------------------------
template <class S>
__device__ __forceinline__ void finalize_node(const S *__restrict__ src, const typename Io<S>::C *side_prev,
                                              const Geom &g, const BcView &bc, uint32_t n, double f[9],
                                              double &rho, double &jx, double &jy, double &ux, double &uy) {
    using C = typename Io<S>::C;
    const uint32_t x = row_of(g, n), y = n - x * g.ly;
    C v[9];
    pull<S>(src, g, x, y, v);
    if (bc.codes) apply_class<S>(bc, src, side_prev, g, n, x, y, v);
#pragma unroll
    for (int i = 0; i < 9; i++) {
        if constexpr (Io<S>::kDev) f[i] = (double)v[i] + kWd[i];
        else f[i] = v[i];
    }
    moments_f64(f, rho, jx, jy);
    velocity_f64(rho, jx, jy, ux, uy);
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_finalize(const S *__restrict__ src,
                                                     const typename Io<S>::C *__restrict__ side_prev, Geom g,
                                                     BcView bc, double *__restrict__ fo, double *__restrict__ ro,
                                                     double *__restrict__ uo, double *__restrict__ partials) {
    const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
    double a0 = 0, a1 = 0, a2 = 0;
    if (n < g.N) {
        double f[9], rho, jx, jy, ux, uy;
        finalize_node<S>(src, side_prev, g, bc, n, f, rho, jx, jy, ux, uy);
        if (fo)
            for (int i = 0; i < 9; i++) fo[(size_t)n * 9 + i] = f[i];
        if (ro) ro[n] = rho;
        if (uo) {
            uo[2 * (size_t)n] = ux;
            uo[2 * (size_t)n + 1] = uy;
        }
        a0 = rho;
        a1 = jx;
        a2 = jy;
    }
    if (partials) block_sum3(a0, a1, a2, partials + 3 * (size_t)blockIdx.x);
}

template <class S>
__global__ void k_probe_cell(const S *__restrict__ src, const typename Io<S>::C *__restrict__ side_prev, Geom g,
                             BcView bc, uint32_t n, double *out) {
    if constexpr (Io<S>::kDev) {
        // the values the fused kernel records for this node
        using C = typename Io<S>::C;
        const uint32_t x = row_of(g, n), y = n - x * g.ly;
        C v[9];
        pull<S>(src, g, x, y, v);
        if (bc.codes) apply_class<S>(bc, src, side_prev, g, n, x, y, v);
        NodeOut<S> o;
        collide_node<S>(v, 1.0, o);
        out[0] = o.rho;
        out[1] = (double)o.ux;
        out[2] = (double)o.uy;
    } else {
        double f[9], rho, jx, jy, ux, uy;
        finalize_node<S>(src, side_prev, g, bc, n, f, rho, jx, jy, ux, uy);
        out[0] = rho;
        out[1] = ux;
        out[2] = uy;
    }
}

__global__ void k_sum_partials(const double *__restrict__ partials, int nblocks, double *out) {
    // one block: strided partial sums in a fixed order, then a fixed tree
    double a = 0, b = 0, c = 0;
    for (int j = threadIdx.x; j < nblocks; j += kBlock) {
        a += partials[3 * (size_t)j];
        b += partials[3 * (size_t)j + 1];
        c += partials[3 * (size_t)j + 2];
    }
    block_sum3(a, b, c, out);
}

__global__ __launch_bounds__(kBlock) void k_moments_sum_aos(const double *__restrict__ f, int64_t n,
                                                            double *partials) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double a0 = 0, a1 = 0, a2 = 0;
    if (c < n) {
        double v[9];
        for (int i = 0; i < 9; i++) v[i] = f[c * 9 + i];
        moments_f64(v, a0, a1, a2);
    }
    block_sum3(a0, a1, a2, partials + 3 * (size_t)blockIdx.x);
}

// ---------------------------------------------------------------------------
// stateless AoS kernels (drop-ins for the reference's numpy functions)
// ---------------------------------------------------------------------------
// |u| per node from AoS FP64 velocities: sqrt(ux*ux + uy*uy) with separate
// roundings, as numpy's norm over the last axis of length 2
// (experiments.py:638-640).
__global__ void k_speed(const double *__restrict__ u, double *__restrict__ out, int64_t n) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const double ux = u[2 * i], uy = u[2 * i + 1];
        out[i] = __builtin_sqrt(ux * ux + uy * uy);
    }
}

template <class T>
__device__ __forceinline__ void feq_t(T rho, T ux, T uy, T e[9]) {
#pragma clang fp contract(off)
    const T mag = sqrt(ux * ux + uy * uy);
    const T usq = mag * mag;
    const T cu[9] = {T(0), ux, uy, -ux, -uy, ux + uy, uy - ux, -(ux + uy), ux - uy};
#pragma unroll
    for (int i = 0; i < 9; i++)
        e[i] = ((T)kWd[i] * rho) * (((T(1) + T(3) * cu[i]) + T(4.5) * (cu[i] * cu[i])) - T(1.5) * usq);
}

template <class T>
__global__ void k_stream_aos(const T *__restrict__ in, T *__restrict__ out, int nx, int ny) {
    const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (n >= (int64_t)nx * ny) return;
    const int x = (int)(n / ny), y = (int)(n - (int64_t)x * ny);
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int sx = (x - kCX[i] + nx) % nx, sy = (y - kCY[i] + ny) % ny;
        out[n * 9 + i] = in[((int64_t)sx * ny + sy) * 9 + i];
    }
}

template <class T>
__global__ void k_equilibrium_aos(const T *__restrict__ rho, const T *__restrict__ u, T *__restrict__ feq,
                                  int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    T e[9];
    feq_t<T>(rho[c], u[2 * c], u[2 * c + 1], e);
    for (int i = 0; i < 9; i++) feq[c * 9 + i] = e[i];
}

template <class T>
__global__ void k_density_aos(const T *__restrict__ f, T *__restrict__ rho, int64_t n) {
#pragma clang fp contract(off)
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const T *a = f + c * 9;
    rho[c] = (((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]))) + a[8];
}

template <class T>
__global__ void k_velocity_aos(const T *__restrict__ rho, const T *__restrict__ f, T *__restrict__ u, int64_t n) {
#pragma clang fp contract(off)
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const T *a = f + c * 9;
    const T jx = ((a[1] + a[5]) + a[8]) - ((a[3] + a[6]) + a[7]);
    const T jy = ((a[2] + a[5]) + a[6]) - ((a[4] + a[7]) + a[8]);
    const T r = rho[c];
    u[2 * c] = r != T(0) ? jx / r : T(0);
    u[2 * c + 1] = r != T(0) ? jy / r : T(0);
}

template <class T>
__global__ void k_collide_aos(const T *__restrict__ f, const T *__restrict__ rho, const T *__restrict__ u,
                              T omega, T *__restrict__ out, int64_t n) {
#pragma clang fp contract(off)
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    T e[9];
    feq_t<T>(rho[c], u[2 * c], u[2 * c + 1], e);
    for (int i = 0; i < 9; i++) out[c * 9 + i] = f[c * 9 + i] + (e[i] - f[c * 9 + i]) * omega;
}

__global__ void k_bc_apply(Geom g, BcView bc, const double *__restrict__ f_pre, double *__restrict__ f_post,
                           const double *__restrict__ f_prev) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= bc.nspecial) return;
    const uint32_t n = bc.special[j];
    const ClassEntry &ce = bc.classes[bc.codes[n]];
    const uint32_t x = n / g.ly, y = n - x * g.ly;
#pragma unroll
    for (int d = 0; d < 9; d++) {
        const uint32_t kind = (ce.src >> (2 * d)) & 3u;
        if (kind == kSrcReflect) f_post[(size_t)n * 9 + d] = f_pre[(size_t)n * 9 + kOpp[d]] - ce.k[d];
        else if (kind == kSrcConst) f_post[(size_t)n * 9 + d] = ce.k[d];
        else if (kind == kSrcOutlet) f_post[(size_t)n * 9 + d] = f_prev[((size_t)(g.lx - 2) * g.ly + y) * 9 + d];
    }
}

__global__ void k_bc_pressure(Geom g, BcView bc, double *__restrict__ f, const double *__restrict__ rho,
                              const double *__restrict__ u) {
#pragma clang fp contract(off)
    const uint32_t y = blockIdx.x * kBlock + threadIdx.x;
    if (y >= g.ly) return;
    const uint32_t a = (g.lx - 2) * g.ly + y, b = (1 % g.lx) * g.ly + y;
    double ein[9], ea[9], eout[9], eb[9];
    feq_f64(bc.rho_in, u[2 * (size_t)a], u[2 * (size_t)a + 1], ein);
    feq_f64(rho[a], u[2 * (size_t)a], u[2 * (size_t)a + 1], ea);
    for (int i = 0; i < 9; i++)
        if ((bc.d1 >> i) & 1u) f[(size_t)y * 9 + i] = ein[i] + (f[(size_t)a * 9 + i] - ea[i]);
    feq_f64(bc.rho_out, u[2 * (size_t)b], u[2 * (size_t)b + 1], eout);
    feq_f64(rho[b], u[2 * (size_t)b], u[2 * (size_t)b + 1], eb);
    const size_t m = (size_t)(g.lx - 1) * g.ly + y;
    for (int i = 0; i < 9; i++)
        if ((bc.d2 >> i) & 1u) f[m * 9 + i] = eout[i] + (f[(size_t)b * 9 + i] - eb[i]);
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Blocks of k_fused with V nodes per thread: lx * wpr waves, kBlock / 64 waves
// per block and node slot.
inline int64_t fused_grid(const Geom &g, int V) {
    const int64_t per = (int64_t)(kBlock / 64) * V;
    return ((int64_t)g.lx * g.wpr + per - 1) / per;
}

template <class S, int V>
hipError_t fused_dispatch(const FusedParams<S> &p, unsigned grid, bool probe, bool reduce, hipStream_t s) {
    if (probe && reduce) hipLaunchKernelGGL((k_fused<S, V, true, true>), dim3(grid), dim3(kBlock), 0, s, p);
    else if (probe) hipLaunchKernelGGL((k_fused<S, V, true, false>), dim3(grid), dim3(kBlock), 0, s, p);
    else if (reduce) hipLaunchKernelGGL((k_fused<S, V, false, true>), dim3(grid), dim3(kBlock), 0, s, p);
    else hipLaunchKernelGGL((k_fused<S, V, false, false>), dim3(grid), dim3(kBlock), 0, s, p);
    return hipGetLastError();
}

template <class S>
hipError_t fused_typed(const Geom &g, const StateView &st, int src, const BcView &bc, double omega,
                       const ProbeView &pr, int64_t slot, double *partials, const LaunchCfg &cfg,
                       hipStream_t s, int *nblocks) {
    using C = typename Io<S>::C;
    FusedParams<S> p;
    p.src = static_cast<const S *>(st.plane[src]);
    p.dst = static_cast<S *>(st.plane[src ^ 1]);
    p.side_prev = static_cast<const C *>(st.side[src]);
    p.side_next = static_cast<C *>(st.side[src ^ 1]);
    p.g = g;
    p.bc = bc;
    p.omega = omega;
    const bool probe = pr.cell >= 0 && slot >= 0 && pr.ring;
    p.probe_cell = probe ? pr.cell : -1;
    p.probe_out = probe ? pr.ring + 3 * (slot % pr.capacity) : nullptr;
    p.partials = partials;
    p.xcd = cfg.xcd_remap;
    p.nt = cfg.nt_stores;
    const int V = cfg.cells_per_thread;
    if (V == 0) {
        constexpr int V16 = Pack<S>::V;
        if (g.ly % V16 != 0) return hipErrorInvalidValue;
        const unsigned grid = (unsigned)(((int64_t)g.N / V16 + kBlock - 1) / kBlock);
        if (nblocks) *nblocks = (int)grid;
        const bool reduce = partials != nullptr;
        if (probe && reduce) hipLaunchKernelGGL((k_fused16<S, true, true>), dim3(grid), dim3(kBlock), 0, s, p);
        else if (probe) hipLaunchKernelGGL((k_fused16<S, true, false>), dim3(grid), dim3(kBlock), 0, s, p);
        else if (reduce) hipLaunchKernelGGL((k_fused16<S, false, true>), dim3(grid), dim3(kBlock), 0, s, p);
        else hipLaunchKernelGGL((k_fused16<S, false, false>), dim3(grid), dim3(kBlock), 0, s, p);
        return hipGetLastError();
    }
    const unsigned grid = (unsigned)fused_grid(g, V);
    if (nblocks) *nblocks = (int)grid;
    switch (V) {
        case 1: return fused_dispatch<S, 1>(p, grid, probe, partials != nullptr, s);
        case 2: return fused_dispatch<S, 2>(p, grid, probe, partials != nullptr, s);
        case 4: return fused_dispatch<S, 4>(p, grid, probe, partials != nullptr, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

int64_t batch_job_blocks(const Geom &g, int nodes_per_thread) { return fused_grid(g, nodes_per_thread); }

hipError_t launch_fused_batch(int storage, int v, const BatchJob *jobs, const uint32_t *block_job, unsigned nblocks,
                              int64_t *step, unsigned *done, hipStream_t s) {
    if (nblocks == 0) return hipSuccess;
#define LBM_BATCH(T, VV) hipLaunchKernelGGL((k_fused_batch<T, VV>), dim3(nblocks), dim3(kBlock), 0, s, jobs, block_job, \
                                            step)
    if (v == 1) {
        if (storage == LBM_STORE_F64) LBM_BATCH(double, 1);
        else if (storage == LBM_STORE_F32) LBM_BATCH(float, 1);
        else LBM_BATCH(__half, 1);
    } else if (v == 2) {
        if (storage == LBM_STORE_F64) LBM_BATCH(double, 2);
        else if (storage == LBM_STORE_F32) LBM_BATCH(float, 2);
        else LBM_BATCH(__half, 2);
    } else {
        return hipErrorInvalidValue;
    }
#undef LBM_BATCH
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    (void)done;
    hipLaunchKernelGGL(k_next_step, dim3(1), dim3(1), 0, s, step);
    return hipGetLastError();
}

int vec16_nodes(int storage) { return storage == LBM_STORE_F64 ? 2 : (storage == LBM_STORE_F32 ? 4 : 8); }

int fused_blocks(const Geom &g, const LaunchCfg &cfg) {
    // 16-byte groups (0) need fewer blocks than one node per thread
    return (int)fused_grid(g, cfg.cells_per_thread ? cfg.cells_per_thread : 1);
}

int partial_rows_needed(const Geom &g, const LaunchCfg &cfg) {
    const int a = fused_blocks(g, cfg);
    const int b = (int)grid_for(g.N);
    return a > b ? a : b;
}

hipError_t launch_fused(int storage, const Geom &g, const StateView &st, int src, const BcView &bc,
                        double omega, const ProbeView &pr, int64_t slot, double *partials,
                        const LaunchCfg &cfg, hipStream_t s, int *nblocks) {
    switch (storage) {
        case LBM_STORE_F64: return fused_typed<double>(g, st, src, bc, omega, pr, slot, partials, cfg, s, nblocks);
        case LBM_STORE_F32: return fused_typed<float>(g, st, src, bc, omega, pr, slot, partials, cfg, s, nblocks);
        case LBM_STORE_F16: return fused_typed<__half>(g, st, src, bc, omega, pr, slot, partials, cfg, s, nblocks);
    }
    return hipErrorInvalidValue;
}

namespace {

constexpr size_t kLdsLimit = 160 * 1024;

size_t sweep_lds_for(int storage, uint32_t ly, uint32_t nclasses, uint32_t rows) {
    switch (storage) {
        case LBM_STORE_F64: return sweep_lds_layout<double>(ly, nclasses, rows).total;
        case LBM_STORE_F32: return sweep_lds_layout<float>(ly, nclasses, rows).total;
        default: return sweep_lds_layout<__half>(ly, nclasses, rows).total;
    }
}

// The k_sweep2 instance of a launch.
template <class S, int VV, bool kProbe>
const void *sweep_fn_p(bool tiled, bool big) {
    if (tiled)
        return big ? reinterpret_cast<const void *>(k_sweep2<S, VV, kProbe, true, true>)
                   : reinterpret_cast<const void *>(k_sweep2<S, VV, kProbe, true, false>);
    return big ? reinterpret_cast<const void *>(k_sweep2<S, VV, kProbe, false, true>)
               : reinterpret_cast<const void *>(k_sweep2<S, VV, kProbe, false, false>);
}
template <class S, int VV>
const void *sweep_fn(bool probing, bool tiled, bool big) {
    return probing ? sweep_fn_p<S, VV, true>(tiled, big) : sweep_fn_p<S, VV, false>(tiled, big);
}

// Row ranges the GPU holds at once for the two-step launch of this lattice
// and shape (resident blocks per CU, from LDS and registers, times the CU
// count, divided by the column tiles of a range); 0 if the runtime cannot say.
template <class S>
int sweep_resident_ranges(const Geom &g, const BcView &bc, const LaunchCfg &cfg) {
    const SweepShape sh = sweep_shape(sizeof(S) == 8 ? LBM_STORE_F64 : (sizeof(S) == 4 ? LBM_STORE_F32 : LBM_STORE_F16),
                                      g.ly, cfg.sweep_tile, cfg.sweep_v);
    const bool big = sweep_big(sizeof(S) == 8 ? LBM_STORE_F64 : (sizeof(S) == 4 ? LBM_STORE_F32 : LBM_STORE_F16), g);
    constexpr int V2 = sizeof(S) == 8 ? 1 : 2;
    int dev = 0, cus = 0, per_cu = 0;
    const size_t lds_est = sweep_lds_layout<S>(sh.lw, bc.chunk ? bc.nclasses : 0, 64).total;
    const void *fn = sh.v == 1 ? sweep_fn<S, 1>(false, sh.ntiles > 1, big) : sweep_fn<S, V2>(false, sh.ntiles > 1, big);
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_est) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)sh.threads, lds_est) == hipSuccess)
        return std::max(1, cus * (per_cu > 0 ? per_cu : 1) / (int)sh.ntiles);
    return 0;
}

// Output rows per block of the two-step launch for this lattice and shape.
template <class S>
uint32_t sweep_rows_typed(const Geom &g, const BcView &bc, const LaunchCfg &cfg) {
    const int slots = cfg.sweep_rows <= 0 ? sweep_resident_ranges<S>(g, bc, cfg) : 0;
    return (uint32_t)sweep_rows_for(g, cfg, slots);
}

template <class S>
hipError_t sweep_typed(const Geom &g, const StateView &st, int src, const BcView &bc, double omega,
                       const ProbeView &pr, int64_t slot1, int64_t slot2, const LaunchCfg &cfg, hipStream_t s) {
    using C = typename Io<S>::C;
    SweepParams<S> p;
    p.src = static_cast<const S *>(st.plane[src]);
    p.dst = static_cast<S *>(st.plane[src ^ 1]);
    p.side_prev = static_cast<const C *>(st.side[src]);
    p.side_next = static_cast<C *>(st.side[src ^ 1]);
    p.g = g;
    p.bc = bc;
    p.omega = omega;
    const bool probing = pr.cell >= 0 && pr.ring && (slot1 >= 0 || slot2 >= 0);
    p.probe_cell = probing ? pr.cell : -1;
    p.probe1 = probing && slot1 >= 0 ? pr.ring + 3 * (slot1 % pr.capacity) : nullptr;
    p.probe2 = probing && slot2 >= 0 ? pr.ring + 3 * (slot2 % pr.capacity) : nullptr;
    p.nt = cfg.nt_stores;
    p.xcd = cfg.xcd_remap;
    p.scratch = static_cast<S *>(st.scratch);
    const SweepShape sh = sweep_shape(sizeof(S) == 8 ? LBM_STORE_F64 : (sizeof(S) == 4 ? LBM_STORE_F32 : LBM_STORE_F16),
                                      g.ly, cfg.sweep_tile, cfg.sweep_v);
    const int V = sh.v;
    const unsigned threads = sh.threads;
    const bool tiled = sh.ntiles > 1;
    p.ntiles = sh.ntiles;
    p.tw = sh.tw;
    p.lw = sh.lw;
    if (!p.scratch || (size_t)9 * V * threads * sizeof(S) > kSweepScratchBytes) return hipErrorInvalidValue;
    // state buffers of 4 GiB and more take 64-bit row bases
    const bool big = sweep_big(sizeof(S) == 8 ? LBM_STORE_F64 : (sizeof(S) == 4 ? LBM_STORE_F32 : LBM_STORE_F16), g);
    constexpr int V2 = sizeof(S) == 8 ? 1 : 2;  // FP64 rows are one column per thread
    if (V != 1 && V != V2) return hipErrorInvalidValue;
    auto pick = [&](bool probe) { return V == 1 ? sweep_fn<S, 1>(probe, tiled, big) : sweep_fn<S, V2>(probe, tiled, big); };
    p.rows = sweep_rows_typed<S>(g, bc, cfg);
    // the buffers in the range-interleaved order of exactly these row ranges
    if (cfg.sweep_il) {
        p.il = row_layout(g.lx, p.rows);
        if (p.il.rows != p.rows) return hipErrorInvalidValue;
    }
    const unsigned grid = (g.lx + p.rows - 1) / p.rows * sh.ntiles;
    const size_t lds = sweep_lds_layout<S>(sh.lw, bc.chunk ? bc.nclasses : 0, p.rows).total;
    const void *fn = pick(probing);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    void *args[] = {&p};
    e = hipLaunchKernel(fn, dim3(grid), dim3(threads), args, lds, s);
    return e != hipSuccess ? e : hipGetLastError();
}

}  // namespace

SweepShape sweep_shape(int storage, uint32_t ly, int tile_threads, int tile_v) {
    // one column per thread for FP64, two for FP32/FP16: the shapes that fit
    // the 128 registers of a 1024-thread block without spilling.  A block
    // holds at most tmax threads: 1024, or the tile width asked for; FP32/FP16
    // tiles may ask for one column per thread
    const uint32_t vmax = storage == LBM_STORE_F64 || tile_v == 1 ? 1 : 2;
    const uint32_t tmax = tile_threads > 0 ? std::min((uint32_t)tile_threads, (uint32_t)kSweepThreads)
                                           : (uint32_t)kSweepThreads;
    const uint32_t maxc = tmax * vmax;
    SweepShape sh;
    if (ly <= maxc) {
        sh.ntiles = 1;
        sh.tw = sh.lw = ly;
    } else {
        // the fewest tiles whose output columns plus two halo columns fit a block
        sh.ntiles = (ly + maxc - 3) / (maxc - 2);
        sh.tw = (ly + sh.ntiles - 1) / sh.ntiles;
        sh.lw = sh.tw + 2;
    }
    sh.v = (storage == LBM_STORE_F64 || sh.lw <= tmax) ? 1 : 2;
    sh.threads = ((sh.lw + sh.v - 1) / sh.v + 63) / 64 * 64;
    return sh;
}

bool sweep_supported(int storage, const Geom &g, const BcView &bc, int tile_threads, int tile_v) {
    if (bc.flags & kProgPressure) return false;
    if (bc.chunk && bc.nclasses > kSweepMaxClasses) return false;
    if (g.lx < 3 || g.ly < 1) return false;
    if (tile_threads < 0 || tile_threads % 64 != 0 || tile_threads > kSweepThreads) return false;
    if (tile_v < 0 || tile_v > 2) return false;
    const SweepShape sh = sweep_shape(storage, g.ly, tile_threads, tile_v);
    // every tile holds at least one output column
    if (sh.ntiles > 1 && (uint64_t)sh.tw * (sh.ntiles - 1) >= g.ly) return false;
    return sweep_lds_for(storage, sh.lw, bc.chunk ? bc.nclasses : 0, kSweepMaxRows) <= kLdsLimit;
}

bool sweep_batch_supported(int storage, const Geom &g, const BcView &bc) {
    // whole rows only
    return sweep_supported(storage, g, bc) && sweep_shape(storage, g.ly).ntiles == 1;
}

int sweep_rows_for(const Geom &g, const LaunchCfg &cfg, int slots) {
    int r;
    if (cfg.sweep_rows > 0) {
        r = cfg.sweep_rows;
    } else {
        // the fewest whole waves of `slots` resident blocks (the fewer the
        // blocks, the fewer halo rows are read twice), rows spread evenly over
        // them.  A lattice of fewer than `slots` rows gets one row per block:
        // it is launch-bound, and filling the GPU matters more than the halo
        // rows (300x300 FP64, 3000 steps: 0.030 s with 8 rows per block, 0.013
        // s with 1-2; FP32 0.016 s and 0.0074 s)
        const uint32_t s = slots > 0 ? (uint32_t)slots : 256u;
        const uint32_t waves = (g.lx + s * kSweepMaxRows - 1) / (s * kSweepMaxRows);
        const uint32_t w = waves > 0 ? waves : 1u;
        r = (int)((g.lx + w * s - 1) / (w * s));
        if (r < 1) r = 1;
    }
    if (r > (int)kSweepMaxRows) r = (int)kSweepMaxRows;
    return r < (int)g.lx ? r : (int)g.lx;
}

hipError_t launch_sweep2(int storage, const Geom &g, const StateView &st, int src, const BcView &bc,
                         double omega, const ProbeView &pr, int64_t slot1, int64_t slot2, const LaunchCfg &cfg,
                         hipStream_t s) {
    if (!sweep_supported(storage, g, bc, cfg.sweep_tile, cfg.sweep_v)) return hipErrorInvalidValue;
    switch (storage) {
        case LBM_STORE_F64: return sweep_typed<double>(g, st, src, bc, omega, pr, slot1, slot2, cfg, s);
        case LBM_STORE_F32: return sweep_typed<float>(g, st, src, bc, omega, pr, slot1, slot2, cfg, s);
        case LBM_STORE_F16: return sweep_typed<__half>(g, st, src, bc, omega, pr, slot1, slot2, cfg, s);
    }
    return hipErrorInvalidValue;
}

int sweep2_rows(int storage, const Geom &g, const BcView &bc, const LaunchCfg &cfg) {
    switch (storage) {
        case LBM_STORE_F64: return (int)sweep_rows_typed<double>(g, bc, cfg);
        case LBM_STORE_F32: return (int)sweep_rows_typed<float>(g, bc, cfg);
        case LBM_STORE_F16: return (int)sweep_rows_typed<__half>(g, bc, cfg);
    }
    return 0;
}

int sweep2_resident_ranges(int storage, const Geom &g, const BcView &bc, const LaunchCfg &cfg) {
    switch (storage) {
        case LBM_STORE_F64: return sweep_resident_ranges<double>(g, bc, cfg);
        case LBM_STORE_F32: return sweep_resident_ranges<float>(g, bc, cfg);
        case LBM_STORE_F16: return sweep_resident_ranges<__half>(g, bc, cfg);
    }
    return 0;
}

namespace {
// Row x of src to its slot in dst (kTo) or back (!kTo): one block per row,
// 16-byte copies of its nine planes (9 pitch values, a multiple of 128
// bytes).  Block 0 also copies the outlet buffer (n_side 4-byte words).
template <bool kTo>
__global__ __launch_bounds__(kBlock) void k_row_layout(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                       uint32_t row16, RowLayout L, const uint32_t *__restrict__ side_src,
                                                       uint32_t *__restrict__ side_dst, uint32_t n_side) {
    const uint32_t x = blockIdx.x;
    const uint32_t sl = row_slot(L, x);
    const uint4 *in = src + (size_t)(kTo ? x : sl) * row16;
    uint4 *out = dst + (size_t)(kTo ? sl : x) * row16;
    for (uint32_t i = threadIdx.x; i < row16; i += kBlock) out[i] = in[i];
    if (x == 0)
        for (uint32_t i = threadIdx.x; i < n_side; i += kBlock) side_dst[i] = side_src[i];
}
}  // namespace

hipError_t launch_row_layout(int storage, const Geom &g, const StateView &st, int src, uint32_t rows, bool to_ilv,
                             hipStream_t s) {
    const RowLayout L = row_layout(g.lx, rows);
    if (L.rows != rows || rows == 0) return hipErrorInvalidValue;
    const size_t elem = storage == LBM_STORE_F64 ? 8 : (storage == LBM_STORE_F32 ? 4 : 2);
    const size_t row_bytes = 9 * (size_t)g.pitch * elem;
    if (row_bytes % 16 != 0 || row_bytes / 16 > 0xffffffffull) return hipErrorInvalidValue;
    // outlet buffers: 3 ly values of the compute type (FP64 or FP32)
    const size_t side_bytes = 3 * (size_t)g.ly * (storage == LBM_STORE_F64 ? 8 : 4);
    const void *in = st.plane[src];
    void *out = st.plane[src ^ 1];
    if (to_ilv)
        hipLaunchKernelGGL(k_row_layout<true>, dim3(g.lx), dim3(kBlock), 0, s, static_cast<const uint4 *>(in),
                           static_cast<uint4 *>(out), (uint32_t)(row_bytes / 16), L,
                           static_cast<const uint32_t *>(st.side[src]), static_cast<uint32_t *>(st.side[src ^ 1]),
                           (uint32_t)(side_bytes / 4));
    else
        hipLaunchKernelGGL(k_row_layout<false>, dim3(g.lx), dim3(kBlock), 0, s, static_cast<const uint4 *>(in),
                           static_cast<uint4 *>(out), (uint32_t)(row_bytes / 16), L,
                           static_cast<const uint32_t *>(st.side[src]), static_cast<uint32_t *>(st.side[src ^ 1]),
                           (uint32_t)(side_bytes / 4));
    return hipGetLastError();
}

namespace {
template <class S>
const void *sweepk_fn(int K, bool big) {
    switch (K) {
        case 3: return big ? reinterpret_cast<const void *>(k_sweepk<S, 3, true>)
                           : reinterpret_cast<const void *>(k_sweepk<S, 3, false>);
        case 4: return big ? reinterpret_cast<const void *>(k_sweepk<S, 4, true>)
                           : reinterpret_cast<const void *>(k_sweepk<S, 4, false>);
    }
    return nullptr;
}

size_t sweepk_lds_for(int storage, uint32_t ly, uint32_t nclasses, uint32_t rows, int K) {
    return storage == LBM_STORE_F32 ? sweepk_lds_layout<float>(ly, nclasses, rows, K).total
                                    : sweepk_lds_layout<__half>(ly, nclasses, rows, K).total;
}

template <class S>
hipError_t sweepk_typed(int K, const Geom &g, const StateView &st, int src, const BcView &bc, double omega,
                        const ProbeView &pr, int64_t slot, const LaunchCfg &cfg, hipStream_t s) {
    using C = typename Io<S>::C;
    constexpr int storage = sizeof(S) == 4 ? LBM_STORE_F32 : LBM_STORE_F16;
    SweepKParams<S> p;
    p.src = static_cast<const S *>(st.plane[src]);
    p.dst = static_cast<S *>(st.plane[src ^ 1]);
    p.side_prev = static_cast<const C *>(st.side[src]);
    p.side_next = static_cast<C *>(st.side[src ^ 1]);
    p.g = g;
    p.bc = bc;
    p.omega = omega;
    const bool probing = pr.cell >= 0 && pr.ring && pr.capacity > 0 && slot + K - 1 >= 0;
    p.probe_cell = probing ? pr.cell : -1;
    p.probe_ring = pr.ring;
    p.probe_cap = pr.capacity;
    p.probe_slot = slot;
    p.nt = cfg.nt_stores;
    p.scratch = static_cast<S *>(st.scratch);
    const unsigned threads = (g.ly + 63) / 64 * 64;
    if (!p.scratch || (size_t)9 * threads * sizeof(S) > kSweepScratchBytes) return hipErrorInvalidValue;
    const void *fn = sweepk_fn<S>(K, sweep_big(storage, g));
    if (!fn) return hipErrorInvalidValue;
    const uint32_t ncls = bc.chunk ? bc.nclasses : 0;
    int slots = 0;
    if (cfg.sweep_rows <= 0) {
        int dev = 0, cus = 0, per_cu = 0;
        const size_t lds_est = sweepk_lds_layout<S>(g.ly, ncls, 64, K).total;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_est) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)threads, lds_est) == hipSuccess)
            slots = std::max(1, cus * (per_cu > 0 ? per_cu : 1));
    }
    p.rows = (uint32_t)sweepk_rows(g, sweep_rows_for(g, cfg, slots), K);
    const size_t lds = sweepk_lds_layout<S>(g.ly, ncls, p.rows, K).total;
    if (lds > kLdsLimit) return hipErrorInvalidValue;
    const unsigned grid = (g.lx + p.rows - 1) / p.rows;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    void *args[] = {&p};
    e = hipLaunchKernel(fn, dim3(grid), dim3(threads), args, lds, s);
    return e != hipSuccess ? e : hipGetLastError();
}
}  // namespace

size_t sweepk_lds_bytes(int storage, uint32_t ly, uint32_t nclasses, uint32_t rows, int K) {
    return sweepk_lds_for(storage, ly, nclasses, rows, K);
}

namespace {
const void *sweepk_batch_fn(int storage, int K) {
    if (K != 3) return nullptr;
    if (storage == LBM_STORE_F32) return reinterpret_cast<const void *>(k_sweepk_batch<float, 3>);
    if (storage == LBM_STORE_F16) return reinterpret_cast<const void *>(k_sweepk_batch<__half, 3>);
    return nullptr;
}
}  // namespace

hipError_t prepare_sweepk_batch(int storage, int K, size_t lds) {
    const void *fn = sweepk_batch_fn(storage, K);
    if (!fn || lds > kLdsLimit) return hipErrorInvalidValue;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_sweepk_batch(int storage, int K, const SweepJob *jobs, const uint32_t *block_job, unsigned nblocks,
                               unsigned threads, size_t lds, const int64_t *step, int64_t step_now, hipStream_t s) {
    if (nblocks == 0) return hipSuccess;
    const void *fn = sweepk_batch_fn(storage, K);
    if (!fn || threads == 0 || threads > (unsigned)kSweepThreads || threads % 64 != 0 || lds > kLdsLimit)
        return hipErrorInvalidValue;
    void *args[] = {&jobs, &block_job, &step, &step_now};
    hipError_t e = hipLaunchKernel(fn, dim3(nblocks), dim3(threads), args, lds, s);
    return e != hipSuccess ? e : hipGetLastError();
}

bool sweepk_supported(int storage, const Geom &g, const BcView &bc, int K) {
    if (K < 3 || K > kSweepMaxK) return false;
    if (storage != LBM_STORE_F32 && storage != LBM_STORE_F16) return false;
    if (bc.flags & kProgPressure) return false;
    if (bc.chunk && bc.nclasses > kSweepMaxClasses) return false;
    if (g.lx < 3 || g.ly < 1 || g.ly > (uint32_t)kSweepThreads) return false;
    return sweepk_lds_for(storage, g.ly, bc.chunk ? bc.nclasses : 0, kSweepMaxRows + 4, K) <= kLdsLimit;
}

int sweepk_rows(const Geom &g, int rows, int K) {
    const int lx = (int)g.lx;
    int r = std::max(rows, K - 1);
    while (r < lx) {
        const int last = lx - (lx + r - 1) / r * r + r;  // rows of the last block
        if (last >= K - 1) break;
        r++;
    }
    return std::min(r, lx);
}

hipError_t launch_sweepk(int storage, int K, const Geom &g, const StateView &st, int src, const BcView &bc,
                         double omega, const ProbeView &pr, int64_t probe_slot, const LaunchCfg &cfg,
                         hipStream_t s) {
    if (!sweepk_supported(storage, g, bc, K)) return hipErrorInvalidValue;
    if (storage == LBM_STORE_F32) return sweepk_typed<float>(K, g, st, src, bc, omega, pr, probe_slot, cfg, s);
    return sweepk_typed<__half>(K, g, st, src, bc, omega, pr, probe_slot, cfg, s);
}

namespace {
// The k_sweep2_batch instance of a bucket, or null for a shape it does not have.
const void *sweep_batch_fn(int storage, int v, bool big) {
    auto pick = [&](auto small_fn, auto big_fn) { return big ? reinterpret_cast<const void *>(big_fn)
                                                             : reinterpret_cast<const void *>(small_fn); };
    switch (storage) {
        case LBM_STORE_F64:
            return v == 1 ? pick(k_sweep2_batch<double, 1, false>, k_sweep2_batch<double, 1, true>) : nullptr;
        case LBM_STORE_F32:
            return v == 1 ? pick(k_sweep2_batch<float, 1, false>, k_sweep2_batch<float, 1, true>)
                          : (v == 2 ? pick(k_sweep2_batch<float, 2, false>, k_sweep2_batch<float, 2, true>) : nullptr);
        case LBM_STORE_F16:
            return v == 1 ? pick(k_sweep2_batch<__half, 1, false>, k_sweep2_batch<__half, 1, true>)
                          : (v == 2 ? pick(k_sweep2_batch<__half, 2, false>, k_sweep2_batch<__half, 2, true>) : nullptr);
    }
    return nullptr;
}
}  // namespace

bool sweep_big(int storage, const Geom &g) {
    const uint64_t elem = storage == LBM_STORE_F64 ? 8 : (storage == LBM_STORE_F32 ? 4 : 2);
    return 9 * (uint64_t)g.lx * g.pitch * elem >= (1ull << 32);
}

hipError_t prepare_sweep2_batch(int storage, int v, bool big, size_t lds) {
    if (lds > kLdsLimit) return hipErrorInvalidValue;
    const void *fn = sweep_batch_fn(storage, v, big);
    if (!fn) return hipErrorInvalidValue;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

int sweep2_batch_per_cu(int storage, int v, bool big, unsigned threads, size_t lds) {
    const void *fn = sweep_batch_fn(storage, v, big);
    int per_cu = 0;
    // the limit raised to the most LDS a block may take, never lowered: a
    // batch launching this kernel meanwhile may need more than `lds`
    if (!fn || threads == 0 || lds > kLdsLimit ||
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsLimit) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)threads, lds) != hipSuccess)
        return 1;
    return per_cu > 0 ? per_cu : 1;
}

hipError_t launch_sweep2_batch(int storage, int v, bool big, const SweepJob *jobs, const uint32_t *block_job,
                               unsigned nblocks, unsigned threads, size_t lds, const int64_t *step, int64_t step_now,
                               hipStream_t s) {
    if (nblocks == 0) return hipSuccess;
    if (threads == 0 || threads > (unsigned)kSweepThreads || threads % 64 != 0 || lds > kLdsLimit)
        return hipErrorInvalidValue;
    const void *fn = sweep_batch_fn(storage, v, big);
    if (!fn) return hipErrorInvalidValue;
    void *args[] = {&jobs, &block_job, &step, &step_now};
    hipError_t e = hipLaunchKernel(fn, dim3(nblocks), dim3(threads), args, lds, s);
    return e != hipSuccess ? e : hipGetLastError();
}

namespace {
// FP32 and FP16 only: at the 64 registers that let two 16-wave blocks share a
// CU, the FP64 instance spills about 120 of them.
const void *sweep_packed_fn(int storage) {
    switch (storage) {
        case LBM_STORE_F32: return reinterpret_cast<const void *>(k_sweep2_packed<float>);
        case LBM_STORE_F16: return reinterpret_cast<const void *>(k_sweep2_packed<__half>);
    }
    return nullptr;
}
}  // namespace

uint32_t packed_block_waves() { return kPackedWaves; }

uint32_t packed_sub_lds(int storage, uint32_t ly, uint32_t nclasses, uint32_t rows) {
    size_t b = 0;
    switch (storage) {
        case LBM_STORE_F64: b = sweep_lds_layout<double>(ly, nclasses, rows).total; break;
        case LBM_STORE_F32: b = sweep_lds_layout<float>(ly, nclasses, rows).total; break;
        case LBM_STORE_F16: b = sweep_lds_layout<__half>(ly, nclasses, rows).total; break;
    }
    return (uint32_t)align16(b);
}

hipError_t prepare_sweep2_packed(int storage, size_t lds) {
    const void *fn = sweep_packed_fn(storage);
    if (!fn || lds > kLdsLimit) return hipErrorInvalidValue;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t packed_grid(int storage, size_t lds, unsigned nblocks, bool persistent, unsigned *grid) {
    const void *fn = sweep_packed_fn(storage);
    if (!fn || lds > kLdsLimit) return hipErrorInvalidValue;
    if (!persistent) {
        *grid = nblocks;
        return hipSuccess;
    }
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kPackedWaves * 64, lds);
    if (e != hipSuccess) return e;
    const unsigned resident = (unsigned)std::max(1, cus * std::max(1, per_cu));
    *grid = std::min(nblocks, resident);
    return hipSuccess;
}

hipError_t launch_sweep2_packed(int storage, const SweepJob *jobs, const PackedSub *subs, const PackedBlock *blocks,
                                unsigned nblocks, unsigned grid, size_t lds, const int64_t *step, int64_t step_now,
                                hipStream_t s) {
    if (nblocks == 0) return hipSuccess;
    const void *fn = sweep_packed_fn(storage);
    if (!fn || lds > kLdsLimit || grid == 0) return hipErrorInvalidValue;
    void *args[] = {&jobs, &subs, &blocks, &nblocks, &step, &step_now};
    hipError_t e = hipLaunchKernel(fn, dim3(grid), dim3(kPackedWaves * 64), args, lds, s);
    return e != hipSuccess ? e : hipGetLastError();
}

// Streaming copy, 16 bytes per lane, one chunk per thread: the practical HBM
// ceiling (read + write) that the step kernels are measured against.
__global__ __launch_bounds__(kBlock) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// The two-step kernel's access pattern alone: block b sweeps rows [b R, b R +
// R) of the row-interleaved layout (nine planes of pitch FP32 values per row),
// loading the next row before it writes the current one, a barrier per row,
// one multiply-add per value; blocks in XCD order.  NS float2 column slots per
// lane (pitch / 2 <= NS blockDim.x).  kPull: plane i of output row x is read
// from row x - cx_i (periodic), as the pull reads it; else from row x.  Rows
// live at their slots of layout L.
template <int NS, bool kPull>
__global__ __launch_bounds__(1024) void k_row_sweep_probe(const float2 *__restrict__ src, float2 *__restrict__ dst,
                                                          uint32_t lx, uint32_t pitch2, uint32_t rows, RowLayout L) {
    const uint32_t xa = xcd_block(blockIdx.x, gridDim.x) * rows;
    const uint32_t xe = min(xa + rows, lx);
    const uint32_t T = blockDim.x;
    uint32_t col[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) col[k] = min(threadIdx.x + k * T, pitch2 - 1);
    auto load = [&](uint32_t x, float2 out[NS][9]) {
        const size_t s0 = row_slot(L, x);
        const size_t sm = kPull ? row_slot(L, x == 0 ? lx - 1 : x - 1) : s0;
        const size_t sp = kPull ? row_slot(L, x + 1 == lx ? 0 : x + 1) : s0;
#pragma unroll
        for (int k = 0; k < NS; k++)
#pragma unroll
            for (int i = 0; i < 9; i++) {
                const size_t sr = kCX[i] == 1 ? sm : (kCX[i] == -1 ? sp : s0);
                out[k][i] = src[(sr * 9 + i) * pitch2 + col[k]];
            }
    };
    float2 cur[NS][9], nxt[NS][9];
    load(xa, cur);
    for (uint32_t x = xa; x < xe; x++) {
        if (x + 1 < xe) load(x + 1, nxt);
        // lanes past the last slot write the last one again (same values)
        const size_t sx = row_slot(L, x);
#pragma unroll
        for (int k = 0; k < NS; k++)
#pragma unroll
            for (int i = 0; i < 9; i++)
                dst[(sx * 9 + i) * pitch2 + col[k]] = make_float2(fmaf(cur[k][i].x, 1.0001f, 0.5f),
                                                                  fmaf(cur[k][i].y, 0.9999f, 0.25f));
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NS; k++)
#pragma unroll
            for (int i = 0; i < 9; i++) cur[k][i] = nxt[k][i];
    }
}

hipError_t launch_row_sweep_probe(const void *src, void *dst, uint32_t lx, uint32_t pitch, uint32_t rows,
                                  bool interleaved, bool pull, hipStream_t s) {
    if (lx == 0 || rows == 0 || pitch == 0 || pitch % 2 != 0 || pitch > 4096) return hipErrorInvalidValue;
    RowLayout L;
    if (interleaved) {
        L = row_layout(lx, rows);
        if (L.rows != rows) return hipErrorInvalidValue;
    }
    const uint32_t pitch2 = pitch / 2;
    const unsigned grid = (lx + rows - 1) / rows;
    const auto *in = static_cast<const float2 *>(src);
    auto *out = static_cast<float2 *>(dst);
    if (pitch2 <= 1024) {
        const uint32_t threads = (pitch2 + 63) / 64 * 64;
        if (pull)
            hipLaunchKernelGGL((k_row_sweep_probe<1, true>), dim3(grid), dim3(threads), 0, s, in, out, lx, pitch2, rows, L);
        else
            hipLaunchKernelGGL((k_row_sweep_probe<1, false>), dim3(grid), dim3(threads), 0, s, in, out, lx, pitch2, rows, L);
    } else {
        if (pull)
            hipLaunchKernelGGL((k_row_sweep_probe<2, true>), dim3(grid), dim3(1024), 0, s, in, out, lx, pitch2, rows, L);
        else
            hipLaunchKernelGGL((k_row_sweep_probe<2, false>), dim3(grid), dim3(1024), 0, s, in, out, lx, pitch2, rows, L);
    }
    return hipGetLastError();
}

hipError_t launch_copy16(const void *src, void *dst, int64_t n16, hipStream_t s) {
    if (n16 <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)((n16 + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n16);
    return hipGetLastError();
}

hipError_t launch_add_step(int64_t *step, int64_t k, hipStream_t s) {
    hipLaunchKernelGGL(k_add_step, dim3(1), dim3(1), 0, s, step, k);
    return hipGetLastError();
}

size_t sweep_lds_bytes(int storage, uint32_t ly, uint32_t nclasses, uint32_t rows) {
    return sweep_lds_for(storage, ly, nclasses, rows);
}

hipError_t launch_speed(const double *u, double *out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_speed, dim3(grid_for(n)), dim3(kBlock), 0, s, u, out, n);
    return hipGetLastError();
}

hipError_t launch_collide_first(int storage, const Geom &g, const double *f, const double *rho, const double *u,
                                const StateView &st, int dst, const BcView &bc, double omega, hipStream_t s) {
    const unsigned grid = grid_for(g.N);
    switch (storage) {
        case LBM_STORE_F64:
            hipLaunchKernelGGL(k_collide_first<double>, dim3(grid), dim3(kBlock), 0, s, f, rho, u,
                               static_cast<double *>(st.plane[dst]), static_cast<double *>(st.side[dst]), g, bc,
                               omega);
            break;
        case LBM_STORE_F32:
            hipLaunchKernelGGL(k_collide_first<float>, dim3(grid), dim3(kBlock), 0, s, f, rho, u,
                               static_cast<float *>(st.plane[dst]), static_cast<float *>(st.side[dst]), g, bc,
                               omega);
            break;
        case LBM_STORE_F16:
            hipLaunchKernelGGL(k_collide_first<__half>, dim3(grid), dim3(kBlock), 0, s, f, rho, u,
                               static_cast<__half *>(st.plane[dst]), static_cast<float *>(st.side[dst]), g, bc,
                               omega);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_finalize(int storage, const Geom &g, const StateView &st, int src, const BcView &bc, double *f,
                           double *rho, double *u, double *partials, hipStream_t s, int *nblocks) {
    const unsigned grid = grid_for(g.N);
    if (nblocks) *nblocks = (int)grid;
    switch (storage) {
        case LBM_STORE_F64:
            hipLaunchKernelGGL(k_finalize<double>, dim3(grid), dim3(kBlock), 0, s,
                               static_cast<const double *>(st.plane[src]), static_cast<const double *>(st.side[src]),
                               g, bc, f, rho, u, partials);
            break;
        case LBM_STORE_F32:
            hipLaunchKernelGGL(k_finalize<float>, dim3(grid), dim3(kBlock), 0, s,
                               static_cast<const float *>(st.plane[src]), static_cast<const float *>(st.side[src]), g,
                               bc, f, rho, u, partials);
            break;
        case LBM_STORE_F16:
            hipLaunchKernelGGL(k_finalize<__half>, dim3(grid), dim3(kBlock), 0, s,
                               static_cast<const __half *>(st.plane[src]), static_cast<const float *>(st.side[src]),
                               g, bc, f, rho, u, partials);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_probe_cell(int storage, const Geom &g, const StateView &st, int src, const BcView &bc,
                             const ProbeView &pr, int64_t slot, hipStream_t s) {
    if (pr.cell < 0 || !pr.ring || pr.capacity <= 0) return hipSuccess;
    double *out = pr.ring + 3 * (slot % pr.capacity);
    const uint32_t n = (uint32_t)pr.cell;
    switch (storage) {
        case LBM_STORE_F64:
            hipLaunchKernelGGL(k_probe_cell<double>, dim3(1), dim3(1), 0, s,
                               static_cast<const double *>(st.plane[src]), static_cast<const double *>(st.side[src]),
                               g, bc, n, out);
            break;
        case LBM_STORE_F32:
            hipLaunchKernelGGL(k_probe_cell<float>, dim3(1), dim3(1), 0, s,
                               static_cast<const float *>(st.plane[src]), static_cast<const float *>(st.side[src]), g,
                               bc, n, out);
            break;
        case LBM_STORE_F16:
            hipLaunchKernelGGL(k_probe_cell<__half>, dim3(1), dim3(1), 0, s,
                               static_cast<const __half *>(st.plane[src]), static_cast<const float *>(st.side[src]),
                               g, bc, n, out);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sum_partials(const double *partials, int nblocks, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, partials, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_moments_sum_aos(const double *f, int64_t n, double *partials, hipStream_t s, int *nblocks) {
    const unsigned grid = grid_for(n);
    if (nblocks) *nblocks = (int)grid;
    hipLaunchKernelGGL(k_moments_sum_aos, dim3(grid), dim3(kBlock), 0, s, f, n, partials);
    return hipGetLastError();
}

hipError_t launch_stream_aos(const void *in, void *out, int nx, int ny, int dtype, hipStream_t s) {
    const unsigned grid = grid_for((int64_t)nx * ny);
    if (dtype == LBM_DT_F64)
        hipLaunchKernelGGL(k_stream_aos<double>, dim3(grid), dim3(kBlock), 0, s, static_cast<const double *>(in),
                           static_cast<double *>(out), nx, ny);
    else
        hipLaunchKernelGGL(k_stream_aos<float>, dim3(grid), dim3(kBlock), 0, s, static_cast<const float *>(in),
                           static_cast<float *>(out), nx, ny);
    return hipGetLastError();
}

hipError_t launch_equilibrium_aos(const void *rho, const void *u, void *feq, int64_t n, int dtype, hipStream_t s) {
    if (dtype == LBM_DT_F64)
        hipLaunchKernelGGL(k_equilibrium_aos<double>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                           static_cast<const double *>(rho), static_cast<const double *>(u),
                           static_cast<double *>(feq), n);
    else
        hipLaunchKernelGGL(k_equilibrium_aos<float>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                           static_cast<const float *>(rho), static_cast<const float *>(u), static_cast<float *>(feq),
                           n);
    return hipGetLastError();
}

hipError_t launch_density_aos(const void *f, void *rho, int64_t n, int dtype, hipStream_t s) {
    if (dtype == LBM_DT_F64)
        hipLaunchKernelGGL(k_density_aos<double>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                           static_cast<const double *>(f), static_cast<double *>(rho), n);
    else
        hipLaunchKernelGGL(k_density_aos<float>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                           static_cast<const float *>(f), static_cast<float *>(rho), n);
    return hipGetLastError();
}

hipError_t launch_velocity_aos(const void *rho, const void *f, void *u, int64_t n, int dtype, hipStream_t s) {
    if (dtype == LBM_DT_F64)
        hipLaunchKernelGGL(k_velocity_aos<double>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                           static_cast<const double *>(rho), static_cast<const double *>(f),
                           static_cast<double *>(u), n);
    else
        hipLaunchKernelGGL(k_velocity_aos<float>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                           static_cast<const float *>(rho), static_cast<const float *>(f), static_cast<float *>(u),
                           n);
    return hipGetLastError();
}

hipError_t launch_collide_aos(const void *f, const void *rho, const void *u, double omega, void *out, int64_t n,
                              int dtype, hipStream_t s) {
    if (dtype == LBM_DT_F64)
        hipLaunchKernelGGL(k_collide_aos<double>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                           static_cast<const double *>(f), static_cast<const double *>(rho),
                           static_cast<const double *>(u), omega, static_cast<double *>(out), n);
    else
        hipLaunchKernelGGL(k_collide_aos<float>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                           static_cast<const float *>(f), static_cast<const float *>(rho),
                           static_cast<const float *>(u), (float)omega, static_cast<float *>(out), n);
    return hipGetLastError();
}

hipError_t launch_bc_apply(const Geom &g, const BcView &bc, const double *f_pre, double *f_post, const double *f_prev,
                           hipStream_t s) {
    if (bc.nspecial == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bc_apply, dim3(grid_for(bc.nspecial)), dim3(kBlock), 0, s, g, bc, f_pre, f_post, f_prev);
    return hipGetLastError();
}

hipError_t launch_bc_pressure(const Geom &g, const BcView &bc, double *f_pre, const double *rho, const double *u,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_bc_pressure, dim3(grid_for(g.ly)), dim3(kBlock), 0, s, g, bc, f_pre, rho, u);
    return hipGetLastError();
}

}  // namespace lbm
