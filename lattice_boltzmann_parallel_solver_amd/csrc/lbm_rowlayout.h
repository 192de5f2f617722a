// Range-interleaved row order of a state buffer (host and device).
//
// The two-step sweep splits the lx rows into ranges of R rows (the last one
// possibly shorter) and sweeps all ranges at once, one row per range at a
// time.  In the plain order (slot = row) the rows that all blocks touch at a
// given moment lie one range apart in memory, R x 9 planes apart; stored
// offset-major instead -- offset t of every range in consecutive slots --
// they are neighbours, and the same row sweep moves 5.7-5.8 TB/s on MI355X
// instead of 5.0-5.1 (8192 rows of 2048 FP32 values, 32-36 rows per range).
//
// Slots: the first `whole` offsets (lx mod R, or R when R divides lx) exist
// in all `ranges` ranges, the others only in the ranges - 1 full ones, so the
// lx rows fill exactly lx slots.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

struct RowLayout {
    uint32_t rows = 0;    // R, rows per range; 0: the plain order (slot = row)
    uint32_t ranges = 0;  // ceil(lx / R)
    uint32_t whole = 0;   // offsets that every range has
    uint64_t magic = 0;   // ceil(2^32 / R): x / R = (x * magic) >> 32 for x < 2^25 and R <= 128
};

// Largest rows per range and lattice length the division by `magic` is exact for.
constexpr uint32_t kRowLayoutMaxRows = 128;
constexpr uint32_t kRowLayoutMaxLx = 1u << 24;

__host__ __device__ inline RowLayout row_layout(uint32_t lx, uint32_t rows) {
    RowLayout L;
    if (rows == 0 || rows > kRowLayoutMaxRows || lx == 0 || lx >= kRowLayoutMaxLx) return L;
    L.rows = rows;
    L.ranges = (lx + rows - 1) / rows;
    L.whole = lx % rows == 0 ? rows : lx % rows;
    L.magic = ((1ull << 32) + rows - 1) / rows;
    return L;
}

// Memory slot of row x (x < lx).
__host__ __device__ inline uint32_t row_slot(const RowLayout &L, uint32_t x) {
    if (L.rows == 0) return x;
    const uint32_t b = (uint32_t)(((uint64_t)x * L.magic) >> 32);  // range
    const uint32_t t = x - b * L.rows;                              // offset in the range
    return t < L.whole ? t * L.ranges + b : L.whole * L.ranges + (t - L.whole) * (L.ranges - 1) + b;
}

}  // namespace lbm
