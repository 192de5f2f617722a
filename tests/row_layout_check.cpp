// Host check of the range-interleaved row order (lbm_rowlayout.h): for every
// lattice length up to 2100 and every row count up to 128 the slots are a
// permutation of the rows, the rows of one offset fill consecutive slots in
// range order, and the multiply-shift division equals x / R up to 2^24.
#include <cstdio>
#include <vector>

#include "lbm_rowlayout.h"

int main() {
    long bad = 0, cases = 0;
    for (uint32_t lx = 1; lx <= 2100; lx++) {
        for (uint32_t R = 1; R <= lbm::kRowLayoutMaxRows; R++) {
            const lbm::RowLayout L = lbm::row_layout(lx, R);
            std::vector<int> seen(lx, 0);
            uint32_t next = 0;  // slots visited offset by offset, range by range, must count up
            for (uint32_t t = 0; t < R; t++)
                for (uint32_t b = 0; b * R + t < lx; b++) {
                    const uint32_t x = b * R + t;
                    const uint32_t s = lbm::row_slot(L, x);
                    if (s >= lx || seen[s]++ || s != next) bad++;
                    next++;
                }
            cases++;
        }
    }
    for (uint32_t R = 1; R <= lbm::kRowLayoutMaxRows; R++) {
        const lbm::RowLayout L = lbm::row_layout(lbm::kRowLayoutMaxLx - 1, R);
        for (uint64_t x = 0; x < lbm::kRowLayoutMaxLx; x += 61)
            if (((x * L.magic) >> 32) != x / R) bad++;
    }
    if (lbm::row_layout(100, 0).rows != 0 || lbm::row_slot(lbm::row_layout(100, 0), 57) != 57) bad++;
    std::printf("%ld cases, %ld failures\n", cases, bad);
    return bad != 0;
}
