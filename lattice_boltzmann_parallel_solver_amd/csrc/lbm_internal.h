// Internal declarations shared by the boundary compiler (lbm_bc.cpp, host
// C++) and the HIP runtime (lbm_capi.hip).  Not part of the public ABI.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/lbm_hip.h"

namespace lbm {

// D2Q9 numbering of the reference (lattice_boltzmann_method.py:5-52):
// 0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE.
constexpr int kCX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int kCY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int kOpp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
constexpr double kW[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                          1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};

// Per-direction source of a post-streaming population (2 bits each).
enum SrcKind : uint32_t {
    kSrcStream = 0,   // pulled from the upstream neighbour (no boundary rule)
    kSrcReflect = 1,  // same node, opposite direction, minus k[d]
    kSrcConst = 2,    // the constant k[d]
    kSrcOutlet = 3,   // column nx-2 of the step-input populations
};

// Per-class flags.
enum ClassFlag : uint32_t {
    kFlagOutletSide = 1u << 0,    // node at x = nx-2 feeds the outlet buffer
    kFlagPressureIn = 1u << 1,    // node at x = nx-2 produces the x = 0 pressure fix
    kFlagPressureOut = 1u << 2,   // node at x = 1 produces the x = nx-1 pressure fix
    kSkipShift = 8,               // bits 8..16: directions not stored by this node
};

// Program-level flags (lbm_bc_query).
enum ProgFlag : uint32_t { kProgOutlet = 1u << 0, kProgPressure = 1u << 1 };

struct ClassEntry {  // device layout, 80 bytes
    uint32_t src;
    uint32_t flags;
    double k[9];
};

// Outlet buffer slot of directions 3, 6, 7 (the ones boundary_conditions.py:265 copies).
constexpr int outlet_slot(int d) { return d == 3 ? 0 : (d == 6 ? 1 : 2); }

}  // namespace lbm

struct lbm_bc_prog {
    int nx = 0, ny = 0;
    uint32_t flags = 0;
    std::vector<uint16_t> codes;            // nx*ny
    std::vector<lbm::ClassEntry> classes;   // classes[0] = plain fluid
    std::vector<uint8_t> chunk_flags;       // one per 64 linear nodes
    std::vector<uint8_t> wave_flags;        // one per (row, 64-column wave), x * ceil(ny/64) + y/64
    std::vector<uint32_t> special;          // nodes with a nonzero class
    double rho_in = 0, rho_out = 0;         // pressure fix densities
    uint32_t d1 = 0, d2 = 0;                // pressure fix direction sets
    // device copy used by the stateless lbm_bc_apply / lbm_bc_pressure
    // (uploaded on first use, released by lbm_bc_destroy)
    int dev_id = -1;
    void *dev_mem = nullptr;
    const uint16_t *dev_codes = nullptr;
    const uint8_t *dev_chunk = nullptr;
    const uint8_t *dev_wave = nullptr;
    const lbm::ClassEntry *dev_classes = nullptr;
    const uint32_t *dev_special = nullptr;
};

namespace lbm {
void set_error(const std::string &msg);
}
