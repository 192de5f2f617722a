// Boundary compiler: ordered boundary descriptors -> per-node classes.
//
// The reference applies its boundary closures one after another on whole
// arrays every step (boundary_utils.py:35-53, 87-108, 164-203).  Every rule
// writes f_post[node, d] from values that no rule modifies (the
// post-collision populations of the same node, a constant, or the step-input
// populations), so the result of a composite is fixed per (node, direction) by
// the LAST rule that writes it.  This file evaluates that "last writer" once,
// on the host, and hands the kernels a small class table plus a 16-bit class
// code per node, so the fused step applies every boundary in the same pass
// as streaming and collision.
//
// Host-only C++, built with -ffp-contract=off so the constants match the
// reference's float64 expressions bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "lbm_internal.h"

using namespace lbm;

namespace {

struct Slot {
    uint32_t kind = kSrcStream;
    double k = 0.0;
};

struct NodeRules {
    Slot slot[9];
    uint32_t flags = 0;
};

// Reference speed of sound squared: (1/np.sqrt(3))**2 (boundary_conditions.py:186, :304).
double cs2() {
    const double cs = 1.0 / std::sqrt(3.0);
    return cs * cs;
}

// equilibrium_distr_func for one node (lattice_boltzmann_method.py:162-188),
// same operation order as numpy.
void feq_node(double rho, double ux, double uy, double out[9]) {
    const double mag = std::sqrt(ux * ux + uy * uy);
    const double usq = mag * mag;
    for (int i = 0; i < 9; i++) {
        const double cu = ux * (double)kCX[i] + uy * (double)kCY[i];
        out[i] = (kW[i] * rho) * (((1.0 + 3.0 * cu) + 4.5 * (cu * cu)) - 1.5 * usq);
    }
}

struct Builder {
    int nx, ny;
    std::unordered_map<uint32_t, NodeRules> nodes;
    bool outlet = false, pressure = false;
    double rho_in = 0, rho_out = 0;
    uint32_t d1 = 0, d2 = 0;

    NodeRules &at(int x, int y) { return nodes[(uint32_t)x * (uint32_t)ny + (uint32_t)y]; }

    bool in_mask(const uint8_t *m, int x, int y) const { return m[(size_t)x * ny + y] != 0; }

    // get_wall_indices (boundary_conditions.py:8-28): first full edge wins.
    int wall_dirs(const uint8_t *m, int dirs[3]) const {
        auto full_col = [&](int x) {
            for (int y = 0; y < ny; y++)
                if (!in_mask(m, x, y)) return false;
            return true;
        };
        auto full_row = [&](int y) {
            for (int x = 0; x < nx; x++)
                if (!in_mask(m, x, y)) return false;
            return true;
        };
        static const int sets[4][3] = {{1, 5, 8}, {3, 6, 7}, {4, 7, 8}, {2, 5, 6}};
        int which = -1;
        if (full_col(0)) which = 0;
        else if (full_col(nx - 1)) which = 1;
        else if (full_row(0)) which = 2;
        else if (full_row(ny - 1)) which = 3;
        if (which < 0) return 0;
        for (int j = 0; j < 3; j++) dirs[j] = sets[which][j];
        return 3;
    }

    void reflect(int x, int y, int i, double k) {
        Slot &s = at(x, y).slot[kOpp[i]];
        s.kind = kSrcReflect;
        s.k = k;
    }

    std::string rigid_or_moving(const lbm_bc_desc &d, bool moving) {
        if (!d.mask) return "wall descriptor needs a mask";
        int dirs[3];
        int nd = wall_dirs(d.mask, dirs);
        double k[9] = {0};
        if (moving) {
            // ((2*w_i)*rho_w) * ((c_i . u_w)/cs^2)   (boundary_conditions.py:208-210)
            const double c2 = cs2();
            for (int j = 0; j < nd; j++) {
                int i = dirs[j];
                const double dot = (double)kCX[i] * d.u_w[0] + (double)kCY[i] * d.u_w[1];
                k[i] = ((2.0 * kW[i]) * d.rho_w) * (dot / c2);
            }
        }
        for (int x = 0; x < nx; x++)
            for (int y = 0; y < ny; y++)
                if (in_mask(d.mask, x, y))
                    for (int j = 0; j < nd; j++) reflect(x, y, dirs[j], k[dirs[j]]);
        return "";
    }

    std::string rigid_object(const lbm_bc_desc &d) {
        if (!d.mask) return "rigid object descriptor needs a mask";
        // get_corner_indices (boundary_conditions.py:31-56)
        int x0 = nx, x1 = -1, y0 = ny, y1 = -1;
        for (int x = 0; x < nx; x++)
            for (int y = 0; y < ny; y++)
                if (in_mask(d.mask, x, y)) {
                    x0 = std::min(x0, x);
                    x1 = std::max(x1, x);
                    y0 = std::min(y0, y);
                    y1 = std::max(y1, y);
                }
        if (x1 < 0) return "rigid object mask is empty";
        if (x1 + 1 >= nx) return "rigid object touches the last column (its right face would leave the lattice)";
        struct Corner {
            int x, y, a, b;
        } corners[4] = {{x0, y1, 1, 8}, {x0, y0, 1, 5}, {x1 + 1, y1, 3, 7}, {x1 + 1, y0, 3, 6}};
        // remove_corner_indices_from_boundary (:59-75), on a private copy
        std::vector<uint8_t> m(d.mask, d.mask + (size_t)nx * ny);
        for (auto &c : corners) m[(size_t)c.x * ny + c.y] = 0;
        // left faces on the mask, right faces on the mask rolled +1 in x (:133-156)
        for (int x = 0; x < nx; x++)
            for (int y = 0; y < ny; y++)
                if (m[(size_t)x * ny + y])
                    for (int i : {1, 5, 8}) reflect(x, y, i, 0.0);
        for (int x = 0; x < nx; x++)
            for (int y = 0; y < ny; y++) {
                int xs = (x - 1 + nx) % nx;
                if (m[(size_t)xs * ny + y])
                    for (int i : {3, 6, 7}) reflect(x, y, i, 0.0);
            }
        for (auto &c : corners) {  // (:158-163)
            reflect(c.x, c.y, c.a, 0.0);
            reflect(c.x, c.y, c.b, 0.0);
        }
        return "";
    }

    std::string inlet(const lbm_bc_desc &d) {
        double e[9];
        feq_node(d.rho_in, d.u_in, 0.0, e);  // (boundary_conditions.py:232-237)
        for (int y = 0; y < ny; y++)
            for (int i = 0; i < 9; i++) {
                Slot &s = at(0, y).slot[i];
                s.kind = kSrcConst;
                s.k = e[i];
            }
        return "";
    }

    std::string outlet_bc() {
        if (nx < 2) return "outlet needs nx >= 2";
        outlet = true;
        for (int y = 0; y < ny; y++) {
            for (int i : {3, 6, 7}) {
                Slot &s = at(nx - 1, y).slot[i];
                s.kind = kSrcOutlet;
                s.k = 0.0;
            }
            at(nx - 2, y).flags |= kFlagOutletSide;
        }
        return "";
    }

    std::string pressure_bc(const lbm_bc_desc &d) {
        if (!d.mask) return "pressure descriptor needs a mask";
        if (pressure) return "only one pressure-periodic descriptor per program";
        if (nx < 2) return "pressure-periodic needs nx >= 2";
        // orientation test (boundary_conditions.py:301-318)
        bool xeq = true, yeq = true;
        for (int y = 0; y < ny; y++)
            if ((d.mask[y] != 0) != (d.mask[(size_t)(nx - 1) * ny + y] != 0)) xeq = false;
        for (int x = 0; x < nx; x++)
            if ((d.mask[(size_t)x * ny] != 0) != (d.mask[(size_t)x * ny + ny - 1] != 0)) yeq = false;
        if (xeq) {
            d1 = (1u << 1) | (1u << 5) | (1u << 8);
            d2 = (1u << 3) | (1u << 6) | (1u << 7);
        } else if (yeq) {
            // the reference's y branch builds density arrays of length nx and
            // pairs them with velocity columns of length ny (:312-316, :338)
            if (nx != ny) return "pressure-periodic along y needs nx == ny (as in the reference)";
            d1 = (1u << 2) | (1u << 5) | (1u << 6);
            d2 = (1u << 4) | (1u << 7) | (1u << 8);
        } else {
            return "pressure-periodic mask must have equal first/last columns or rows";
        }
        const double c2 = cs2();
        rho_in = d.p_in / c2;
        rho_out = d.p_out / c2;
        pressure = true;
        for (int y = 0; y < ny; y++) {
            at(nx - 2, y).flags |= kFlagPressureIn;
            at(1 % nx, y).flags |= kFlagPressureOut;
            at(0, y).flags |= d1 << kSkipShift;
            at(nx - 1, y).flags |= d2 << kSkipShift;
        }
        return "";
    }
};

}  // namespace

extern "C" int lbm_bc_compile(lbm_bc_prog **out, int nx, int ny, const lbm_bc_desc *descs, int n) {
    if (!out) {
        set_error("lbm_bc_compile: out is NULL");
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    if (nx < 1 || ny < 1 || (int64_t)nx * ny > 0x7fffffffLL) {
        set_error("lbm_bc_compile: lattice must satisfy 1 <= nx*ny < 2^31");
        return LBM_ERR_ARG;
    }
    if (n < 0 || (n > 0 && !descs)) {
        set_error("lbm_bc_compile: bad descriptor list");
        return LBM_ERR_ARG;
    }
    Builder b{nx, ny, {}, false, false, 0, 0, 0, 0};
    for (int j = 0; j < n; j++) {
        const lbm_bc_desc &d = descs[j];
        std::string err;
        switch (d.kind) {
            case LBM_BC_RIGID_WALL: err = b.rigid_or_moving(d, false); break;
            case LBM_BC_MOVING_WALL: err = b.rigid_or_moving(d, true); break;
            case LBM_BC_RIGID_OBJECT: err = b.rigid_object(d); break;
            case LBM_BC_INLET: err = b.inlet(d); break;
            case LBM_BC_OUTLET: err = b.outlet_bc(); break;
            case LBM_BC_PRESSURE_PERIODIC: err = b.pressure_bc(d); break;
            default: err = "unknown descriptor kind " + std::to_string(d.kind);
        }
        if (!err.empty()) {
            set_error("lbm_bc_compile: descriptor " + std::to_string(j) + ": " + err);
            return LBM_ERR_ARG;
        }
    }

    auto *p = new (std::nothrow) lbm_bc_prog();
    if (!p) {
        set_error("lbm_bc_compile: out of host memory");
        return LBM_ERR_NOMEM;
    }
    const size_t N = (size_t)nx * ny;
    p->nx = nx;
    p->ny = ny;
    p->codes.assign(N, 0);
    // one flag per 64 nodes, padded to a multiple of 8 so a wave of the
    // 16-byte kernel (up to 8 chunks) reads its flags with one aligned load
    // one byte per 64 nodes, padded to a multiple of 8 plus 8 spare zero bytes:
    // a wave whose lanes all lie past the last node may still read the flag
    // of the chunk after the last one
    p->chunk_flags.assign((((N + 63) / 64 + 7) & ~size_t(7)) + 8, 0);
    // one byte per row and 64-column wave (the unit of the row-aligned
    // kernels), plus 8 spare zero bytes
    const size_t wpr = ((size_t)ny + 63) / 64;
    p->wave_flags.assign((size_t)nx * wpr + 8, 0);
    p->classes.push_back(ClassEntry{0, 0, {0}});
    p->rho_in = b.rho_in;
    p->rho_out = b.rho_out;
    p->d1 = b.d1;
    p->d2 = b.d2;
    p->flags = (b.outlet ? kProgOutlet : 0) | (b.pressure ? kProgPressure : 0);

    // deduplicate node rule sets into classes; nodes are visited in index
    // order so the class numbering is deterministic
    std::vector<uint32_t> keys;
    keys.reserve(b.nodes.size());
    for (auto &kv : b.nodes) keys.push_back(kv.first);
    std::sort(keys.begin(), keys.end());
    std::map<std::vector<uint64_t>, uint16_t> index;
    for (uint32_t node : keys) {
        const NodeRules &r = b.nodes[node];
        ClassEntry c{0, r.flags, {0}};
        for (int d = 0; d < 9; d++) {
            c.src |= (r.slot[d].kind & 3u) << (2 * d);
            c.k[d] = r.slot[d].kind == kSrcStream ? 0.0 : r.slot[d].k;
        }
        if (c.src == 0 && c.flags == 0) continue;
        std::vector<uint64_t> key(11);
        key[0] = c.src;
        key[1] = c.flags;
        for (int d = 0; d < 9; d++) std::memcpy(&key[2 + d], &c.k[d], 8);
        auto it = index.find(key);
        uint16_t code;
        if (it == index.end()) {
            if (p->classes.size() >= 65535) {
                delete p;
                set_error("lbm_bc_compile: more than 65534 distinct boundary node classes");
                return LBM_ERR_ARG;
            }
            code = (uint16_t)p->classes.size();
            p->classes.push_back(c);
            index.emplace(std::move(key), code);
        } else {
            code = it->second;
        }
        p->codes[node] = code;
        p->chunk_flags[node / 64] = 1;
        p->wave_flags[(size_t)(node / (uint32_t)ny) * wpr + (node % (uint32_t)ny) / 64] = 1;
        p->special.push_back(node);
    }
    *out = p;
    return LBM_OK;
}

extern "C" int lbm_bc_query(const lbm_bc_prog *prog, int *nclasses, int64_t *nspecial, uint32_t *flags) {
    if (!prog) {
        set_error("lbm_bc_query: prog is NULL");
        return LBM_ERR_ARG;
    }
    if (nclasses) *nclasses = (int)prog->classes.size();
    if (nspecial) *nspecial = (int64_t)prog->special.size();
    if (flags) *flags = prog->flags;
    return LBM_OK;
}

extern "C" int lbm_bc_export(const lbm_bc_prog *prog, uint16_t *codes, uint32_t *class_src,
                             uint32_t *class_flags, double *class_k, double *params) {
    if (!prog) {
        set_error("lbm_bc_export: prog is NULL");
        return LBM_ERR_ARG;
    }
    if (codes) std::memcpy(codes, prog->codes.data(), prog->codes.size() * sizeof(uint16_t));
    for (size_t c = 0; c < prog->classes.size(); c++) {
        if (class_src) class_src[c] = prog->classes[c].src;
        if (class_flags) class_flags[c] = prog->classes[c].flags;
        if (class_k)
            for (int d = 0; d < 9; d++) class_k[9 * c + d] = prog->classes[c].k[d];
    }
    if (params) {
        params[0] = prog->rho_in;
        params[1] = prog->rho_out;
        params[2] = (double)prog->d1;
        params[3] = (double)prog->d2;
    }
    return LBM_OK;
}
