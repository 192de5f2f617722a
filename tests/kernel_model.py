"""Numpy model of the device dataflow (test helper, CPU only).

It runs the fused-step algorithm of lbm_kernels.hip on the tables that the
boundary compiler exports (BoundaryProgram.export()), using the oracle's
arithmetic: state = post-collision populations after the pressure fix,
pull streaming, per-node classes (stream / reflect / constant / outlet copy),
moments, equilibrium, BGK, stores with the skip mask, the pressure-fix
producers at x = lx-2 and x = 1 and the outlet side buffer at x = lx-2.
Comparing it with the oracle's reference-order step checks the compiler's
tables and the phase-shifted formulation without a GPU.
"""

import numpy as np

from oracle import lbm_oracle as O

SRC_STREAM, SRC_REFLECT, SRC_CONST, SRC_OUTLET = 0, 1, 2, 3
FLAG_OUTLET_SIDE, FLAG_P_IN, FLAG_P_OUT = 1, 2, 4
SLOT = {3: 0, 6: 1, 7: 2}


class DeviceModel:
    def __init__(self, lx, ly, tables=None):
        self.lx, self.ly = lx, ly
        if tables is None:
            tables = {"codes": np.zeros((lx, ly), np.uint16), "src": np.zeros(1, np.uint32),
                      "flags": np.zeros(1, np.uint32), "k": np.zeros((1, 9)), "params": np.zeros(4)}
        self.t = tables
        self.codes = tables["codes"]
        self.flags = tables["flags"][self.codes]  # per node
        self.rho_in, self.rho_out = tables["params"][0], tables["params"][1]
        self.d1, self.d2 = int(tables["params"][2]), int(tables["params"][3])

    def _store(self, s_new, feq, u):
        """Skip mask + pressure producers, as store_node()."""
        out = s_new.copy()
        fl = self.flags
        for x, y in zip(*np.nonzero(fl & FLAG_P_IN)):
            e = O.equilibrium(np.array(self.rho_in), u[x, y])
            for i in range(9):
                if self.d1 >> i & 1:
                    out[0, y, i] = e[i] + (s_new[x, y, i] - feq[x, y, i])
        for x, y in zip(*np.nonzero(fl & FLAG_P_OUT)):
            e = O.equilibrium(np.array(self.rho_out), u[x, y])
            for i in range(9):
                if self.d2 >> i & 1:
                    out[self.lx - 1, y, i] = e[i] + (s_new[x, y, i] - feq[x, y, i])
        return out

    def _side(self, fvals):
        side = np.zeros((self.ly, 3))
        for x, y in zip(*np.nonzero(self.flags & FLAG_OUTLET_SIDE)):
            side[y] = fvals[x, y, [3, 6, 7]]
        return side

    def collide_first(self, f, rho, u, omega):
        feq = O.equilibrium(rho, u)
        s = O.collide(f, feq, omega)
        return self._store(s, feq, u), self._side(f)

    def pull_bc(self, s, side_prev):
        g = O.streaming(s)
        src, kk = self.t["src"], self.t["k"]
        for x, y in zip(*np.nonzero(self.codes)):
            c = self.codes[x, y]
            for d in range(9):
                kind = (int(src[c]) >> (2 * d)) & 3
                if kind == SRC_REFLECT:
                    g[x, y, d] = s[x, y, O.OPP[d]] - kk[c, d]
                elif kind == SRC_CONST:
                    g[x, y, d] = kk[c, d]
                elif kind == SRC_OUTLET:
                    g[x, y, d] = side_prev[y, SLOT[d]]
        return g

    def fused(self, s, side_prev, omega):
        g = self.pull_bc(s, side_prev)
        rho = O.density(g)
        u = O.velocity(rho, g)
        feq = O.equilibrium(rho, u)
        s_new = O.collide(g, feq, omega)
        return self._store(s_new, feq, u), self._side(g)

    def finalize(self, s, side_prev):
        g = self.pull_bc(s, side_prev)
        rho = O.density(g)
        return g, rho, O.velocity(rho, g)

    def run(self, f, rho, u, omega, steps):
        s, side = self.collide_first(f, rho, u, omega)
        for _ in range(steps - 1):
            s, side = self.fused(s, side, omega)
        return self.finalize(s, side)
