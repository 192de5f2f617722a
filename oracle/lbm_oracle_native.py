"""ctypes binding of the C oracle (oracle/lbm_oracle_c.c).

TEST INFRASTRUCTURE ONLY: the checker used by tests/ for long runs and by
bench.py's cpu_baseline leg.  It runs the same arithmetic as
oracle/lbm_oracle.py; tests/test_oracle.py checks the two bit for bit.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import lbm_oracle as O

_HERE = os.path.dirname(os.path.abspath(__file__))
_SRC = os.path.join(_HERE, "lbm_oracle_c.c")
_LIB = os.path.join(_HERE, "liblbm_oracle.so")

OP_REFLECT, OP_INLET, OP_OUTLET, OP_PRESSURE = 1, 2, 3, 4


class _Op(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("dirs", ctypes.c_int), ("dirs2", ctypes.c_int),
                ("ncells", ctypes.c_long), ("cells", ctypes.POINTER(ctypes.c_int64)),
                ("k", ctypes.c_double * 9), ("rho_in", ctypes.c_double),
                ("rho_out", ctypes.c_double)]


def build(force: bool = False) -> str:
    """Compile the C oracle with gcc (no FMA contraction, no fast-math)."""
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(_SRC):
        cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
               "-o", _LIB, _SRC, "-lm"]
        subprocess.check_call(cmd)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.POINTER(ctypes.c_double)
        L.oracle_run.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_double, ctypes.c_long,
                                 ctypes.POINTER(_Op), ctypes.c_int, ctypes.c_long, P]
        L.oracle_run.restype = ctypes.c_int
        L.oracle_set_square_pow.argtypes = [ctypes.c_int]
        L.oracle_equilibrium.argtypes = [P, P, P, ctypes.c_long]
        L.oracle_norm_fma.argtypes = [P, P, P, ctypes.c_long]
        L.oracle_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ops_from(boundary, ly):
    """Translate a numpy-oracle composite's oracle_ops into C op records."""
    keep = []
    ops = []
    if boundary is None:
        return (_Op * 1)(), 0, keep
    for kind, payload in boundary.oracle_ops:
        if kind == "reflect":
            for mask, dirs, k in payload:
                idx = np.ascontiguousarray(np.flatnonzero(mask.ravel()).astype(np.int64))
                keep.append(idx)
                op = _Op()
                op.kind = OP_REFLECT
                op.dirs = sum(1 << i for i in dirs)
                op.ncells = idx.size
                op.cells = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
                for i in range(9):
                    op.k[i] = k[i]
                ops.append(op)
        elif kind == "inlet":
            op = _Op()
            op.kind = OP_INLET
            for i in range(9):
                op.k[i] = payload[i]
            ops.append(op)
        elif kind == "outlet":
            op = _Op()
            op.kind = OP_OUTLET
            ops.append(op)
        elif kind == "pressure":
            rin, rout, d1, d2 = payload
            op = _Op()
            op.kind = OP_PRESSURE
            op.dirs = sum(1 << i for i in d1)
            op.dirs2 = sum(1 << i for i in d2)
            op.rho_in = rin
            op.rho_out = rout
            ops.append(op)
        else:
            raise ValueError(kind)
    arr = (_Op * max(1, len(ops)))(*ops)
    return arr, len(ops), keep


def run(f, rho, u, omega, steps, boundary=None, probe=None, square=None):
    """Same contract as lbm_oracle.run but in C; returns (f, rho, u, trace[steps,3]).

    trace rows are (rho, ux, uy) at the probe cell after each step.
    """
    L = lib()
    L.oracle_set_square_pow(1 if (square or O.SQUARE) == "pow" else 0)
    lx, ly = f.shape[0], f.shape[1]
    f = np.ascontiguousarray(f, dtype=np.float64).copy()
    rho = np.ascontiguousarray(rho, dtype=np.float64).copy()
    u = np.ascontiguousarray(u, dtype=np.float64).copy()
    ops, nops, keep = _ops_from(boundary, ly)
    pidx = -1 if probe is None else int(probe[0]) * ly + int(probe[1])
    trace = np.zeros((max(1, steps), 3))
    rc = L.oracle_run(lx, ly, _dptr(f), _dptr(rho), _dptr(u), float(omega), int(steps), ops, nops,
                      pidx, _dptr(trace))
    if rc != 0:
        raise MemoryError("oracle_run failed")
    del keep
    return f, rho, u, trace[:steps]


def norm_fma(ux, uy):
    """sqrt(fma(uy, uy, ux*ux)) elementwise: numpy's 1-D norm through an FMA BLAS dot."""
    ux = np.ascontiguousarray(ux, dtype=np.float64)
    uy = np.ascontiguousarray(uy, dtype=np.float64)
    out = np.empty_like(ux)
    lib().oracle_norm_fma(_dptr(ux), _dptr(uy), _dptr(out), ux.size)
    return out


def norm_plain(ux, uy):
    """sqrt(ux*ux + uy*uy) elementwise: the norm without a fused multiply-add."""
    ux = np.asarray(ux, dtype=np.float64)
    uy = np.asarray(uy, dtype=np.float64)
    return np.sqrt(ux * ux + uy * uy)
