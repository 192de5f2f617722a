"""ctypes binding of liblbm_hip.so (the C ABI declared in include/lbm_hip.h).

The library is built in-tree by ``python -m lattice_boltzmann_parallel_solver_amd.build``.
There is no CPU fallback: if the library is missing, or no HIP device is
visible when a compute entry point is used, the call raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm_hip.so")

# status codes (include/lbm_hip.h)
LBM_OK = 0
LBM_ERR_ARG = -1
LBM_ERR_HIP = -2
LBM_ERR_STATE = -3
LBM_ERR_NOMEM = -4
LBM_ERR_UNSTABLE = -5

STORE_F64, STORE_F32, STORE_F16 = 0, 1, 2
STORAGE = {"f64": STORE_F64, "f32": STORE_F32, "f16": STORE_F16}
DT_F64, DT_F32 = 0, 1
HOST, DEVICE = 0, 1

BC_RIGID_WALL = 1
BC_MOVING_WALL = 2
BC_RIGID_OBJECT = 3
BC_INLET = 4
BC_OUTLET = 5
BC_PRESSURE_PERIODIC = 6


class LbmError(RuntimeError):
    """A C ABI call returned an error code."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class LbmUnstable(LbmError):
    """Non-finite moments: the run blew up."""


class BcDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("mask", ctypes.POINTER(ctypes.c_uint8)),
        ("u_w", ctypes.c_double * 2),
        ("rho_w", ctypes.c_double),
        ("rho_in", ctypes.c_double),
        ("u_in", ctypes.c_double),
        ("p_in", ctypes.c_double),
        ("p_out", ctypes.c_double),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes)
_SIGNATURES = {
    "lbm_last_error": (ctypes.c_char_p, []),
    "lbm_version": (_I, []),
    "lbm_device_count": (_I, [ctypes.POINTER(_I)]),
    "lbm_bc_compile": (_I, [ctypes.POINTER(_P), _I, _I, ctypes.POINTER(BcDesc), _I]),
    "lbm_bc_destroy": (_I, [_P]),
    "lbm_bc_query": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I64), ctypes.POINTER(ctypes.c_uint32)]),
    "lbm_bc_export": (_I, [_P, _P, _P, _P, _P, _P]),
    "lbm_create": (_I, [ctypes.POINTER(_P), _I, _I, _I, _I]),
    "lbm_destroy": (_I, [_P]),
    "lbm_set_stream": (_I, [_P, _P]),
    "lbm_set_bc": (_I, [_P, _P]),
    "lbm_set_state": (_I, [_P, _P, _P, _P, _I]),
    "lbm_set_state_equilibrium": (_I, [_P, _P, _P, _I]),
    "lbm_run": (_I, [_P, _D, _I64]),
    "lbm_get_state": (_I, [_P, _P, _P, _P, _I]),
    "lbm_set_probe": (_I, [_P, _I, _I, _I64]),
    "lbm_read_probe": (_I, [_P, _P, _I64, ctypes.POINTER(_I64)]),
    "lbm_moments_sum": (_I, [_P, _PD]),
    "lbm_synchronize": (_I, [_P]),
    "lbm_speed": (_I, [_P, _P, _I]),
    "lbm_checkpoint_bytes": (_I, [_P, ctypes.POINTER(_I64)]),
    "lbm_checkpoint_save": (_I, [_P, _P, _I64]),
    "lbm_checkpoint_load": (_I, [_P, _P, _I64]),
    "lbm_steps_done": (_I, [_P, ctypes.POINTER(_I64)]),
    "lbm_set_launch": (_I, [_P, _I, _I, _I]),
    "lbm_set_temporal": (_I, [_P, _I, _I]),
    "lbm_set_sweep_tile": (_I, [_P, _I, _I]),
    "lbm_set_row_layout": (_I, [_P, _I]),
    "lbm_batch_create": (_I, [ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_double), _I]),
    "lbm_batch_destroy": (_I, [_P]),
    "lbm_batch_set_launch": (_I, [_P, _I, _I]),
    "lbm_batch_set_temporal": (_I, [_P, _I, _I]),
    "lbm_batch_set_packing": (_I, [_P, _I]),
    "lbm_batch_run": (_I, [_P, _I64]),
    "lbm_stream": (_I, [_P, _P, _I, _I, _I, _P]),
    "lbm_equilibrium": (_I, [_P, _P, _P, _I64, _I, _P]),
    "lbm_density": (_I, [_P, _P, _I64, _I, _P]),
    "lbm_velocity": (_I, [_P, _P, _P, _I64, _I, _P]),
    "lbm_collide": (_I, [_P, _P, _P, _D, _P, _I64, _I, _P]),
    "lbm_bc_apply": (_I, [_P, _P, _P, _P, _P]),
    "lbm_bc_pressure": (_I, [_P, _P, _P, _P, _P]),
    "lbm_moments_sum_noalloc": (_I, [_P, _I64, _P, _P]),
    "lbm_copy_probe": (_I, [_P, _P, _I64, _P]),
    "lbm_row_sweep_probe": (_I, [_P, _P, _I, _I, _I, _P]),
    "lbm_row_sweep_probe_ordered": (_I, [_P, _P, _I, _I, _I, _I, _P]),
}

_lib = None
_lock = threading.Lock()


def load():
    """Load liblbm_hip.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with "
                    "`python -m lattice_boltzmann_parallel_solver_amd.build`")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(rc: int) -> None:
    """Raise LbmError for a nonzero status code."""
    if rc != LBM_OK:
        msg = load().lbm_last_error().decode(errors="replace")
        if rc == LBM_ERR_UNSTABLE:
            raise LbmUnstable(rc, msg)
        raise LbmError(rc, msg)


def declared_symbols(header: str | None = None) -> list:
    """Function names declared in include/lbm_hip.h."""
    import re

    if header is None:
        header = os.path.join(os.path.dirname(_HERE), "include", "lbm_hip.h")
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))
