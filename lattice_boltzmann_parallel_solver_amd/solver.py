"""Device-resident lattice: the performance API over the C ABI.

``Lattice`` keeps one D2Q9 lattice on one GPU and advances it with the fused
stream+collide kernel.  Its state semantics are those of the reference's
``lattice_boltzmann_step`` loop (lattice_boltzmann_method.py:191-228):
``set_state(f, rho, u)`` loads the step input, ``run(omega, n)`` performs n
reference steps, ``state()`` returns (f_n, rho_n, u_n).

``BoundaryProgram`` is a compiled, ordered list of boundary descriptors
(boundary_conditions.py / boundary_utils.py semantics, last writer wins).
"""

from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import check

_c_p = ctypes.c_void_p


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _is_torch(x) -> bool:
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)


def _stream_of_torch() -> int:
    import torch

    return torch.cuda.current_stream().cuda_stream


class BoundaryProgram:
    """Compiled boundary descriptors for an lx x ly lattice (lbm_bc_compile)."""

    def __init__(self, lx: int, ly: int, descriptors: Sequence[dict]):
        L = _lib.load()
        self.lx, self.ly = int(lx), int(ly)
        n = len(descriptors)
        arr = (_lib.BcDesc * max(1, n))()
        keep = []
        for j, d in enumerate(descriptors):
            arr[j].kind = int(d["kind"])
            arr[j].reserved = 0
            mask = d.get("mask")
            if mask is not None:
                m = np.ascontiguousarray(np.asarray(mask) != 0, dtype=np.uint8)
                if m.shape != (self.lx, self.ly):
                    raise ValueError(f"mask shape {m.shape} does not match lattice {(self.lx, self.ly)}")
                keep.append(m)
                arr[j].mask = m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
            uw = d.get("u_w", (0.0, 0.0))
            arr[j].u_w[0] = float(uw[0])
            arr[j].u_w[1] = float(uw[1])
            arr[j].rho_w = float(d.get("rho_w", 0.0))
            arr[j].rho_in = float(d.get("rho_in", 0.0))
            arr[j].u_in = float(d.get("u_in", 0.0))
            arr[j].p_in = float(d.get("p_in", 0.0))
            arr[j].p_out = float(d.get("p_out", 0.0))
        h = _c_p()
        check(L.lbm_bc_compile(ctypes.byref(h), self.lx, self.ly, arr, n))
        self._h = h
        del keep

    @property
    def handle(self):
        return self._h

    def query(self) -> Tuple[int, int, int]:
        nc = ctypes.c_int()
        ns = ctypes.c_int64()
        fl = ctypes.c_uint32()
        check(_lib.load().lbm_bc_query(self._h, ctypes.byref(nc), ctypes.byref(ns), ctypes.byref(fl)))
        return nc.value, ns.value, fl.value

    @property
    def has_outlet(self) -> bool:
        return bool(self.query()[2] & 1)

    @property
    def has_pressure(self) -> bool:
        return bool(self.query()[2] & 2)

    def export(self) -> dict:
        """Compiled tables: codes (lx, ly), per-class src/flags/k, pressure params."""
        nc, _, _ = self.query()
        codes = np.zeros((self.lx, self.ly), dtype=np.uint16)
        src = np.zeros(nc, dtype=np.uint32)
        flags = np.zeros(nc, dtype=np.uint32)
        k = np.zeros((nc, 9), dtype=np.float64)
        params = np.zeros(4, dtype=np.float64)
        check(_lib.load().lbm_bc_export(self._h, _ptr(codes), _ptr(src), _ptr(flags), _ptr(k), _ptr(params)))
        return {"codes": codes, "src": src, "flags": flags, "k": k, "params": params}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().lbm_bc_destroy(h)
            except Exception:
                pass
            self._h = None


class Lattice:
    """One lattice resident on one GPU (lbm_create / lbm_run / lbm_get_state).

    storage: "f64" (FP64 populations, reference arithmetic order), "f32"
    (FP32 deviations f - w), "f16" (FP16 deviations, FP32 arithmetic).
    boundary: a BoundaryProgram, an object with ``program(lx, ly)``, or None
    (fully periodic).
    """

    def __init__(self, lx: int, ly: int, boundary=None, storage: str = "f64", device: int = 0,
                 stream: Optional[int] = None):
        L = _lib.load()
        if storage not in _lib.STORAGE:
            raise ValueError(f"storage must be one of {sorted(_lib.STORAGE)}")
        self.lx, self.ly, self.storage = int(lx), int(ly), storage
        h = _c_p()
        check(L.lbm_create(ctypes.byref(h), self.lx, self.ly, _lib.STORAGE[storage], int(device)))
        self._h = h
        self.device = int(device)
        if stream is not None:
            check(L.lbm_set_stream(h, _c_p(stream)))
        self._program = None
        self.set_boundary(boundary)

    # -- configuration --------------------------------------------------------
    def set_boundary(self, boundary) -> None:
        if boundary is None:
            prog = None
        elif isinstance(boundary, BoundaryProgram):
            prog = boundary
        elif hasattr(boundary, "program"):
            prog = boundary.program(self.lx, self.ly)
        else:
            raise TypeError("boundary must be a BoundaryProgram or carry .program(lx, ly)")
        check(_lib.load().lbm_set_bc(self._h, prog.handle if prog is not None else None))
        self._program = prog

    def set_launch(self, cells_per_thread: int = 1, xcd_remap: bool = True, nt_stores: bool = False) -> None:
        """Launch shape of the fused step: nodes per thread (1, 2, 4; 0 = aligned
        16-byte groups), XCD-aware block order, non-temporal stores."""
        check(_lib.load().lbm_set_launch(self._h, int(cells_per_thread), int(bool(xcd_remap)), int(bool(nt_stores))))

    def set_temporal(self, steps_per_launch: int = 2, rows_per_block: int = 0) -> None:
        """Advance `steps_per_launch` (1 to 4) steps per kernel launch; two or
        more steps keep the intermediate states on chip (identical results,
        1/K of the population traffic per step).  Rows wider than one kernel
        block run as column tiles with two steps; three and four need FP32 or
        FP16 rows of at most 1024 columns.  Raises LbmError where the kernel
        does not apply (pressure fix, fewer than three rows, more than 128
        boundary classes, LDS)."""
        check(_lib.load().lbm_set_temporal(self._h, int(steps_per_launch), int(rows_per_block)))

    def set_sweep_tile(self, threads: int = 0, cols_per_thread: int = 0) -> None:
        """Two-step launches in column tiles of blocks of at most `threads`
        threads (a multiple of 64; 0 = whole rows up to 1024 threads), each
        tile with a halo column on each side; cols_per_thread 1 or 2 (0: 2 for
        FP32/FP16, FP64 always 1).  Identical results."""
        check(_lib.load().lbm_set_sweep_tile(self._h, int(threads), int(cols_per_thread)))

    def set_row_layout(self, interleave: bool = True) -> None:
        """Long two-step runs (64 steps or more) sweep the state in a row order
        that stores the rows the blocks touch at once next to each other,
        moving it there and back around the run (default on; used for FP32
        rows of more than 1024 columns and FP64 column tiles with 16 or more
        rows per block, where it measured faster).  Identical results; the
        state is in the plain order between runs."""
        check(_lib.load().lbm_set_row_layout(self._h, 1 if interleave else 0))

    # -- state ------------------------------------------------------------------
    def _shape_check(self, f=None, rho=None, u=None):
        if f is not None:
            assert tuple(f.shape) == (self.lx, self.ly, 9), f"f must have shape {(self.lx, self.ly, 9)}"
        if rho is not None:
            assert tuple(rho.shape) == (self.lx, self.ly), f"density must have shape {(self.lx, self.ly)}"
        if u is not None:
            assert tuple(u.shape) == (self.lx, self.ly, 2), f"velocity must have shape {(self.lx, self.ly, 2)}"

    def set_state(self, f, rho, u) -> None:
        """Load the step input (f, rho, u); numpy arrays or float64 device tensors."""
        self._shape_check(f, rho, u)
        L = _lib.load()
        if _is_torch(f):
            import torch

            ts = [t.detach().to(dtype=torch.float64).contiguous() for t in (f, rho, u)]
            for t in ts:
                if t.device.type != "cuda":
                    raise ValueError("torch inputs must be device tensors")
            torch.cuda.current_stream().synchronize()
            check(L.lbm_set_state(self._h, ts[0].data_ptr(), ts[1].data_ptr(), ts[2].data_ptr(), _lib.DEVICE))
        else:
            a = [np.ascontiguousarray(x, dtype=np.float64) for x in (f, rho, u)]
            check(L.lbm_set_state(self._h, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), _lib.HOST))

    def set_equilibrium(self, rho, u) -> None:
        """Load f = equilibrium(rho, u) with (rho, u) as the step input."""
        self._shape_check(None, rho, u)
        L = _lib.load()
        if _is_torch(rho):
            import torch

            ts = [t.detach().to(dtype=torch.float64).contiguous() for t in (rho, u)]
            torch.cuda.current_stream().synchronize()
            check(L.lbm_set_state_equilibrium(self._h, ts[0].data_ptr(), ts[1].data_ptr(), _lib.DEVICE))
        else:
            a = [np.ascontiguousarray(x, dtype=np.float64) for x in (rho, u)]
            check(L.lbm_set_state_equilibrium(self._h, _ptr(a[0]), _ptr(a[1]), _lib.HOST))

    def run(self, omega: float, steps: int) -> None:
        """Advance `steps` reference steps (asynchronous)."""
        check(_lib.load().lbm_run(self._h, float(omega), int(steps)))

    def state(self, want=("f", "rho", "u")):
        """(f, rho, u) of the current step as float64 numpy arrays."""
        f = np.empty((self.lx, self.ly, 9)) if "f" in want else None
        rho = np.empty((self.lx, self.ly)) if "rho" in want else None
        u = np.empty((self.lx, self.ly, 2)) if "u" in want else None
        check(_lib.load().lbm_get_state(self._h, None if f is None else _ptr(f), None if rho is None else _ptr(rho),
                                        None if u is None else _ptr(u), _lib.HOST))
        return f, rho, u

    def speed(self) -> np.ndarray:
        """|u| of every node of the current state, (lx, ly) float64: the field
        the reference saves every 100 steps (experiments.py:638-640)."""
        out = np.empty((self.lx, self.ly))
        check(_lib.load().lbm_speed(self._h, _ptr(out), _lib.HOST))
        return out

    def checkpoint(self) -> bytes:
        """The device state as bytes (64-byte header + payload); see
        `restore`.  Resuming continues bit for bit in every storage."""
        L = _lib.load()
        n = ctypes.c_int64()
        check(L.lbm_checkpoint_bytes(self._h, ctypes.byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        check(L.lbm_checkpoint_save(self._h, buf.ctypes.data, n.value))
        return buf.tobytes()

    def restore(self, blob) -> None:
        """Load a `checkpoint` taken from a lattice of the same shape and
        storage (the boundary and probe are this lattice's own)."""
        buf = np.frombuffer(bytes(blob), dtype=np.uint8)
        check(_lib.load().lbm_checkpoint_load(self._h, buf.ctypes.data, buf.size))

    def save_checkpoint(self, path: str) -> None:
        with open(path, "wb") as fh:
            fh.write(self.checkpoint())

    def load_checkpoint(self, path: str) -> None:
        with open(path, "rb") as fh:
            self.restore(fh.read())

    def state_into(self, f=None, rho=None, u=None) -> None:
        """Write (f, rho, u) into float64 device tensors (any may be None)."""
        import torch

        for t in (f, rho, u):
            if t is not None and (t.dtype != torch.float64 or not t.is_contiguous() or t.device.type != "cuda"):
                raise ValueError("outputs must be contiguous float64 device tensors")
        self._shape_check(f, rho, u)
        torch.cuda.current_stream().synchronize()
        check(_lib.load().lbm_get_state(self._h, None if f is None else f.data_ptr(),
                                        None if rho is None else rho.data_ptr(),
                                        None if u is None else u.data_ptr(), _lib.DEVICE))

    # -- diagnostics --------------------------------------------------------------
    def set_probe(self, x: int, y: int, capacity: int) -> None:
        check(_lib.load().lbm_set_probe(self._h, int(x), int(y), int(capacity)))

    def probe(self, max_rows: Optional[int] = None) -> np.ndarray:
        """Recorded (rho, ux, uy) rows, oldest first."""
        n = self.steps if max_rows is None else int(max_rows)
        out = np.zeros((max(1, n), 3))
        cnt = ctypes.c_int64()
        check(_lib.load().lbm_read_probe(self._h, _ptr(out), n, ctypes.byref(cnt)))
        return out[:cnt.value]

    def moments_sum(self) -> Tuple[float, float, float]:
        """FP64 (sum rho, sum rho*ux, sum rho*uy) of the current state."""
        out = (ctypes.c_double * 3)()
        check(_lib.load().lbm_moments_sum(self._h, out))
        return out[0], out[1], out[2]

    def synchronize(self) -> None:
        check(_lib.load().lbm_synchronize(self._h))

    @property
    def steps(self) -> int:
        s = ctypes.c_int64()
        check(_lib.load().lbm_steps_done(self._h, ctypes.byref(s)))
        return s.value

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().lbm_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several lattices advanced together, one kernel launch per step
    (lbm_batch_*).  Lattices must share device and storage; omegas[i] is the
    relaxation of lattices[i].  Close the batch before its lattices."""

    def __init__(self, lattices: Sequence[Lattice], omegas: Sequence[float], nodes_per_thread: int = 1,
                 use_graph: bool = True, temporal: int = 1, rows_per_block: int = 0, packed: int = 0):
        if len(lattices) != len(omegas) or not lattices:
            raise ValueError("need one omega per lattice and at least one lattice")
        L = _lib.load()
        self.lattices = list(lattices)
        handles = (_c_p * len(lattices))(*[lat._h.value for lat in lattices])
        oms = (ctypes.c_double * len(lattices))(*[float(o) for o in omegas])
        h = _c_p()
        check(L.lbm_batch_create(ctypes.byref(h), handles, oms, len(lattices)))
        self._h = h
        self.set_launch(nodes_per_thread, use_graph)
        self.set_temporal(temporal, rows_per_block)
        self.set_packing(packed)

    def set_launch(self, nodes_per_thread: int = 1, use_graph: bool = True) -> None:
        check(_lib.load().lbm_batch_set_launch(self._h, int(nodes_per_thread), int(bool(use_graph))))

    def set_temporal(self, steps_per_launch: int = 2, rows_per_block: int = 0) -> None:
        """Two steps per launch for every lattice the sweep kernel supports
        (identical results); the others keep one step per launch.  3: in large
        batches, FP32 lattices of 385-512 or 641-1024 columns take three steps
        per launch, the rest as with 2."""
        check(_lib.load().lbm_batch_set_temporal(self._h, int(steps_per_launch), int(rows_per_block)))

    def set_packing(self, packed: int = 0) -> None:
        """With two steps per launch: one launch per group of equal row width
        (0, the default and the fastest measured), or one packed launch per
        step pair for the FP32/FP16 lattices of at most 1024 columns (1: one
        block per packed block, 2: a persistent grid).  Identical results."""
        check(_lib.load().lbm_batch_set_packing(self._h, int(packed)))

    def run(self, steps: int) -> None:
        """Advance every lattice `steps` reference steps (synchronous)."""
        check(_lib.load().lbm_batch_run(self._h, int(steps)))

    @property
    def nodes(self) -> int:
        return sum(l.lx * l.ly for l in self.lattices)

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.load().lbm_batch_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
