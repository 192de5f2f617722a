"""Multi-process CPU oracle: the reference's MPI domain decomposition restated
with forked processes and shared memory.

TEST INFRASTRUCTURE ONLY.  Used by ``tests/`` (it must equal the
single-process oracle bit for bit) and by the ``cpu_baseline`` leg of
``bench.py`` (the reference's parallel numpy path timed on all host cores).

What it restates
----------------
* Process grid: ``get_xy_size`` (``src/parallelization_utils.py:55-92``).
  One or two processes give 1 x P; a prime P > 2 is rejected.  Otherwise
  both sides start at ceil(sqrt(P)); while their product differs from P the
  x side shrinks (product too large) or the y side grows (too small).
* Block sizes: ``get_local_coords`` (``src/parallelization_utils.py:95-120``).
  Each block is lx // x_size by ly // y_size cells, and the last block in
  each direction also takes the remainder.
* Halo exchange: ``communication`` (``src/parallelization_utils.py:6-52``)
  swaps one ghost layer of post-collision populations with the four
  neighbours before streaming (``src/lattice_boltzmann_method.py:217-218``).
  Here every process writes its block of the post-collision field into a
  shared array, waits at a barrier and gathers its ghost frame (edges and
  corners, periodic wrap) from it.  The data reaching each block is the same.
* The step itself is the single-process oracle's (``lbm_oracle.step``)
  restricted to the block, with the boundary ops of a composite's
  ``oracle_ops`` applied in their order to the block's part of each mask.
  The pre-streaming pressure fix reads columns owned by other blocks and is
  not restated here (``NotImplementedError``); the channel workloads do not
  use it.

Processes are forked, so shared memory is an anonymous mapping created
before the fork; call this before anything in the process initialises a GPU.
"""

from __future__ import annotations

import math
import mmap
import multiprocessing as mp
import os
import time
from typing import Callable, Optional, Tuple

import numpy as np

from . import lbm_oracle as O


def get_xy_size(processes: int) -> Tuple[int, int]:
    """(x_size, y_size) of the Cartesian process grid (parallelization_utils.py:55-92)."""
    def is_prime(n):
        return n > 2 and all(n % d for d in range(2, int(math.isqrt(n)) + 1))

    if processes in (1, 2):
        return (1, processes) if processes == 2 else (1, 1)
    if is_prime(processes):
        raise ValueError("a prime number of processes (other than 1 and 2) has no 2-D grid here")
    lo = hi = math.ceil(math.sqrt(processes))
    while lo * hi != processes:
        if lo * hi > processes:
            lo -= 1
        else:
            hi += 1
    return lo, hi


def block_of(rank: int, lx: int, ly: int, x_size: int, y_size: int) -> Tuple[int, int, int, int]:
    """(x0, nx, y0, ny) of a rank's block; ranks run y-fastest over the grid and
    the last block in each direction takes the remainder (parallelization_utils.py:95-120)."""
    cx, cy = divmod(rank, y_size)
    bx, by = lx // x_size, ly // y_size
    nx = lx - bx * (x_size - 1) if cx + 1 == x_size else bx
    ny = ly - by * (y_size - 1) if cy + 1 == y_size else by
    return cx * bx, nx, cy * by, ny


class _Shared:
    """Float64 arrays in one anonymous shared mapping (inherited by forked children)."""

    def __init__(self, shapes):
        sizes = [int(np.prod(s)) * 8 for s in shapes]
        self._map = mmap.mmap(-1, max(1, sum(sizes)))
        self.arrays = []
        off = 0
        for s, n in zip(shapes, sizes):
            self.arrays.append(np.frombuffer(self._map, dtype=np.float64, count=n // 8, offset=off).reshape(s))
            off += n


def _local_ops(ops, x0, nx, y0, ny, lx):
    out = []
    for kind, data in ops or ():
        if kind == "reflect":
            rules = []
            for mask, dirs, k in data:
                m = np.asarray(mask)[x0:x0 + nx, y0:y0 + ny]
                if m.any():
                    rules.append((m.copy(), tuple(dirs), tuple(k)))
            if rules:
                out.append(("reflect", rules))
        elif kind == "inlet":
            if x0 == 0:
                out.append(("inlet", np.asarray(data, dtype=np.float64)))
        elif kind == "outlet":
            if x0 + nx == lx:
                if nx < 2:
                    raise ValueError("the block holding the outlet column needs at least two columns")
                out.append(("outlet", None))
        else:
            raise NotImplementedError(f"boundary op {kind!r} spans blocks")
    return out


def _worker(rank, grid, shape, omega, warmup, steps, ops, arrays, barrier, times):
    try:
        _work(rank, grid, shape, omega, warmup, steps, ops, arrays, barrier, times)
    except BaseException:
        barrier.abort()  # release the other processes instead of leaving them at the barrier
        raise


def _work(rank, grid, shape, omega, warmup, steps, ops, arrays, barrier, times):
    lx, ly = shape
    x0, nx, y0, ny = block_of(rank, lx, ly, *grid)
    F, PRE, RHO, U = arrays[0:2], arrays[2], arrays[3:5], arrays[5:7]
    local = _local_ops(ops, x0, nx, y0, ny, lx)
    xs = (np.arange(x0 - 1, x0 + nx + 1)) % lx
    ys = (np.arange(y0 - 1, y0 + ny + 1)) % ly
    bx, by = slice(x0, x0 + nx), slice(y0, y0 + ny)
    total = warmup + steps
    for s in range(total):
        if s == warmup:
            barrier.wait()
            t0 = time.perf_counter()
        src, dst = s % 2, (s + 1) % 2
        f_in = F[src][bx, by]
        pre = O.collide(f_in, O.equilibrium(RHO[src][bx, by], U[src][bx, by]), omega)
        PRE[bx, by] = pre
        barrier.wait()
        post = np.ascontiguousarray(O.streaming(PRE[np.ix_(xs, ys)])[1:-1, 1:-1])
        for kind, data in local:
            if kind == "reflect":
                for m, dirs, k in data:
                    for i in dirs:
                        post[m, O.OPP[i]] = pre[m, i] - k[i] if k[i] != 0.0 else pre[m, i]
            elif kind == "inlet":
                post[0, :, :] = data
            else:
                for i in (3, 6, 7):
                    post[-1, :, i] = f_in[-2, :, i]
        rho = O.density(post)
        F[dst][bx, by] = post
        RHO[dst][bx, by] = rho
        U[dst][bx, by] = O.velocity(rho, post)
        barrier.wait()
    if steps:
        times[rank] = time.perf_counter() - t0


def run(f: np.ndarray, rho: np.ndarray, u: np.ndarray, omega: float, steps: int,
        boundary: Optional[Callable] = None, processes: Optional[int] = None, warmup: int = 0):
    """`warmup + steps` oracle steps on `processes` forked processes.

    Returns (f, rho, u, seconds of the last `steps` steps, (x_size, y_size))."""
    processes = processes or default_processes()
    grid = get_xy_size(processes)
    lx, ly = rho.shape
    if lx < grid[0] or ly < grid[1]:
        raise ValueError("more processes than cells along an axis")
    ops = getattr(boundary, "oracle_ops", None) if boundary is not None else None
    if boundary is not None and ops is None:
        raise TypeError("boundary has no oracle_ops description")
    sh = _Shared([(lx, ly, 9)] * 3 + [(lx, ly)] * 2 + [(lx, ly, 2)] * 2 + [(processes,)])
    A = sh.arrays
    A[0][...] = f
    A[3][...] = rho
    A[5][...] = u
    arrays = [A[0], A[1], A[2], A[3], A[4], A[5], A[6]]
    ctx = mp.get_context("fork")
    barrier = ctx.Barrier(processes)
    procs = [ctx.Process(target=_worker, args=(r, grid, (lx, ly), omega, warmup, steps, ops, arrays, barrier, A[7]))
             for r in range(processes)]
    for p in procs:
        p.start()
    for p in procs:
        p.join()
    bad = [p.exitcode for p in procs if p.exitcode != 0]
    if bad:
        raise RuntimeError(f"oracle worker failed (exit codes {bad})")
    out = (warmup + steps) % 2
    return (A[0 + out].copy(), A[3 + out].copy(), A[5 + out].copy(), float(A[7].max()), grid)


def default_processes() -> int:
    """Host cores available to this process (at most 16), lowered to the
    nearest count that get_xy_size accepts."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    n = max(1, min(16, n))
    while True:
        try:
            get_xy_size(n)
            return n
        except ValueError:
            n -= 1
