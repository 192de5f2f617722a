"""The measurement probes behind bench.py's ceilings compute what they claim:
the streaming copy copies, and the row-sweep probe writes every value of
every row once, transformed by its one multiply-add, for pitches of one and
two column slots per lane and row counts that leave a short last block."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from lattice_boltzmann_parallel_solver_amd import _lib

    return _lib, _lib.load()


def test_copy_probe_copies():
    import torch

    lib, L = _lib()
    a = torch.arange(4096 * 5, dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    s = torch.cuda.current_stream().cuda_stream
    lib.check(L.lbm_copy_probe(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), a.numel() * 4,
                               ctypes.c_void_p(s)))
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert L.lbm_copy_probe(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), 20, ctypes.c_void_p(s)) != 0


CX = [0, 1, 0, -1, 0, 1, -1, -1, 1]


def _slot(lx, rows, x, interleaved):
    """Slot of row x in the range-interleaved order (offset t of every
    rows-row range together; the offsets the last range lacks packed over the
    full ranges), or x in the plain order."""
    if not interleaved:
        return x
    ranges = -(-lx // rows)
    whole = rows if lx % rows == 0 else lx % rows
    b, t = divmod(x, rows)
    return t * ranges + b if t < whole else whole * ranges + (t - whole) * (ranges - 1) + b


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("lx,pitch,rows", [(37, 128, 8), (64, 2048, 32), (9, 2304, 4), (5, 64, 32), (70, 2048, 34)])
def test_row_sweep_probe_writes_every_row_once(lx, pitch, rows, mode):
    """Every output row x, plane i, holds the transformed input of row x (or,
    with the pull pattern, of row x - cx_i), each at its row's place in the
    plain or the interleaved order."""
    import torch

    lib, L = _lib()
    rng = np.random.default_rng(lx + pitch)
    host = rng.standard_normal(9 * lx * pitch).astype(np.float32)
    src = torch.from_numpy(host).cuda()
    dst = torch.full_like(src, float("nan"))
    s = torch.cuda.current_stream().cuda_stream
    lib.check(L.lbm_row_sweep_probe_ordered(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), lx,
                                            pitch, rows, mode, ctypes.c_void_p(s)))
    torch.cuda.synchronize()
    il, pull = bool(mode & 1), bool(mode & 2)
    rows_in = host.reshape(lx, 9, pitch).astype(np.float64)
    want = np.empty_like(rows_in)
    for x in range(lx):
        for i in range(9):
            xs = (x - CX[i]) % lx if pull else x
            v = rows_in[_slot(lx, rows, xs, il), i]
            w = np.empty_like(v)
            w[0::2] = v[0::2] * np.float32(1.0001) + 0.5
            w[1::2] = v[1::2] * np.float32(0.9999) + 0.25
            want[_slot(lx, rows, x, il), i] = w
    got = dst.cpu().numpy().reshape(lx, 9, pitch)
    assert np.allclose(got, want, rtol=1e-6, atol=1e-6)
    assert L.lbm_row_sweep_probe_ordered(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), lx, pitch,
                                         rows, 4, ctypes.c_void_p(s)) != 0
    assert L.lbm_row_sweep_probe(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), lx, 63, rows,
                                 ctypes.c_void_p(s)) != 0
