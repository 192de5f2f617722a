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


@pytest.mark.parametrize("interleaved", [0, 1])
@pytest.mark.parametrize("lx,pitch,rows", [(37, 128, 8), (64, 2048, 32), (9, 2304, 4), (5, 64, 32), (70, 2048, 34)])
def test_row_sweep_probe_writes_every_row_once(lx, pitch, rows, interleaved):
    import torch

    lib, L = _lib()
    rng = np.random.default_rng(lx + pitch)
    src = torch.from_numpy(rng.standard_normal(9 * lx * pitch).astype(np.float32)).cuda()
    dst = torch.full_like(src, float("nan"))
    s = torch.cuda.current_stream().cuda_stream
    lib.check(L.lbm_row_sweep_probe_ordered(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), lx,
                                            pitch, rows, interleaved, ctypes.c_void_p(s)))
    torch.cuda.synchronize()
    x = src.view(-1, 2)
    want = torch.stack([torch.addcmul(torch.full_like(x[:, 0], 0.5), x[:, 0], torch.full_like(x[:, 0], 1.0001)),
                        torch.addcmul(torch.full_like(x[:, 1], 0.25), x[:, 1], torch.full_like(x[:, 1], 0.9999))], 1)
    assert torch.allclose(dst.view(-1, 2), want, rtol=1e-6, atol=1e-6)
    assert L.lbm_row_sweep_probe(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), lx, 63, rows,
                                 ctypes.c_void_p(s)) != 0
