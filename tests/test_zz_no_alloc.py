"""lbm_moments_sum_noalloc, the reduction that allocates nothing (its block
partial sums live in a fixed device array), against lbm_moments_sum on a
context holding the same populations as its step input: the FP64 sums are
identical, on two calls in a row with the second one smaller (nothing left
in the shared array by the first call is read).  Also: an empty input gives
zeros, a size past LBM_NOALLOC_MAX_CELLS is refused with a message, and a
repeated call leaves the device's free memory unchanged."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_CELLS = 44739072  # LBM_NOALLOC_MAX_CELLS in include/lbm_hip.h


def _noalloc(f_dev, ncells, out_dev):
    import torch

    from lattice_boltzmann_parallel_solver_amd import _lib

    stream = torch.cuda.current_stream().cuda_stream
    return _lib.load().lbm_moments_sum_noalloc(ctypes.c_void_p(f_dev.data_ptr()), ncells,
                                               ctypes.c_void_p(out_dev.data_ptr()), ctypes.c_void_p(stream))


def _context_sums(f):
    from lattice_boltzmann_parallel_solver_amd import Lattice

    lx, ly = f.shape[:2]
    lat = Lattice(lx, ly, storage="f64")
    lat.set_state(f, f.sum(axis=-1), np.zeros((lx, ly, 2)))
    sums = np.array(lat.moments_sum())
    lat.close()
    return sums


def test_noalloc_sums_equal_the_context_sums_on_two_calls():
    import torch

    from lattice_boltzmann_parallel_solver_amd import _lib

    rng = np.random.default_rng(7)
    out = torch.empty(3, dtype=torch.float64, device="cuda")
    seen = []
    for lx, ly in ((1500, 1003), (37, 19)):  # the second call is smaller
        print(f"case: lbm_moments_sum_noalloc on {lx}x{ly} AoS FP64 populations", flush=True)
        f = 0.05 + 0.1 * rng.random((lx, ly, 9))
        f_dev = torch.from_numpy(f).to("cuda")
        out.fill_(np.nan)
        torch.cuda.synchronize()
        _lib.check(_noalloc(f_dev, lx * ly, out))
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        want = _context_sums(f)
        assert np.array_equal(got, want), (lx, ly, got, want)
        seen.append(got)
    assert not np.array_equal(seen[0], seen[1])


def test_noalloc_edge_cases_and_no_device_allocation():
    import torch

    from lattice_boltzmann_parallel_solver_amd import _lib
    from lattice_boltzmann_parallel_solver_amd._lib import LBM_ERR_ARG

    f = 0.1 + 0.01 * np.arange(2 * 3 * 9, dtype=np.float64).reshape(2, 3, 9)
    f_dev = torch.from_numpy(f).to("cuda")
    out = torch.full((3,), np.nan, dtype=torch.float64, device="cuda")
    print("case: lbm_moments_sum_noalloc with 0 cells, 6 cells, and past the limit", flush=True)
    _lib.check(_noalloc(f_dev, 0, out))
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), np.zeros(3))
    _lib.check(_noalloc(f_dev, 6, out))  # warm: the code object and its array are loaded
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        _lib.check(_noalloc(f_dev, 6, out))
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
    assert np.array_equal(out.cpu().numpy(), _context_sums(f))
    rc = _noalloc(f_dev, MAX_CELLS + 1, out)  # refused before any launch
    assert rc == LBM_ERR_ARG
    assert "LBM_NOALLOC_MAX_CELLS" in _lib.load().lbm_last_error().decode()
