"""Drop-in for the reference's ``src/lattice_boltzmann_method.py``.

Same function names, argument meaning and assertions; every array operation
runs as a HIP kernel (liblbm_hip.so).  numpy inputs return numpy outputs
(copied through the device), float64/float32 torch device tensors return
device tensors.  ``lattice_boltzmann_step`` runs the whole step (collision
with the given density/velocity, streaming, boundary, moments) on the GPU
through a cached device context; the boundary must be one of this package's
boundary objects (they compile to a device program), not an arbitrary
Python callable.
"""

from __future__ import annotations

from collections import OrderedDict
from typing import Callable, Tuple

import numpy as np

from . import _lib
from ._device import back, dtype_code, stream, to_device, torch_cuda
from ._lib import check


def get_velocity_sets() -> np.ndarray:
    """D2Q9 velocities c_i, 0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE
    (lattice_boltzmann_method.py:5-26)."""
    return np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]])


def vel_to_opp_vel_mapping() -> np.ndarray:
    """Opposite direction of each c_i (lattice_boltzmann_method.py:29-39)."""
    return np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])


def get_w_i() -> np.ndarray:
    """D2Q9 weights (lattice_boltzmann_method.py:42-52)."""
    return np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36])


def reynolds_number(L: int, u: float, v: float) -> float:
    """Re = L u / nu (lattice_boltzmann_method.py:55-71)."""
    return np.divide(L * u, v)


def strouhal_number(f: float, L: int, u: float) -> float:
    """St = f L / u (lattice_boltzmann_method.py:74-90)."""
    return np.divide(f * L, u)


def compute_density(prob_densitiy_func):
    """rho = sum_i f_i on the GPU (lattice_boltzmann_method.py:93-105)."""
    assert prob_densitiy_func.shape[-1] == 9
    torch = torch_cuda()
    t, kind = to_device(prob_densitiy_func)
    out = torch.empty(tuple(t.shape[:-1]), dtype=t.dtype, device=t.device)
    check(_lib.load().lbm_density(t.data_ptr(), out.data_ptr(), t.numel() // 9, dtype_code(t), stream()))
    return back(out, kind)


def compute_velocity_field(density_func, prob_density_func):
    """u = (sum_i c_i f_i)/rho, 0 where rho == 0, with the given rho
    (lattice_boltzmann_method.py:108-137)."""
    assert prob_density_func.shape[-1] == 9
    assert prob_density_func.ndim == 3
    assert tuple(density_func.shape) == tuple(prob_density_func.shape[:-1])
    torch = torch_cuda()
    f, kind = to_device(prob_density_func)
    rho, _ = to_device(density_func, dtype=f.dtype)
    out = torch.empty(tuple(f.shape[:-1]) + (2,), dtype=f.dtype, device=f.device)
    check(_lib.load().lbm_velocity(rho.data_ptr(), f.data_ptr(), out.data_ptr(), f.numel() // 9, dtype_code(f),
                                   stream()))
    return back(out, kind)


def streaming(prob_density_func):
    """Periodic streaming f_i(x + c_i) <- f_i(x), bit-exact index shift
    (lattice_boltzmann_method.py:140-159)."""
    assert prob_density_func.shape[-1] == 9
    assert prob_density_func.ndim == 3
    torch = torch_cuda()
    t, kind = to_device(prob_density_func)
    out = torch.empty_like(t)
    lx, ly = int(t.shape[0]), int(t.shape[1])
    if lx * ly:
        check(_lib.load().lbm_stream(t.data_ptr(), out.data_ptr(), lx, ly, dtype_code(t), stream()))
    return back(out, kind)


def equilibrium_distr_func(density_func, velocity_field):
    """Second-order equilibrium, any leading shape (lattice_boltzmann_method.py:162-188)."""
    assert tuple(density_func.shape) == tuple(velocity_field.shape[:-1])
    torch = torch_cuda()
    rho, kind = to_device(density_func)
    u, _ = to_device(velocity_field, dtype=rho.dtype)
    out = torch.empty(tuple(rho.shape) + (9,), dtype=rho.dtype, device=rho.device)
    check(_lib.load().lbm_equilibrium(rho.data_ptr(), u.data_ptr(), out.data_ptr(), rho.numel(), dtype_code(rho),
                                      stream()))
    return back(out, kind)


# cached device contexts for lattice_boltzmann_step, keyed by shape and boundary program
_STEP_CONTEXTS: "OrderedDict[tuple, object]" = OrderedDict()
_MAX_CONTEXTS = 4


def _step_context(lx: int, ly: int, boundary):
    from .solver import Lattice

    prog = None if boundary is None else boundary.program(lx, ly)
    key = (lx, ly, id(prog))
    ctx = _STEP_CONTEXTS.get(key)
    if ctx is None or ctx._program is not prog:
        ctx = Lattice(lx, ly, boundary=prog, storage="f64")
        _STEP_CONTEXTS[key] = ctx
        while len(_STEP_CONTEXTS) > _MAX_CONTEXTS:
            _STEP_CONTEXTS.popitem(last=False)[1].close()
    _STEP_CONTEXTS.move_to_end(key)
    return ctx


def lattice_boltzmann_step(f, density, velocity, omega: float,
                           boundary: Callable = None,
                           parallel_communication: Callable = None) -> Tuple:
    """One lattice-Boltzmann step on the GPU (lattice_boltzmann_method.py:191-228).

    collide with the *given* density/velocity -> stream -> boundary -> moments,
    returning new (f, density, velocity).  FP64, in the reference's arithmetic
    order.  ``boundary`` must be a boundary object of this package (for
    example from boundary_utils); ``parallel_communication`` must be None: the
    whole lattice lives on one GPU.
    """
    assert tuple(f.shape[0:2]) == tuple(density.shape)
    assert tuple(f.shape[0:2]) == tuple(velocity.shape[0:2])
    assert 0 < omega < 2
    if parallel_communication is not None:
        raise NotImplementedError("the whole lattice lives on one GPU; parallel_communication must be None")
    if boundary is not None and not hasattr(boundary, "program"):
        raise TypeError("boundary must be a boundary object of lattice_boltzmann_parallel_solver_amd "
                        "(boundary_utils.compose(...) builds one from primitives); arbitrary Python "
                        "callables cannot run inside the fused GPU step")
    lx, ly = int(f.shape[0]), int(f.shape[1])
    ctx = _step_context(lx, ly, boundary)
    torch_in = False
    try:
        import torch

        torch_in = isinstance(f, torch.Tensor)
    except ImportError:  # pragma: no cover
        pass
    ctx.set_state(f, density, velocity)
    ctx.run(omega, 1)
    if torch_in:
        torch = torch_cuda()
        fo = torch.empty((lx, ly, 9), dtype=torch.float64, device="cuda")
        ro = torch.empty((lx, ly), dtype=torch.float64, device="cuda")
        uo = torch.empty((lx, ly, 2), dtype=torch.float64, device="cuda")
        ctx.state_into(fo, ro, uo)
        return fo, ro, uo
    return ctx.state()
