"""Drop-in for the reference's ``src/boundary_conditions.py``.

Each factory returns a callable with the reference closure's signature and
in-place semantics (it mutates and returns its population argument); the
call runs as a HIP kernel.  The same object also describes itself to the
fused GPU step: ``descriptor(lx, ly)`` gives its C ABI descriptor and
``program(lx, ly)`` a compiled single-rule program, so composites
(boundary_utils) can run inside ``lattice_boltzmann_step`` and ``Lattice``.

The index helpers (get_wall_indices, get_corner_indices,
remove_corner_indices_from_boundary) are host-side set-up, as in the
reference.
"""

from __future__ import annotations

from typing import Tuple

import numpy as np

from . import _lib
from ._device import stream, to_device, torch_cuda, write_back
from ._lib import check


def get_wall_indices(boundary: np.ndarray) -> np.ndarray:
    """Directions a straight wall bounces back; the first full edge decides
    (boundary_conditions.py:8-28)."""
    if np.all(boundary[0, :]):
        return np.array([1, 5, 8])
    if np.all(boundary[-1, :]):
        return np.array([3, 6, 7])
    if np.all(boundary[:, 0]):
        return np.array([4, 7, 8])
    if np.all(boundary[:, -1]):
        return np.array([2, 5, 6])
    return np.array([])


def get_corner_indices(boundary: np.ndarray) -> np.ndarray:
    """Corner nodes of an obstacle with the two directions each reflects
    (boundary_conditions.py:31-56)."""
    assert len(boundary.shape) == 2
    idx = np.argwhere(boundary)
    x0, x1 = np.amin(idx[:, 0]), np.amax(idx[:, 0])
    y0, y1 = np.amin(idx[:, 1]), np.amax(idx[:, 1])
    return np.array([((x0, y1), (1, 8)), ((x0, y0), (1, 5)), ((x1 + 1, y1), (3, 7)), ((x1 + 1, y0), (3, 6))])


def remove_corner_indices_from_boundary(boundary: np.ndarray, corner_indices: np.ndarray) -> np.ndarray:
    """Clear the corner nodes in place (boundary_conditions.py:59-75)."""
    assert len(boundary.shape) == 2
    for corner, _ in corner_indices:
        boundary[corner[0], corner[1]] = False
    return boundary


class BoundaryRule:
    """One boundary primitive: a GPU-backed closure plus its descriptor."""

    kind: int = 0

    def __init__(self):
        self._programs = {}

    def descriptor(self, lx: int, ly: int) -> dict:
        raise NotImplementedError

    def descriptors(self, lx: int, ly: int) -> list:
        return [self.descriptor(lx, ly)]

    def program(self, lx: int, ly: int):
        """Compiled program of this rule alone (cached per lattice shape)."""
        from .solver import BoundaryProgram

        key = (int(lx), int(ly))
        p = self._programs.get(key)
        if p is None:
            p = BoundaryProgram(lx, ly, self.descriptors(lx, ly))
            self._programs[key] = p
        return p

    def _apply_post(self, f_pre, f_post, f_prev=None):
        """Apply the post-streaming rules to f_post in place on the GPU."""
        lx, ly = int(f_post.shape[0]), int(f_post.shape[1])
        torch = torch_cuda()
        prog = self.program(lx, ly)
        post, _ = to_device(f_post, dtype=torch.float64)
        pre, _ = to_device(f_pre, dtype=torch.float64) if f_pre is not None else (post, None)
        prev = to_device(f_prev, dtype=torch.float64)[0] if f_prev is not None else None
        check(_lib.load().lbm_bc_apply(prog.handle, pre.data_ptr(), post.data_ptr(),
                                       None if prev is None else prev.data_ptr(), stream()))
        write_back(f_post, post)
        return f_post


class _MaskRule(BoundaryRule):
    def __init__(self, boundary: np.ndarray):
        super().__init__()
        assert boundary.dtype == "bool"
        self.boundary = np.array(boundary, dtype=bool, copy=True)

    def _check(self, f_pre, f_post):
        assert self.boundary.shape == tuple(f_pre.shape[0:2])
        assert self.boundary.shape == tuple(f_post.shape[0:2])


class RigidWall(_MaskRule):
    """Same-node bounce-back on a straight wall (boundary_conditions.py:78-113)."""

    kind = _lib.BC_RIGID_WALL

    def descriptor(self, lx, ly):
        return {"kind": self.kind, "mask": self.boundary}

    def __call__(self, f_pre_streaming, f_post_streaming):
        self._check(f_pre_streaming, f_post_streaming)
        return self._apply_post(f_pre_streaming, f_post_streaming)


class MovingWall(_MaskRule):
    """Bounce-back with wall momentum, f*_i - 2 w_i rho_w (c_i.u_w)/cs^2
    (boundary_conditions.py:170-214)."""

    kind = _lib.BC_MOVING_WALL

    def __init__(self, boundary, u_w, avg_density):
        super().__init__(boundary)
        u_w = np.asarray(u_w, dtype=np.float64).reshape(-1)
        assert u_w.shape == (2,)
        if np.ndim(avg_density) != 0:
            raise ValueError("avg_density must be a scalar")
        self.u_w = u_w
        self.avg_density = float(avg_density)

    def descriptor(self, lx, ly):
        return {"kind": self.kind, "mask": self.boundary, "u_w": tuple(self.u_w), "rho_w": self.avg_density}

    def __call__(self, f_pre_streaming, f_post_streaming):
        self._check(f_pre_streaming, f_post_streaming)
        return self._apply_post(f_pre_streaming, f_post_streaming)


class RigidObject(_MaskRule):
    """Bounce-back on an obstacle between columns, with corner rules
    (boundary_conditions.py:116-167).  Like the reference, constructing it
    clears the corner nodes in the caller's mask."""

    kind = _lib.BC_RIGID_OBJECT

    def __init__(self, boundary):
        assert boundary.dtype == "bool"
        original = np.array(boundary, copy=True)
        corners = get_corner_indices(boundary)
        remove_corner_indices_from_boundary(boundary, corners)
        super().__init__(original)

    def descriptor(self, lx, ly):
        return {"kind": self.kind, "mask": self.boundary}

    def __call__(self, f_pre_streaming, f_post_streaming):
        self._check(f_pre_streaming, f_post_streaming)
        return self._apply_post(f_pre_streaming, f_post_streaming)


class Inlet(BoundaryRule):
    """Equilibrium inlet on column x = 0 (boundary_conditions.py:217-254)."""

    kind = _lib.BC_INLET

    def __init__(self, lattice_grid_shape: Tuple[int, int], density_in: float, velocity_in: float):
        super().__init__()
        self.shape = tuple(int(s) for s in lattice_grid_shape)
        self.density_in = float(density_in)
        self.velocity_in = float(velocity_in)

    def descriptor(self, lx, ly):
        assert (lx, ly) == self.shape, f"inlet built for {self.shape}, lattice is {(lx, ly)}"
        return {"kind": self.kind, "rho_in": self.density_in, "u_in": self.velocity_in}

    def __call__(self, f_post_streaming):
        assert tuple(f_post_streaming.shape[0:2]) == self.shape
        return self._apply_post(None, f_post_streaming)


class Outlet(BoundaryRule):
    """Copy directions 3, 6, 7 from column x = -2 of the step input
    (boundary_conditions.py:257-283)."""

    kind = _lib.BC_OUTLET

    def descriptor(self, lx, ly):
        return {"kind": self.kind}

    def __call__(self, f_previous, f_post_streaming):
        assert tuple(f_previous.shape) == tuple(f_post_streaming.shape)
        return self._apply_post(f_post_streaming, f_post_streaming, f_prev=f_previous)


class PeriodicPressure(_MaskRule):
    """Pre-streaming periodic boundary with a pressure drop
    (boundary_conditions.py:286-348)."""

    kind = _lib.BC_PRESSURE_PERIODIC

    def __init__(self, boundary, p_in, p_out):
        assert boundary.dtype == "bool"
        assert np.all(boundary[0, :] == boundary[-1, :]) or np.all(boundary[:, 0] == boundary[:, -1])
        super().__init__(boundary)
        self.p_in = float(p_in)
        self.p_out = float(p_out)

    def descriptor(self, lx, ly):
        return {"kind": self.kind, "mask": self.boundary, "p_in": self.p_in, "p_out": self.p_out}

    def __call__(self, f_pre_streaming, density, velocity):
        assert self.boundary.shape == tuple(f_pre_streaming.shape[0:2])
        torch = torch_cuda()
        lx, ly = self.boundary.shape
        prog = self.program(lx, ly)
        f, _ = to_device(f_pre_streaming, dtype=torch.float64)
        rho, _ = to_device(density, dtype=torch.float64)
        u, _ = to_device(velocity, dtype=torch.float64)
        check(_lib.load().lbm_bc_pressure(prog.handle, f.data_ptr(), rho.data_ptr(), u.data_ptr(), stream()))
        write_back(f_pre_streaming, f)
        return f_pre_streaming


def rigid_wall(boundary: np.ndarray) -> RigidWall:
    """boundary_conditions.py:78-113."""
    return RigidWall(boundary)


def rigid_object(boundary: np.ndarray) -> RigidObject:
    """boundary_conditions.py:116-167."""
    return RigidObject(boundary)


def moving_wall(boundary: np.ndarray, u_w: np.ndarray, avg_density) -> MovingWall:
    """boundary_conditions.py:170-214."""
    return MovingWall(boundary, u_w, avg_density)


def inlet(lattice_grid_shape: Tuple[int, int], density_in: float, velocity_in: float) -> Inlet:
    """boundary_conditions.py:217-254."""
    return Inlet(lattice_grid_shape, density_in, velocity_in)


def outlet() -> Outlet:
    """boundary_conditions.py:257-283."""
    return Outlet()


def periodic_with_pressure_variations(boundary: np.ndarray, p_in: float, p_out: float) -> PeriodicPressure:
    """boundary_conditions.py:286-348."""
    return PeriodicPressure(boundary, p_in, p_out)
