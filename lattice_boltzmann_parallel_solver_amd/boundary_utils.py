"""Drop-in for the reference's ``src/boundary_utils.py``: boundary composites.

A composite is an ordered list of boundary rules.  Called on arrays it
behaves like the reference's closure ``boundary(f_pre, f_post, density,
velocity, f_prev) -> f_post`` (run as HIP kernels); handed to
``lattice_boltzmann_step`` or ``Lattice`` it compiles to one device program
(last writer wins per node and direction) applied inside the fused step.

The MPI decomposition of the reference (parallel_von_karman_boundary_conditions
with ghost layers, boundary_utils.py:113-205) reduces here to one domain: the
whole lattice lives on one GPU.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np

from .boundary_conditions import (BoundaryRule, PeriodicPressure, inlet, moving_wall, outlet,
                                  periodic_with_pressure_variations, rigid_object, rigid_wall)


class Composite(BoundaryRule):
    """Ordered boundary rules applied as one (boundary_utils.py:35-53, 87-108, 164-203)."""

    def __init__(self, rules: Sequence[BoundaryRule]):
        super().__init__()
        self.rules = list(rules)
        if sum(isinstance(r, PeriodicPressure) for r in self.rules) > 1:
            raise ValueError("at most one periodic_with_pressure_variations rule per composite")

    def descriptors(self, lx, ly):
        out = []
        for r in self.rules:
            out.extend(r.descriptors(lx, ly))
        return out

    @property
    def has_pressure(self) -> bool:
        return any(isinstance(r, PeriodicPressure) for r in self.rules)

    def __call__(self, f_pre_streaming, f_post_streaming, density=None, velocity=None, f_prev=None):
        """Apply the composite to arrays as the reference's closure does.

        With a pressure rule, f_pre is fixed in place and re-streamed into a new
        f_post (boundary_utils.py:103-104); the post-streaming rules then
        update f_post in place and it is returned.
        """
        lx, ly = int(f_post_streaming.shape[0]), int(f_post_streaming.shape[1])
        prog = self.program(lx, ly)
        if self.has_pressure:
            from .lattice_boltzmann_method import streaming

            if density is None or velocity is None:
                raise ValueError("a pressure-periodic composite needs density and velocity")
            pressure = next(r for r in self.rules if isinstance(r, PeriodicPressure))
            f_pre_streaming = pressure(f_pre_streaming, density, velocity)
            f_post_streaming = streaming(f_pre_streaming)
        if prog.has_outlet and f_prev is None:
            raise ValueError("a composite with an outlet needs f_prev")
        return self._apply_post(f_pre_streaming, f_post_streaming, f_prev=f_prev)


def compose(*rules: BoundaryRule) -> Composite:
    """Composite of primitives in application order, e.g. the serial von
    Karman closure inlet -> outlet -> rigid_object of
    milestoneQuickFunctionCalls.py:304-311."""
    return Composite(rules)


def _edge(lx, ly, where):
    m = np.zeros((lx, ly), dtype=bool)
    if where == "top":
        m[:, -1] = True
    elif where == "bottom":
        m[:, 0] = True
    return m


def couette_flow_boundary_conditions(lx: int, ly: int, U: float, avg_density: float) -> Composite:
    """Rigid top wall, then the bottom wall moving with (U, 0)
    (boundary_utils.py:9-55)."""
    return Composite([rigid_wall(_edge(lx, ly, "top")),
                      moving_wall(_edge(lx, ly, "bottom"), np.array([U, 0]), avg_density)])


def poiseuille_flow_boundary_conditions(lx: int, ly: int, p_in: float, p_out: float) -> Composite:
    """Pressure-periodic in x, then bottom and top rigid walls
    (boundary_utils.py:58-110)."""
    m = np.zeros((lx, ly), dtype=bool)
    m[0, :] = True
    m[-1, :] = True
    return Composite([periodic_with_pressure_variations(m, p_in, p_out),
                      rigid_wall(_edge(lx, ly, "bottom")),
                      rigid_wall(_edge(lx, ly, "top"))])


def plate_mask(lx: int, ly: int, plate_size: int, y_offset: int = 0) -> np.ndarray:
    """Plate column x = lx//4 over rows [ly//2 - d//2, ly//2 + d//2) + y_offset
    (milestoneQuickFunctionCalls.py:308-309)."""
    if plate_size < 2:
        raise ValueError("plate_size must be >= 2")
    m = np.zeros((lx, ly), dtype=bool)
    y0 = ly // 2 - plate_size // 2 + y_offset
    if y0 < 0 or y0 + 2 * (plate_size // 2) > ly:
        raise ValueError("plate does not fit in the channel")
    m[lx // 4, y0:y0 + 2 * (plate_size // 2)] = True
    return m


def von_karman_boundary_conditions(lx: int, ly: int, density_in: float, velocity_in: float, plate_size: int,
                                   walls: bool = False, y_offset: int = 0) -> Composite:
    """Inlet, outlet and the plate at x = lx//4; with walls=True rigid bottom
    and top walls follow, and y_offset shifts the plate.

    walls=False, y_offset=0 is the reference's channel (periodic in y,
    boundary_utils.py:113-205 on one domain, milestoneQuickFunctionCalls.py:304-311).
    """
    rules = [inlet((lx, ly), density_in, velocity_in), outlet(),
             rigid_object(plate_mask(lx, ly, plate_size, y_offset))]
    if walls:
        rules.append(rigid_wall(_edge(lx, ly, "bottom")))
        rules.append(rigid_wall(_edge(lx, ly, "top")))
    return Composite(rules)


def parallel_von_karman_boundary_conditions(coord2d, n_local_x: int, n_local_y: int, lx: int, ly: int,
                                            x_size: int, y_size: int, density_in: float, velocity_in: float,
                                            plate_size: int) -> Composite:
    """boundary_utils.py:113-205 for the one-domain case (x_size = y_size = 1).

    The lattice is not split: the returned composite acts on the full
    lx x ly lattice without ghost layers.
    """
    if x_size != 1 or y_size != 1 or tuple(coord2d) != (0, 0) or (n_local_x, n_local_y) != (lx, ly):
        raise NotImplementedError("the whole lattice lives on one GPU: use x_size = y_size = 1")
    return von_karman_boundary_conditions(lx, ly, density_in, velocity_in, plate_size)
