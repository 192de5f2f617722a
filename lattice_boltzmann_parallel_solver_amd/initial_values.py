"""Drop-in for the reference's ``src/initial_values.py`` (host-side set-up).

These build the initial density and velocity fields of the reference's
experiments; they run once per simulation on the host, as in the reference,
and the GPU computes f = equilibrium(rho, u) from them
(``Lattice.set_equilibrium`` or ``equilibrium_distr_func``).
"""

from __future__ import annotations

from typing import Tuple

import numpy as np


def sinusoidal_density_x(lattice_grid_shape: Tuple[int, int], initial_p0: float, epsilon: float):
    """rho(x) = p0 + eps sin(2 pi x / lx), u = 0 (initial_values.py:38-64)."""
    assert initial_p0 + epsilon < 1, "rho can potentially be >= 1"
    assert initial_p0 - epsilon > 0, "rho can potentially be <= 0"
    lx, ly = lattice_grid_shape
    x = np.arange(0, lx)
    column = initial_p0 + epsilon * np.sin(np.divide(2 * np.pi * x, lx))
    rho = np.repeat(column[:, None], ly, axis=1)
    return rho, np.zeros((lx, ly, 2))


def sinusoidal_velocity_x(lattice_grid_shape: Tuple[int, int], epsilon: float):
    """rho = 1, ux(y) = eps sin(2 pi y / ly), uy = 0 (initial_values.py:67-93)."""
    assert np.abs(epsilon) < 0.1, "|u| can potentially be >= 0.1"
    lx, ly = lattice_grid_shape
    y = np.arange(0, ly)
    row = epsilon * np.sin(np.divide(2 * np.pi * y, ly))
    u = np.zeros((lx, ly, 2))
    u[..., 0] = row[None, :]
    return np.ones((lx, ly)), u


def density_1_velocity_0_initial(lattice_grid_shape: Tuple[int, int]):
    """rho = 1, u = 0 (initial_values.py:96-106)."""
    return np.ones(lattice_grid_shape), np.zeros(tuple(lattice_grid_shape) + (2,))


def density_1_velocity_x_u0_velocity_y_0_initial(lattice_grid_shape: Tuple[int, int], u0: float):
    """rho = 1, u = (u0, 0) (initial_values.py:109-123)."""
    u = np.empty(tuple(lattice_grid_shape) + (2,))
    u[..., 0] = u0
    u[..., 1] = 0
    return np.ones(lattice_grid_shape), u
