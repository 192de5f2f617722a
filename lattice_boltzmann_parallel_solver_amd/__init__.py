"""MI355X-native D2Q9 lattice-Boltzmann step (gfx950 HIP kernels behind the
reference solver's Python API).

Modules mirroring the reference's ``src/``:
  lattice_boltzmann_method  step, streaming, equilibrium, moments
  boundary_conditions       boundary primitives
  boundary_utils            boundary composites
  initial_values            initial fields
Device-resident API:
  solver.Lattice            one lattice on one GPU, fused step, probes, sums
  solver.BoundaryProgram    compiled boundary descriptors
  solver.Batch              many lattices advanced by one launch per step
"""

from .solver import Batch, BoundaryProgram, Lattice  # noqa: F401

__all__ = ["Batch", "BoundaryProgram", "Lattice"]
