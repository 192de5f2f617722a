"""Velocity-magnitude snapshots and checkpoints of a device-resident lattice.

The reference's von Karman driver saves |u| every 100 steps as a ``.npy``
file written with MPI-IO (experiments.py:638-647 calling save_mpiio,
parallelization_utils.py:203-253), then renders it.  Here the field is
computed on the GPU (``Lattice.speed``) and written by ``write_npy``, which
lays out the file as save_mpiio does: format 1.0 magic, a little-endian
16-bit header length, the header dictionary as ``str(dict)`` padded with
spaces and a newline to a multiple of 16 bytes, then the C-order data.
Rendering to images is left to the caller (``on_snapshot``).

Checkpoints (``Lattice.save_checkpoint`` / ``load_checkpoint``) hold the
device state itself, so a resumed run continues bit for bit.
"""

from __future__ import annotations

import os
from typing import Callable, List, Optional, Tuple

import numpy as np

from .solver import Lattice

_MAGIC = b"\x93NUMPY\x01\x00"


def npy_header(shape: Tuple[int, ...], dtype=np.float64) -> bytes:
    """Header of a version-1.0 .npy file as save_mpiio builds it
    (parallelization_utils.py:226-233)."""
    descr = np.lib.format.dtype_to_descr(np.dtype(dtype))
    text = str({"descr": descr, "fortran_order": False, "shape": tuple(int(n) for n in shape)})
    while (len(text) + len(_MAGIC) + 2) % 16 != 15:
        text += " "
    text += "\n"
    return _MAGIC + np.int16(len(text)).astype("<i2").tobytes() + text.encode("latin-1")


def write_npy(path: str, array: np.ndarray) -> None:
    """Write a 2-D array to `path` in the layout of save_mpiio."""
    a = np.ascontiguousarray(array)
    if a.ndim != 2:
        raise ValueError("write_npy writes two-dimensional arrays")
    with open(path, "wb") as fh:
        fh.write(npy_header(a.shape, a.dtype))
        fh.write(a.tobytes(order="C"))


def speed_snapshots(lattice: Lattice, omega: float, steps: int, every: int = 100,
                    directory: Optional[str] = None, name: str = "vel_norm_{step}.npy",
                    on_snapshot: Optional[Callable[[int, np.ndarray], None]] = None) -> List[int]:
    """Advance `steps` steps and take |u| after step i + 1 whenever i % every
    == 0 (the reference's loop, experiments.py:699-707 with its i % 100 test),
    i.e. after steps 1, every + 1, 2 every + 1, ...  Each snapshot is written
    to `directory`/`name` (if a directory is given) and passed to
    `on_snapshot(step, field)`.  Returns the step counts of the snapshots."""
    if every < 1:
        raise ValueError("every must be >= 1")
    taken = []
    done = 0
    for i in range(0, steps, every):
        k = i + 1 - done
        lattice.run(omega, k)
        done += k
        field = lattice.speed()
        if directory is not None:
            write_npy(os.path.join(directory, name.format(step=done)), field)
        if on_snapshot is not None:
            on_snapshot(done, field)
        taken.append(done)
    if done < steps:
        lattice.run(omega, steps - done)
    return taken
