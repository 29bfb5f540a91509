"""Central divergence (reference: numerics/fluid_ops.py:8-41).

``divergence(u, v, dx, dy)`` is the operator the step uses for the dt
safeguard and the continuity residual (``solvers/solver.py:45,68,114``):
central differences inside, one-sided differences across each edge, both
one-sided at the corners. It runs on the GPU through the same cell function
as the fused step (``div_central`` in ``csrc/pf_cells.hpp``). The reference
module's other helpers (``gradient``, ``laplacian`` and the legacy state
wrappers) are not on the projection step and are not provided.
"""
from __future__ import annotations

import numpy as np

from .._ops import apply_op

__all__ = ["divergence"]


def divergence(u: np.ndarray, v: np.ndarray, dx: float, dy: float) -> np.ndarray:
    """Central divergence of (u, v) on (ny, nx) arrays; needs at least 2 x 2 cells."""
    u = np.asarray(u, dtype=float)
    if u.ndim != 2 or u.shape[0] < 2 or u.shape[1] < 2:
        raise IndexError("central divergence needs at least 2x2 cells")
    return apply_op("div_central", u.shape, dx, dy, u, v)
