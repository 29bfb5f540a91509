"""Advection operators (reference: numerics/operators/advection.py:4-84).

QUICK (edge-padded, cell-centred velocity on both faces, branch on >= 0, the
reference's negative-flow face weights reproduced as written) and first-order
upwind (branch on > 0). Evaluated on the GPU by the cell functions the fused
predictor uses (float64).
"""
from __future__ import annotations

import numpy as np

from ..._ops import apply_op

__all__ = ["quick_face_interpolation_1d", "advect_upwind", "advect_quick"]


def quick_face_interpolation_1d(phi, i, flow_positive):
    """QUICK face value between i and i+1 in 1D (advection.py:4-15)."""
    if flow_positive:
        return (3 / 8) * phi[i - 1] + (6 / 8) * phi[i] - (1 / 8) * phi[i + 1]
    return (3 / 8) * phi[i + 2] + (6 / 8) * phi[i + 1] - (1 / 8) * phi[i]


def advect_upwind(u, v, field, dx, dy):
    """First-order upwind convective derivative u*df/dx + v*df/dy."""
    f = np.asarray(field, dtype=float)
    return apply_op("adv_upwind", f.shape, dx, dy, u, v, f)


def advect_quick(u, v, field, dx, dy):
    """QUICK convective derivative (u*phi_e - u*phi_w)/dx + (v*phi_n - v*phi_s)/dy."""
    f = np.asarray(field, dtype=float)
    return apply_op("adv_quick", f.shape, dx, dy, u, v, f)
