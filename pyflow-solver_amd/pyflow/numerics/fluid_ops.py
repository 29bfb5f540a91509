"""Fluid operators (reference: numerics/fluid_ops.py).

``divergence`` (central differences inside, one-sided on edges; used for the
continuity metric and the dt safeguard) runs on the GPU through the same cell
function as the fused step. ``gradient``, ``laplacian`` and the legacy helpers
are not on the projection step; they are small NumPy utilities kept for API
compatibility (fluid_ops.py:43-125).
"""
from __future__ import annotations

import numpy as np

from .._ops import apply_op
from ..core.ghost_fields import State, interior_view

__all__ = ["divergence", "gradient", "laplacian", "legacy_divergence_state", "legacy_grad_p_state",
           "apply_pressure_correction"]


def _divergence_impl(u: np.ndarray, v: np.ndarray, dx: float, dy: float) -> np.ndarray:
    u = np.asarray(u, dtype=float)
    if u.shape[0] < 2 or u.shape[1] < 2:
        raise IndexError("central divergence needs at least 2x2 cells")
    return apply_op("div_central", u.shape, dx, dy, u, v)


def divergence(u: np.ndarray, v: np.ndarray, dx: float, dy: float) -> np.ndarray:
    """Central divergence of (u, v) (fluid_ops.py:8-41)."""
    return _divergence_impl(u, v, dx, dy)


def gradient(p: np.ndarray, dx: float, dy: float) -> tuple[np.ndarray, np.ndarray]:
    """Central-difference gradient with one-sided edges (fluid_ops.py:43-67)."""
    ny, nx = p.shape
    dpdx = np.zeros_like(p)
    dpdy = np.zeros_like(p)
    if nx > 2:
        dpdx[:, 1:-1] = (p[:, 2:] - p[:, 0:-2]) / (2.0 * dx)
    if ny > 2:
        dpdy[1:-1, :] = (p[2:, :] - p[0:-2, :]) / (2.0 * dy)
    dpdx[:, 0] = (p[:, 1] - p[:, 0]) / dx
    dpdx[:, -1] = (p[:, -1] - p[:, -2]) / dx
    dpdy[0, :] = (p[1, :] - p[0, :]) / dy
    dpdy[-1, :] = (p[-1, :] - p[-2, :]) / dy
    return dpdx, dpdy


def laplacian(p: np.ndarray, dx: float, dy: float) -> np.ndarray:
    """5-point Laplacian inside, boundary rows/columns copied inward (fluid_ops.py:69-91)."""
    ny, nx = p.shape
    lap = np.zeros_like(p)
    if nx > 2 and ny > 2:
        lap[1:-1, 1:-1] = ((p[1:-1, 2:] - 2.0 * p[1:-1, 1:-1] + p[1:-1, 0:-2]) / (dx * dx)
                           + (p[2:, 1:-1] - 2.0 * p[1:-1, 1:-1] + p[0:-2, 1:-1]) / (dy * dy))
    lap[0, :] = lap[1, :]
    lap[-1, :] = lap[-2, :]
    lap[:, 0] = lap[:, 1]
    lap[:, -1] = lap[:, -2]
    return lap


def legacy_divergence_state(state: State, dx: float = 1.0, dy: float = 1.0) -> np.ndarray:
    return _divergence_impl(interior_view(state.fields["u"]), interior_view(state.fields["v"]), dx, dy)


def legacy_grad_p_state(state: State, dx: float = 1.0, dy: float = 1.0):
    return gradient(interior_view(state.fields["p"]), dx, dy)


def apply_pressure_correction(state: State, p: np.ndarray, dt: float, dx: float, dy: float) -> None:
    interior_view(state.fields["p"])[:] = p
    ui = interior_view(state.fields["u"])
    vi = interior_view(state.fields["v"])
    dpdx, dpdy = gradient(p, dx, dy)
    ui[:] = ui - dt * dpdx
    vi[:] = vi - dt * dpdy
