"""Coherent discrete operators (reference: numerics/core_ops_coherent.py:45-112).

Forward-difference gradient with zero normal derivative on the high faces,
backward-difference divergence with zero exterior flux on the low faces; their
composition is the reflective-Neumann 5-point Laplacian. Evaluated on the GPU
by the same cell functions the fused pressure solve uses (float64).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .._ops import apply_op

__all__ = ["coherence_diagnostics", "divergence_coherent", "gradient_coherent", "laplacian_coherent"]

Array2D = np.ndarray


def gradient_coherent(p: Array2D, dx: float, dy: float) -> tuple[Array2D, Array2D]:
    """Forward difference gradient; dpdx[:, -1] = 0 and dpdy[-1, :] = 0."""
    p = np.asarray(p, dtype=float)
    return apply_op("grad_coherent", p.shape, dx, dy, p, nout=2)


def divergence_coherent(u: Array2D, v: Array2D, dx: float, dy: float) -> Array2D:
    """Backward difference divergence consistent with the forward gradient."""
    u = np.asarray(u, dtype=float)
    return apply_op("div_coherent", u.shape, dx, dy, u, v)


def laplacian_coherent(p: Array2D, dx: float, dy: float) -> Array2D:
    """divergence_coherent(gradient_coherent(p))."""
    p = np.asarray(p, dtype=float)
    return apply_op("lap_coherent", p.shape, dx, dy, p)


@dataclass(slots=True)
class CoherenceDiagnostics:
    max_abs_diff: float
    l2_rel_diff: float
    lap_norm: float
    divgrad_norm: float
    constant_laplacian_max: float
    mean_laplacian: float


def coherence_diagnostics(nx: int, ny: int, dx: float, dy: float, seed: int = 0) -> CoherenceDiagnostics:
    """Operator coherence metrics on a seeded random field (core_ops_coherent.py:94-112)."""
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((ny, nx))
    gx, gy = gradient_coherent(p, dx, dy)
    lap = laplacian_coherent(p, dx, dy)
    divgrad = divergence_coherent(gx, gy, dx, dy)
    diff = lap - divgrad
    lap_norm = float(np.linalg.norm(lap)) or 1.0
    c = np.full((ny, nx), 7.123, dtype=p.dtype)
    lap_c = laplacian_coherent(c, dx, dy)
    return CoherenceDiagnostics(
        max_abs_diff=float(np.max(np.abs(diff))),
        l2_rel_diff=float(np.linalg.norm(diff)) / lap_norm,
        lap_norm=lap_norm,
        divgrad_norm=float(np.linalg.norm(divgrad)),
        constant_laplacian_max=float(np.max(np.abs(lap_c))),
        mean_laplacian=float(np.mean(lap)),
    )
