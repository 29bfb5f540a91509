"""Jacobi preconditioner (reference: linear_solvers/preconditioners.py:8-25)."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

__all__ = ["jacobi_preconditioner", "JacobiOperator", "is_pyflow_jacobi"]


class JacobiOperator(spla.LinearOperator):
    """``M x = invD * x`` with ``invD = where(D != 0, 1/D, 0)``."""

    def __init__(self, inv_diag: np.ndarray):
        self.inv_diag = np.asarray(inv_diag, dtype=np.float64).reshape(-1)
        n = self.inv_diag.size
        super().__init__(dtype=np.float64, shape=(n, n))

    def _matvec(self, x):
        return self.inv_diag * np.asarray(x).reshape(-1)


def jacobi_preconditioner(A: sp.spmatrix):
    """Diagonal (Jacobi) preconditioner of a sparse matrix."""
    D = np.asarray(A.diagonal()).flatten()
    with np.errstate(divide="ignore"):
        invD = np.where(D != 0, 1.0 / D, 0.0)
    return JacobiOperator(invD)


def is_pyflow_jacobi(M, nx: int, ny: int, dx: float, dy: float) -> bool:
    """True when M is the Jacobi operator of assemble_negative_laplacian(nx, ny, dx, dy):
    1 on boundary cells and 1/(2/dx**2 + 2/dy**2) inside, which the GPU applies analytically."""
    inv = getattr(M, "inv_diag", None)
    if inv is None or inv.size != nx * ny:
        return False
    g = inv.reshape(ny, nx)
    D = 2.0 / dx ** 2 + 2.0 / dy ** 2
    inner = 1.0 / D if D != 0 else 0.0
    expect = np.ones((ny, nx))
    if nx > 2 and ny > 2:
        expect[1:-1, 1:-1] = inner
    return bool(np.array_equal(g, expect))
