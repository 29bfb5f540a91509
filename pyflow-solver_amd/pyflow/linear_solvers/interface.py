"""Linear solve wrapper (reference: linear_solvers/interface.py:9-53).

``solve(A, b, method='cg', ...)`` with ``A`` one of PyFlow's pressure operators
(``PinnedNegLaplacian`` / ``DirichletPosLaplacian``, the operators
``solve_pressure_poisson`` uses) runs the whole Jacobi-PCG on the GPU: the
fused matrix-free kernels of libpyflow_hip with x0 = 0 and SciPy's ``cg``
recurrence and stopping rule. Any other ``A`` (an arbitrary user matrix or
operator) is outside the accelerated path and is handed to SciPy unchanged,
exactly as the reference does.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Literal

import numpy as np

from .. import _hip as H
from .. import _device as D
from .._device import DeviceCavity, cavity_params, precision_code

__all__ = ["LinearSolveResult", "solve", "PinnedNegLaplacian", "DirichletPosLaplacian", "device_cg"]


@dataclass(slots=True)
class LinearSolveResult:
    x: np.ndarray
    converged: bool
    iterations: int
    residual_norm: float
    method: str


def _call_solver(solver_fn, A, b, x0, tol, maxiter, **kwargs):
    try:  # SciPy >= 1.12 uses rtol
        return solver_fn(A, b, x0=x0, rtol=tol, maxiter=maxiter, **kwargs)
    except TypeError:
        return solver_fn(A, b, x0=x0, tol=tol, maxiter=maxiter, **kwargs)


def _linear_operator_base():
    import scipy.sparse.linalg as spla
    return spla.LinearOperator


class _PyflowOperator(_linear_operator_base()):
    """Matrix-free PyFlow pressure operator; matvec runs on the GPU (float64)."""

    op = H.OP_NEG_LAP_PINNED
    _opk = "neg_lap_pinned"

    def __init__(self, nx: int, ny: int, dx: float, dy: float):
        super().__init__(dtype=np.float64, shape=(nx * ny, nx * ny))
        self.nx, self.ny, self.dx, self.dy = int(nx), int(ny), float(dx), float(dy)

    def _matvec(self, x):
        from .._ops import apply_op
        x = np.asarray(x, dtype=np.float64).reshape(-1)
        return apply_op(self._opk, (self.ny, self.nx), self.dx, self.dy, x).reshape(-1)


class PinnedNegLaplacian(_PyflowOperator):
    """``-laplacian_coherent`` with row 0 replaced by the identity (pressure_solver.py:97-105)."""

    op = H.OP_NEG_LAP_PINNED
    _opk = "neg_lap_pinned"


class DirichletPosLaplacian(_PyflowOperator):
    """``+laplacian_coherent`` with identity rows on all boundary cells (pressure_solver.py:69-77)."""

    op = H.OP_POS_LAP_DIRICHLET
    _opk = "pos_lap_dirichlet"


_cache: dict = {}


def device_cg(op: _PyflowOperator, b: np.ndarray, tol: float, maxiter: int, jacobi: bool,
              precision=None, inner_rtol: float = None):
    """SciPy-``cg``-equivalent PCG on the GPU. Returns (x, info, true_iters, ||b-Ax||, converged)."""
    prec = precision_code(precision)
    key = (op.nx, op.ny, prec)
    params = cavity_params(op.nx, op.ny, dx=op.dx, dy=op.dy, disable_advection=True, lin_tol=tol,
                           lin_maxiter=maxiter, jacobi=jacobi, op=op.op, mode=H.MODE_SOLVE,
                           inner_rtol=D.DEFAULT_INNER_RTOL if inner_rtol is None else inner_rtol)
    cav = _cache.get(key)
    if cav is None:
        if len(_cache) > 8:
            for c in _cache.values():
                c.close()
            _cache.clear()
        cav = DeviceCavity(params, prec)
        _cache[key] = cav
    cav.set_params(params)
    cav.set_rhs(np.asarray(b, dtype=np.float64).reshape(op.ny, op.nx))
    d = cav.run(1)
    x = cav.solution().reshape(-1)
    return x, int(d.rp_info), int(d.rp_iterations), float(d.rp_residual), bool(d.rp_converged)


def solve(A, b, method: Literal["cg", "bicgstab"] = "bicgstab", tol: float = 1e-8, maxiter: int = 500,
          **kwargs) -> LinearSolveResult:
    """Generic linear solve; PyFlow operators with method='cg' run on the GPU."""
    b = np.asarray(b)
    M = kwargs.get("M")
    if method == "cg" and isinstance(A, _PyflowOperator) and set(kwargs) <= {"M"}:
        from .preconditioners import is_pyflow_jacobi
        if M is None or is_pyflow_jacobi(M, A.nx, A.ny, A.dx, A.dy):
            x, info, _its, rnorm, converged = device_cg(A, b, tol, maxiter, jacobi=M is not None)
            iterations = info if info > 0 else 0
            return LinearSolveResult(x=x, converged=converged, iterations=iterations, residual_norm=rnorm,
                                     method=method)
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    x0 = np.zeros_like(b)
    if not isinstance(A, spla.LinearOperator) and not sp.issparse(A):
        A = sp.csr_matrix(A)
    if method == "cg":
        x, info = _call_solver(spla.cg, A, b, x0, tol, maxiter, **kwargs)
    else:
        x, info = _call_solver(spla.bicgstab, A, b, x0, tol, maxiter, **kwargs)
    r = b - (A @ x)
    rnorm = float(np.linalg.norm(r))
    bnorm = float(np.linalg.norm(b)) or 1.0
    converged = (info == 0) and (rnorm <= tol * bnorm)
    iterations = info if info > 0 else 0
    return LinearSolveResult(x=x, converged=converged, iterations=iterations, residual_norm=rnorm, method=method)
