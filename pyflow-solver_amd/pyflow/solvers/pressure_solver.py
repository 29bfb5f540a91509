"""Pressure projection (reference: solvers/pressure_solver.py:15-219).

``solve_pressure_poisson`` keeps the reference signature, both modes and the
diagnostics keys. The whole projection (coherent divergence -> RHS with mean
removal and pinned cell -> Jacobi-PCG -> gradient correction) runs on the GPU.
``assemble_negative_laplacian`` is kept for API compatibility (tests and
tools use it); the solver itself never assembles a matrix: the Jacobi weights
are applied analytically on the device.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .. import _hip as H
from .. import _device as D
from .._device import cavity_params
from .._mirror import config_precision, mirror_for
from ..core.ghost_fields import State

__all__ = ["assemble_negative_laplacian", "solve_pressure_poisson"]


def assemble_negative_laplacian(nx: int, ny: int, dx: float, dy: float):
    """Sparse matrix of pressure_solver.py:15-40: identity rows on boundary cells,
    5-point ``-Laplace`` rows (centre ``2/dx**2 + 2/dy**2``) inside."""
    N = nx * ny
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    idx = (jj * nx + ii).ravel()
    bnd = ((ii == 0) | (ii == nx - 1) | (jj == 0) | (jj == ny - 1)).ravel()
    rows, cols, vals = [idx[bnd]], [idx[bnd]], [np.ones(int(bnd.sum()))]
    inn = idx[~bnd]
    cx, cy = -1.0 / dx ** 2, -1.0 / dy ** 2
    rows += [inn] * 5
    cols += [inn, inn - 1, inn + 1, inn - nx, inn + nx]
    vals += [np.full(inn.size, 2.0 / dx ** 2 + 2.0 / dy ** 2), np.full(inn.size, cx), np.full(inn.size, cx),
             np.full(inn.size, cy), np.full(inn.size, cy)]
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    return A.tocsr()


def _jacobi_flag(preconditioner, cfg, nx, ny, dx, dy, default_on: bool):
    """Device Jacobi setting for the given preconditioner argument."""
    if preconditioner is None:
        return bool(getattr(cfg, "enable_jacobi_pc", True)) if default_on else False
    from ..linear_solvers.preconditioners import is_pyflow_jacobi
    if is_pyflow_jacobi(preconditioner, nx, ny, dx, dy):
        return True
    raise NotImplementedError(
        "solve_pressure_poisson runs its PCG on the GPU and accepts preconditioner=None or the "
        "PyFlow Jacobi preconditioner of assemble_negative_laplacian(nx, ny, dx, dy)")


def solve_pressure_poisson(state: State, dt: float, dx: float, dy: float, cfg, preconditioner=None,
                           rhs_override=None):
    """Pressure projection to reduce velocity divergence.

    Normal mode: solve ``A p = -(1/dt) div_c(u*)`` with ``A = -Laplace`` (Neumann, cell 0
    pinned), mean-free RHS, Jacobi PCG; then ``u -= dt*grad_c(p)``.
    Manufactured mode (``rhs_override``): ``+Laplace`` with identity rows on every boundary
    cell, no Jacobi unless a preconditioner is passed, no velocity correction.
    """
    u = state._raw("u")
    ny, nx = u.shape[0] - 2, u.shape[1] - 2
    manufactured = rhs_override is not None
    verbose = getattr(cfg, "diagnostics", True) and not getattr(cfg, "force_quiet", False)
    lin_tol = getattr(cfg, "lin_tol", 1e-10)
    if manufactured:
        params = cavity_params(nx, ny, dx=dx, dy=dy, disable_advection=True, lid_velocity=None,
                               lin_tol=lin_tol, lin_maxiter=getattr(cfg, "lin_maxiter", 500),
                               jacobi=_jacobi_flag(preconditioner, cfg, nx, ny, dx, dy, False),
                               op=H.OP_POS_LAP_DIRICHLET, mode=H.MODE_SOLVE,
                               inner_rtol=float(getattr(cfg, "inner_rtol", D.DEFAULT_INNER_RTOL)))
    else:
        params = cavity_params(nx, ny, dx=dx, dy=dy, disable_advection=True, lid_velocity=None,
                               lin_tol=lin_tol, lin_maxiter=getattr(cfg, "lin_maxiter", 400),
                               jacobi=_jacobi_flag(preconditioner, cfg, nx, ny, dx, dy, True),
                               op=H.OP_NEG_LAP_PINNED, mode=H.MODE_PROJECT,
                               inner_rtol=float(getattr(cfg, "inner_rtol", D.DEFAULT_INNER_RTOL)))
    mirror = mirror_for(state, params, config_precision(cfg, state))
    mirror.before(params)
    if manufactured:
        rhs = np.copy(np.asarray(rhs_override, dtype=np.float64)).reshape(-1)
        mirror.cav.set_rhs(rhs.reshape(ny, nx))
    else:
        mirror.cav.set_dt(float(dt))
    d = mirror.cav.run(1)
    mirror.after()
    mirror.ensure_host()
    p = state._raw("p")[1:-1, 1:-1].copy()
    div_norm = d.div_norm if not manufactured else d.rp_residual
    div_before = None if manufactured else d.div_initial
    div_after_raw = None if manufactured else d.div_norm
    if verbose:
        if not manufactured:
            print(f"DIVERGENCE_NORM_BEFORE={div_before}")
            print(f"[pressure_solver] RHS mean subtracted (compatibility) = {d.rhs_mean:.6e}")
            print(f"DIVERGENCE_NORM_AFTER_RAW={div_after_raw}")
            print(f"DIVERGENCE_NORM_FINAL={div_norm}")
    # Linear residual diagnostics: ||A p - rhs|| equals the true CG residual
    # because the pinned component of x is exactly 0 (p = x - x[0] = x).
    lin_residual_norm = d.rp_residual
    rhs_norm = d.bnorm or 1.0
    lin_residual_rel = lin_residual_norm / rhs_norm
    if verbose:
        print(f"LIN_RESIDUAL_NORM={lin_residual_norm}")
        print(f"LIN_RESIDUAL_REL={lin_residual_rel}")
    rp_iters = d.rp_info if d.rp_info > 0 else 0
    if rp_iters == 0 and d.rp_residual > lin_tol:
        rp_iters = 1
    red_factor = None
    if div_before is not None and div_norm not in (None, 0.0):
        red_factor = div_before / div_norm
    diagnostics = {
        "Rp_iterations": rp_iters,
        "Rp_residual": d.rp_residual,
        "Rp_converged": bool(d.rp_converged),
        "Rp_method": "cg",
        "divergence_norm": div_norm,
        "lin_residual_norm": lin_residual_norm,
        "lin_residual_rel": lin_residual_rel,
        "divergence_reduction_factor": red_factor,
        "divergence_norm_initial": div_before,
        "divergence_norm_after_raw": div_after_raw,
        "divergence_norm_final": div_norm,
        "Rp_true_iterations": int(d.rp_iterations),
    }
    return p, diagnostics
