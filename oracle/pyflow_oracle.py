"""FP64 NumPy oracle for PyFlow's projection time step.

TEST INFRASTRUCTURE ONLY. This module is the checker that the HIP path is
compared against. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it. The product package
(``pyflow``) never imports it and has no CPU fallback.

It restates, in plain vectorised NumPy on float64 arrays, the arithmetic of the
reference implementation (beautifulplanet/PyFlow-Solver, ``src/pyflow``). Every
function cites the reference lines it follows. Element-wise expressions keep
the reference's operation order (no re-association), so every operator below
is bit-identical to the reference's on the same inputs; only reductions
(``np.dot``/``np.linalg.norm``/``np.mean``) carry BLAS/pairwise summation order.

Pinning (what checks each function, so a wrong sign/index/term fails a test):
  * coherent gradient/divergence/laplacian  -- golden vectors written by the
    reference itself (tests/golden), lap == div(grad) identity, constant-field
    null space, brute-force assembled 5-point Neumann matrix on small grids.
  * central divergence                      -- golden vectors; u=X, v=Y gives 2
    in the interior; linear fields exact on edges/corners; right-edge column 0.
  * advect_quick / advect_upwind            -- golden vectors from the reference
    (the only spec of the non-standard QUICK face formulas) plus the
    x^2+y^2 accuracy ordering test.
  * pinned SpMV / Dirichlet SpMV            -- dense assembly vs. closed form.
  * cg                                      -- bit-identical to the installed
    SciPy ``cg`` (the reference's dependency) on the same operator; dense
    direct solve on tiny systems.
  * solve_pressure_poisson / step           -- golden trajectories written by
    the reference (tests/golden/make_golden.py), plus the reference's own
    invariant tests restated in tests/.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np

__all__ = [
    "gradient_coherent", "divergence_coherent", "laplacian_coherent",
    "divergence_central", "advect_quick", "advect_upwind", "advect_quick_loops",
    "advect_upwind_loops", "jacobi_inv_diag_assembled",
    "neg_laplacian_pinned", "pos_laplacian_dirichlet", "jacobi_inv_diag",
    "assemble_negative_laplacian_diag", "cg", "SolveResult", "solve",
    "solve_pressure_poisson", "OracleState", "allocate_state", "step",
    "run_steps",
]


# --------------------------------------------------------------------------
# Coherent operator pair (reference: numerics/core_ops_coherent.py:45-84)
# --------------------------------------------------------------------------
def gradient_coherent(p: np.ndarray, dx: float, dy: float):
    """Forward differences, zero normal derivative on the high faces.

    Follows ``core_ops_coherent.py:45-61``: ``gx[:, :-1] = (p[:,1:]-p[:,:-1])/dx``,
    ``gx[:, -1] = 0``; same along y. Degenerate shapes (<2) give zeros.
    """
    ny, nx = p.shape
    gx = np.zeros_like(p)
    gy = np.zeros_like(p)
    if ny < 2 or nx < 2:
        return gx, gy
    gx[:, :-1] = (p[:, 1:] - p[:, :-1]) / dx
    gy[:-1, :] = (p[1:, :] - p[:-1, :]) / dy
    return gx, gy


def divergence_coherent(u: np.ndarray, v: np.ndarray, dx: float, dy: float):
    """Backward differences with zero exterior flux on the low faces.

    Follows ``core_ops_coherent.py:63-79``. The reference accumulates into a
    zero array, x part first, then y part: ``div = (0 + xpart) + ypart``.
    """
    ny, nx = u.shape
    if ny < 2 or nx < 2:
        return np.zeros_like(u)
    xpart = np.empty_like(u)
    xpart[:, 1:] = (u[:, 1:] - u[:, :-1]) / dx
    xpart[:, 0] = u[:, 0] / dx
    ypart = np.empty_like(v)
    ypart[1:, :] = (v[1:, :] - v[:-1, :]) / dy
    ypart[0, :] = v[0, :] / dy
    return (np.zeros_like(u) + xpart) + ypart


def laplacian_coherent(p: np.ndarray, dx: float, dy: float):
    """``divergence_coherent(gradient_coherent(p))`` (core_ops_coherent.py:81-84)."""
    gx, gy = gradient_coherent(p, dx, dy)
    return divergence_coherent(gx, gy, dx, dy)


# --------------------------------------------------------------------------
# Central divergence used for continuity and the dt safeguard
# (reference: numerics/fluid_ops.py:8-28, _divergence_impl)
# --------------------------------------------------------------------------
def divergence_central(u: np.ndarray, v: np.ndarray, dx: float, dy: float):
    """Central differences inside, one-sided normal differences on edges.

    Follows ``fluid_ops.py:8-28`` assignment by assignment: interior, left
    column, row 0, row ny-1, then the four corners. The reference never assigns
    the right column between the corners, so those cells stay 0.
    """
    ny, nx = u.shape
    div = np.zeros_like(u)
    div[1:-1, 1:-1] = (u[1:-1, 2:] - u[1:-1, 0:-2]) / (2.0 * dx) + (v[2:, 1:-1] - v[0:-2, 1:-1]) / (2.0 * dy)
    div[1:-1, 0] = (u[1:-1, 1] - u[1:-1, 0]) / dx + (v[2:, 0] - v[0:-2, 0]) / (2.0 * dy)
    div[0, 1:-1] = (u[0, 2:] - u[0, :-2]) / (2.0 * dx) + (v[1, 1:-1] - v[0, 1:-1]) / dy
    div[-1, 1:-1] = (u[-1, 2:] - u[-1, :-2]) / (2.0 * dx) + (v[-1, 1:-1] - v[-2, 1:-1]) / dy
    div[0, 0] = (u[0, 1] - u[0, 0]) / dx + (v[1, 0] - v[0, 0]) / dy
    div[0, -1] = (u[0, -1] - u[0, -2]) / dx + (v[1, -1] - v[0, -1]) / dy
    div[-1, 0] = (u[-1, 1] - u[-1, 0]) / dx + (v[-1, 0] - v[-2, 0]) / dy
    div[-1, -1] = (u[-1, -1] - u[-1, -2]) / dx + (v[-1, -1] - v[-2, -1]) / dy
    return div


# --------------------------------------------------------------------------
# Advection (reference: numerics/operators/advection.py:17-84)
# --------------------------------------------------------------------------
_C3, _C6, _C1 = 3 / 8, 6 / 8, 1 / 8


def _quick_face(a, b, c):
    # ((3/8)*a + (6/8)*b) - (1/8)*c, evaluated left to right as in Python.
    return (_C3 * a + _C6 * b) - _C1 * c


def advect_quick(u, v, f, dx, dy):
    """QUICK convective term exactly as ``advection.py:45-84`` computes it.

    Edge padding by 2 (``np.pad(..., 2, mode='edge')``); both faces use the
    cell-centred velocity; branches on ``>= 0``; for negative flow the east
    face is ``3/8 W + 6/8 P - 1/8 E`` (non-standard, reproduced as is).
    """
    ny, nx = f.shape
    ph = np.pad(f, 2, mode="edge")
    P = ph[2:ny + 2, 2:nx + 2]
    E = ph[2:ny + 2, 3:nx + 3]
    W = ph[2:ny + 2, 1:nx + 1]
    WW = ph[2:ny + 2, 0:nx]
    N = ph[3:ny + 3, 2:nx + 2]
    S = ph[1:ny + 1, 2:nx + 2]
    SS = ph[0:ny, 2:nx + 2]
    pos_u = u >= 0
    pos_v = v >= 0
    phi_e = np.where(pos_u, _quick_face(E, P, W), _quick_face(W, P, E))
    phi_w = np.where(pos_u, _quick_face(P, W, WW), _quick_face(WW, W, P))
    phi_n = np.where(pos_v, _quick_face(N, P, S), _quick_face(S, P, N))
    phi_s = np.where(pos_v, _quick_face(P, S, SS), _quick_face(SS, S, P))
    return (u * phi_e - u * phi_w) / dx + (v * phi_n - v * phi_s) / dy


def advect_upwind(u, v, f, dx, dy):
    """First-order upwind term exactly as ``advection.py:17-43`` computes it.

    Branches on ``> 0``; at i=0 (j=0) the backward difference is ``(f-f)/d``;
    the forward difference clamps the neighbour index at the high edge.
    """
    ny, nx = f.shape
    fW = np.concatenate([f[:, :1], f[:, :-1]], axis=1)        # f[j, i-1] (i=0 -> f[j,0])
    fE = np.concatenate([f[:, 1:], f[:, -1:]], axis=1)        # f[j, min(i+1,nx-1)]
    fS = np.concatenate([f[:1, :], f[:-1, :]], axis=0)
    fN = np.concatenate([f[1:, :], f[-1:, :]], axis=0)
    fx = np.where(u > 0, (f - fW) / dx, (fE - f) / dx)
    fy = np.where(v > 0, (f - fS) / dy, (fN - f) / dy)
    return u * fx + v * fy


# --------------------------------------------------------------------------
# Loop-faithful forms: the reference's cost shape, one Python iteration per
# cell (advection.py:17-84 loops over j, i; pressure_solver.py:15-40 assembles
# a LIL matrix cell by cell on every step). Same per-cell arithmetic as the
# vectorised forms above, so the results are bitwise equal (tested); they
# exist to time the reference's own algorithm on the host (CPU baseline ii).
# --------------------------------------------------------------------------
def advect_quick_loops(u, v, f, dx, dy):
    ny, nx = f.shape
    ph = np.pad(f, 2, mode="edge")
    out = np.zeros_like(f)
    for j in range(ny):
        for i in range(nx):
            r, c = j + 2, i + 2
            uc, vc = u[j, i], v[j, i]
            if uc >= 0:
                fe = (_C3 * ph[r, c + 1] + _C6 * ph[r, c]) - _C1 * ph[r, c - 1]
                fw = (_C3 * ph[r, c] + _C6 * ph[r, c - 1]) - _C1 * ph[r, c - 2]
            else:
                fe = (_C3 * ph[r, c - 1] + _C6 * ph[r, c]) - _C1 * ph[r, c + 1]
                fw = (_C3 * ph[r, c - 2] + _C6 * ph[r, c - 1]) - _C1 * ph[r, c]
            if vc >= 0:
                fn = (_C3 * ph[r + 1, c] + _C6 * ph[r, c]) - _C1 * ph[r - 1, c]
                fs = (_C3 * ph[r, c] + _C6 * ph[r - 1, c]) - _C1 * ph[r - 2, c]
            else:
                fn = (_C3 * ph[r - 1, c] + _C6 * ph[r, c]) - _C1 * ph[r + 1, c]
                fs = (_C3 * ph[r - 2, c] + _C6 * ph[r - 1, c]) - _C1 * ph[r, c]
            out[j, i] = (uc * fe - uc * fw) / dx + (vc * fn - vc * fs) / dy
    return out


def advect_upwind_loops(u, v, f, dx, dy):
    ny, nx = f.shape
    out = np.zeros_like(f)
    for j in range(ny):
        for i in range(nx):
            c = f[j, i]
            if u[j, i] > 0:
                fx = (c - f[j, i - 1]) / dx if i > 0 else (c - c) / dx
            else:
                fx = (f[j, min(i + 1, nx - 1)] - c) / dx
            if v[j, i] > 0:
                fy = (c - f[j - 1, i]) / dy if j > 0 else (c - c) / dy
            else:
                fy = (f[min(j + 1, ny - 1), i] - c) / dy
            out[j, i] = u[j, i] * fx + v[j, i] * fy
    return out


def jacobi_inv_diag_assembled(nx: int, ny: int, dx: float, dy: float):
    """Jacobi weights the way the reference obtains them every step: a LIL
    matrix filled cell by cell (identity rows on the boundary, 5-point rows
    inside), converted to CSR, its diagonal inverted."""
    import scipy.sparse as sp
    A = sp.lil_matrix((nx * ny, nx * ny))
    cx, cy = 1.0 / dx ** 2, 1.0 / dy ** 2
    for j in range(ny):
        for i in range(nx):
            k = j * nx + i
            if i in (0, nx - 1) or j in (0, ny - 1):
                A[k, k] = 1.0
                continue
            A[k, k] = 2.0 / dx ** 2 + 2.0 / dy ** 2
            A[k, k - 1] = -cx
            A[k, k + 1] = -cx
            A[k, k - nx] = -cy
            A[k, k + nx] = -cy
    D = np.asarray(A.tocsr().diagonal()).ravel()
    with np.errstate(divide="ignore"):
        return np.where(D != 0, 1.0 / D, 0.0)


# --------------------------------------------------------------------------
# Pressure operators and Jacobi (reference: solvers/pressure_solver.py:15-106,
# linear_solvers/preconditioners.py:8-25)
# --------------------------------------------------------------------------
def neg_laplacian_pinned(x: np.ndarray, nx: int, ny: int, dx: float, dy: float):
    """``_NegLaplacianPinned._matvec`` (pressure_solver.py:97-105)."""
    lap = laplacian_coherent(x.reshape(ny, nx), dx, dy)
    out = (-lap).reshape(-1)
    out[0] = x[0]
    return out


def pos_laplacian_dirichlet(x: np.ndarray, nx: int, ny: int, dx: float, dy: float):
    """``_PosLaplacianDirichlet._matvec`` (pressure_solver.py:69-77): identity rows on all boundary cells."""
    out = laplacian_coherent(x.reshape(ny, nx), dx, dy).reshape(-1).copy()
    mask = np.zeros((ny, nx), dtype=bool)
    mask[0, :] = True
    mask[-1, :] = True
    mask[:, 0] = True
    mask[:, -1] = True
    b = mask.reshape(-1)
    out[b] = x[b]
    return out


def assemble_negative_laplacian_diag(nx: int, ny: int, dx: float, dy: float):
    """Diagonal of ``assemble_negative_laplacian`` (pressure_solver.py:15-40).

    1.0 on every boundary cell, ``2.0/dx**2 + 2.0/dy**2`` inside.
    """
    D = np.ones((ny, nx))
    if nx > 2 and ny > 2:
        D[1:-1, 1:-1] = 2.0 / dx ** 2 + 2.0 / dy ** 2
    return D.reshape(-1)


def jacobi_inv_diag(nx: int, ny: int, dx: float, dy: float):
    """``invD = where(D != 0, 1/D, 0)`` (preconditioners.py:18-21)."""
    D = assemble_negative_laplacian_diag(nx, ny, dx, dy)
    with np.errstate(divide="ignore"):
        return np.where(D != 0, 1.0 / D, 0.0)


# --------------------------------------------------------------------------
# Conjugate gradients: restatement of SciPy's ``cg`` (scipy/sparse/linalg/
# _isolve/iterative.py:305-422 in the installed 1.15.3; same recurrence in the
# pinned scipy==1.13.1). Textbook Hestenes-Stiefel PCG, x0 = 0, convergence
# test on the recursively updated residual at the top of each iteration.
# --------------------------------------------------------------------------
def cg(matvec: Callable[[np.ndarray], np.ndarray], b: np.ndarray, rtol: float,
       maxiter: int, psolve: Callable[[np.ndarray], np.ndarray] | None = None,
       history: list | None = None):
    """Return ``(x, info, iterations_done)``.

    ``atol = max(0, rtol*||b||)`` (``_get_atol_rtol``); ``||b|| == 0`` returns
    ``x = b``; loop ``it < maxiter``: stop with info 0 when ``||r|| < atol``;
    ``z = M r``; ``rho = r.z``; ``p = z`` (first) or ``p = beta*p + z``;
    ``q = A p``; ``alpha = rho/(p.q)``; ``x += alpha p``; ``r -= alpha q``.
    Loop exhaustion returns ``info = maxiter``.
    """
    b = np.asarray(b, dtype=float)
    bnrm2 = np.linalg.norm(b)
    atol = max(0.0, float(rtol) * float(bnrm2))
    if bnrm2 == 0:
        return b, 0, 0
    x = np.zeros_like(b)
    r = b.copy()
    p = None
    rho_prev = None
    for it in range(maxiter):
        rn = np.linalg.norm(r)
        if history is not None:
            history.append(float(rn))
        if rn < atol:
            return x, 0, it
        z = psolve(r) if psolve is not None else r
        rho = np.dot(r, z)
        if it > 0:
            beta = rho / rho_prev
            p *= beta
            p += z
        else:
            p = np.empty_like(r)
            p[:] = z[:]
        q = matvec(p)
        alpha = rho / np.dot(p, q)
        x += alpha * p
        r -= alpha * q
        rho_prev = rho
    return x, maxiter, maxiter


@dataclass
class SolveResult:
    x: np.ndarray
    converged: bool
    iterations: int          # reference semantics (interface.py:52)
    residual_norm: float
    method: str
    true_iterations: int = 0


def solve(matvec, b, tol, maxiter, psolve=None):
    """``linear_solvers.interface.solve(..., method='cg')`` (interface.py:23-53)."""
    x, info, its = cg(matvec, b, tol, maxiter, psolve)
    r = b - matvec(x)
    rnorm = float(np.linalg.norm(r))
    bnorm = float(np.linalg.norm(b)) or 1.0
    converged = (info == 0) and (rnorm <= tol * bnorm)
    iterations = info if info > 0 else 0
    return SolveResult(x=x, converged=converged, iterations=iterations,
                       residual_norm=rnorm, method="cg", true_iterations=its)


# --------------------------------------------------------------------------
# Pressure projection (reference: solvers/pressure_solver.py:42-219)
# --------------------------------------------------------------------------
def solve_pressure_poisson(u, v, p, dt, dx, dy, cfg, rhs_override=None,
                           use_jacobi=None, loops=False):
    """In-place projection on interior arrays ``u, v, p`` of shape (ny, nx).

    Returns ``(p_solution, diagnostics)`` with the reference's keys plus
    ``Rp_true_iterations``.
    """
    ny, nx = u.shape
    N = nx * ny
    if rhs_override is not None:
        mv = lambda x: pos_laplacian_dirichlet(x, nx, ny, dx, dy)  # noqa: E731
        rhs = np.copy(rhs_override).reshape(-1).astype(float)
        lin_tol = getattr(cfg, "lin_tol", 1e-10)
        lin_maxiter = getattr(cfg, "lin_maxiter", 500)
        res = solve(mv, rhs, lin_tol, lin_maxiter, None)
        psol = res.x.reshape(ny, nx)
        p[:] = psol
        div_norm = float(np.linalg.norm(mv(psol.reshape(-1)) - rhs))
        div_before = None
        div_after_raw = None
    else:
        mv = lambda x: neg_laplacian_pinned(x, nx, ny, dx, dy)  # noqa: E731
        div_u = divergence_coherent(u, v, dx, dy)
        div_before = float(np.linalg.norm(div_u))
        rhs = (-div_u / max(dt, 1e-14)).reshape(-1)
        rhs_mean = float(np.mean(rhs))
        rhs -= rhs_mean
        rhs[0] = 0.0
        lin_tol = getattr(cfg, "lin_tol", 1e-10)
        lin_maxiter = getattr(cfg, "lin_maxiter", 400)
        if use_jacobi is None:
            use_jacobi = getattr(cfg, "enable_jacobi_pc", True)
        psolve = None
        if use_jacobi:
            invD = (jacobi_inv_diag_assembled if loops else jacobi_inv_diag)(nx, ny, dx, dy)
            psolve = lambda r: invD * r  # noqa: E731
        res = solve(mv, rhs, lin_tol, lin_maxiter, psolve)
        psol = res.x.reshape(ny, nx)
        psol -= psol.flat[0]
        p[:] = psol
        gx, gy = gradient_coherent(p, dx, dy)
        u[:] -= dt * gx
        v[:] -= dt * gy
        div_after = divergence_coherent(u, v, dx, dy)
        div_after_raw = float(np.linalg.norm(div_after))
        div_norm = float(np.linalg.norm(div_after))
    pvec = psol.reshape(-1)
    lin_res = float(np.linalg.norm(mv(pvec) - rhs))
    rhs_norm = float(np.linalg.norm(rhs)) or 1.0
    rp_iters = res.iterations
    if rp_iters == 0 and res.residual_norm > lin_tol:
        rp_iters = 1
    red = None
    if div_before is not None and div_norm not in (None, 0.0):
        red = div_before / div_norm
    diag = {
        "Rp_iterations": rp_iters,
        "Rp_residual": res.residual_norm,
        "Rp_converged": res.converged,
        "Rp_method": "cg",
        "divergence_norm": div_norm,
        "lin_residual_norm": lin_res,
        "lin_residual_rel": lin_res / rhs_norm,
        "divergence_reduction_factor": red,
        "divergence_norm_initial": div_before,
        "divergence_norm_after_raw": div_after_raw,
        "divergence_norm_final": div_norm,
        "Rp_true_iterations": res.true_iterations,
        "rhs_norm": float(np.linalg.norm(rhs)),
    }
    return psol, diag


# --------------------------------------------------------------------------
# State and time step (reference: core/ghost_fields.py:47-60,
# solvers/solver.py:15-193, residuals/manager.py:2-16)
# --------------------------------------------------------------------------
@dataclass
class OracleState:
    nx: int
    ny: int
    fields: dict
    meta: dict = field(default_factory=dict)


def allocate_state(nx: int, ny: int) -> OracleState:
    return OracleState(nx, ny, {k: np.zeros((ny + 2, nx + 2)) for k in ("u", "v", "p")})


def _interior(a):
    return a[1:-1, 1:-1]


def step(cfg: Any, state: OracleState, iteration: int = 0, ru_first: float = 1.0, loops: bool = False):
    """One projection step; returns ``(residuals, diagnostics)``.

    Follows ``solver.py:15-156``: dt controller (CFL limit, diffusion limit,
    growth cap, divergence-rise safeguard), predictor ``u* = u - dt*conv``,
    projection, lid BC after projection, ``Ru = ||du||/N``, continuity
    ``||central_div(u,v;1,1)||/N``, NaN flag. ``ru_first`` is the tracker's
    first ``Ru`` value (the reference seeds it with 1.0). ``loops`` selects the
    loop-faithful advection and Jacobi assembly (same results, the
    reference's cost).
    """
    u = _interior(state.fields["u"])
    v = _interior(state.fields["v"])
    p = _interior(state.fields["p"])
    ny, nx = u.shape
    lx = getattr(cfg, "lx", float(nx - 1))
    ly = getattr(cfg, "ly", float(ny - 1))
    dx = lx / max(nx - 1, 1)
    dy = ly / max(ny - 1, 1)
    umax = float(np.max(np.abs(u))) + 1e-12
    vmax = float(np.max(np.abs(v))) + 1e-12
    cfl_target = getattr(cfg, "cfl_target", 0.5)
    cfl_growth = getattr(cfg, "cfl_growth", 1.05)
    Re = float(getattr(cfg, "Re", 100.0))
    nu = 1.0 / Re
    dt_prev = float(state.meta.get("dt_prev", 1e-2))
    dt_cfl = min(cfl_target * dx / umax, cfl_target * dy / vmax)
    dt_diff = 0.25 * min(dx * dx, dy * dy) / max(nu, 1e-12)
    dt_cand = min(dt_cfl, dt_diff)
    dt = dt_prev * cfl_growth if dt_cand > dt_prev * cfl_growth else dt_cand
    div0 = float(np.linalg.norm(divergence_central(u, v, 1.0, 1.0)))
    hist = list(state.meta.get("divergence_history", []))
    hist.append(div0)
    hist = hist[-5:]
    state.meta["divergence_history"] = hist
    safeguard = False
    if len(hist) >= 3:
        a, b, c = hist[-3], hist[-2], hist[-1]
        if a > 0 and b > a * 1.01 and c > b * 1.01:
            dt *= 0.5
            safeguard = True
    state.meta["dt_prev"] = dt
    state.meta["dx"] = dx
    state.meta["dy"] = dy
    state.meta["nu"] = nu
    cfl_now = max(umax * dt / dx, vmax * dt / dy)
    u_old = u.copy()
    v_old = v.copy()
    if getattr(cfg, "disable_advection", False):
        us, vs = u_old.copy(), v_old.copy()
    else:
        scheme = getattr(cfg, "advection_scheme", "upwind")
        if loops:
            adv = advect_quick_loops if scheme == "quick" else advect_upwind_loops
        else:
            adv = advect_quick if scheme == "quick" else advect_upwind
        us = u_old - dt * adv(u_old, v_old, u_old, dx, dy)
        vs = v_old - dt * adv(u_old, v_old, v_old, dx, dy)
    u[:, :] = us
    v[:, :] = vs
    _, pdiag = solve_pressure_poisson(u, v, p, dt, dx, dy, cfg, loops=loops)
    lid = getattr(cfg, "lid_velocity", None)
    if lid is not None:
        state.fields["u"][-2, 1:-1] = lid
    N = u.size

    def l2(d):
        d = d.ravel()
        return float(np.sqrt(np.sum(d * d)) / max(1, d.size))

    Ru = l2(u - u_old)
    Rv = l2(v - v_old)
    continuity = float(np.linalg.norm(divergence_central(u, v, 1.0, 1.0)) / max(1.0, nx * ny))
    nan = not (np.isfinite(u).all() and np.isfinite(v).all() and np.isfinite(p).all())
    residuals = {"Ru": Ru, "Rv": Rv, "Rp": 0.0, "continuity": continuity}
    drop = 0.0
    if Ru > 0 and ru_first > 0:
        drop = math.log10(ru_first) - math.log10(Ru)
    diagnostics = {
        "iteration": iteration,
        "dt": dt,
        "CFL": cfl_now,
        "plateau_Ru": False,
        "max_residual": max(residuals.values()),
        "residual_drop_Ru_orders": max(0.0, drop),
        "nu": nu,
        "Rp_iterations": pdiag["Rp_iterations"],
        "Rp_residual": pdiag["Rp_residual"],
        "Rp_converged": pdiag["Rp_converged"],
        "Rp_method": "cg",
        "divergence_norm": pdiag["divergence_norm"],
        "nan_detected": nan,
        # extras (not in the reference's dict)
        "Rp_true_iterations": pdiag["Rp_true_iterations"],
        "divergence_norm_initial": pdiag["divergence_norm_initial"],
        "Rp_rhs_norm": pdiag["rhs_norm"],
        "safeguard": safeguard,
        "N": N,
    }
    return residuals, diagnostics


def run_steps(cfg: Any, state: OracleState, nsteps: int, start: int = 0):
    """Drive ``step`` like ``SimulationDriver.run`` (drivers/simulation_driver.py:76-110)."""
    out = []
    for it in range(start, start + nsteps):
        res, diag = step(cfg, state, it)
        out.append((res, diag))
    return out
