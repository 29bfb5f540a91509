"""Projection time step (reference: solvers/solver.py:15-193).

``step`` keeps the reference signature and return values; the numerics run on
the GPU as one device phase machine (dt controller, predictor, pinned-Neumann
Jacobi-PCG pressure solve, correction, lid BC, residuals), see
``pyflow-solver_amd/csrc/pf_core.hip``. The host side only builds the
reference's dictionaries, prints, logs and writes the emergency checkpoint.
"""
from __future__ import annotations

import math
from typing import Any

from .. import _hip as H
from .. import _device as D
from .._device import cavity_params
from .._mirror import config_precision, mirror_for
from ..core.ghost_fields import State
from ..logging.structured import JsonlLogger

__all__ = ["step", "step_params"]


def _grid(state: State, config):
    u = state._raw("u")
    ny, nx = u.shape[0] - 2, u.shape[1] - 2
    lx = getattr(config, "lx", float(nx - 1))
    ly = getattr(config, "ly", float(ny - 1))
    dx = lx / max(nx - 1, 1)
    dy = ly / max(ny - 1, 1)
    return nx, ny, dx, dy


def step_params(config, nx: int, ny: int, dx: float, dy: float, mode: int = H.MODE_STEP) -> H.CavityParams:
    """Device parameters from the attributes ``step`` reads (solver.py:16-97, pressure_solver.py:134-139).

    Optional attributes a config may add (the reference has none of them):
    ``precision`` ("fp64" / "mixed"), ``inner_rtol``, and the tiling
    ``tile_rows``, ``max_workgroups``, ``pcg_per_round`` (0 or absent: the
    library's default). A state that already ran keeps the tiling it first
    ran with (``state.meta['kernel_config']``)."""
    p = cavity_params(
        nx, ny, dx=dx, dy=dy,
        scheme=getattr(config, "advection_scheme", "upwind"),
        disable_advection=bool(getattr(config, "disable_advection", False)),
        lid_velocity=getattr(config, "lid_velocity", None),
        Re=float(getattr(config, "Re", 100.0)),
        cfl_target=getattr(config, "cfl_target", 0.5),
        cfl_growth=getattr(config, "cfl_growth", 1.05),
        lin_tol=getattr(config, "lin_tol", 1e-10),
        lin_maxiter=getattr(config, "lin_maxiter", 400),
        jacobi=bool(getattr(config, "enable_jacobi_pc", True)),
        mode=mode,
        inner_rtol=float(getattr(config, "inner_rtol", D.DEFAULT_INNER_RTOL)),
    )
    for k in ("tile_rows", "max_workgroups", "pcg_per_round"):
        setattr(p, k, int(getattr(config, k, 0) or 0))
    return p


def step(config, state: State, tracker, iteration: int) -> tuple[State, dict[str, float], dict[str, Any]]:
    diagnostics_enabled = getattr(config, "diagnostics", True)
    nx, ny, dx, dy = _grid(state, config)
    if nx < 2 or ny < 2:
        raise ValueError("step needs at least 2x2 interior cells")
    if diagnostics_enabled:
        print(f"[Step {iteration}] Grid: nx={nx}, ny={ny}, dx={dx:.4g}, dy={dy:.4g}, Re={getattr(config, 'Re', 100.0)}")
    Re = float(getattr(config, "Re", 100.0))
    nu = 1.0 / Re
    params = step_params(config, nx, ny, dx, dy)
    mirror = mirror_for(state, params, config_precision(config, state))
    mirror.before(params)
    hist = state.meta.get("divergence_history", [])
    mirror.cav.set_meta(float(state.meta.get("dt_prev", 1e-2)), list(hist)[-5:])
    d = mirror.cav.run(1)
    mirror.after()

    dt = d.dt
    hist.append(d.div_before)
    hist = hist[-5:]
    state.meta["divergence_history"] = hist
    if d.safeguard and diagnostics_enabled:
        a, b, c = hist[-3], hist[-2], hist[-1]
        print(f"[Step {iteration}] dt safeguard triggered (divergence rising {a:.3g}->{b:.3g}->{c:.3g}); new dt={dt:.3g}")
    state.meta["dt_prev"] = dt
    state.meta["dx"] = dx
    state.meta["dy"] = dy
    state.meta["nu"] = nu
    current_cfl = d.cfl
    if diagnostics_enabled:
        print(f"[Step {iteration}] dt={dt:.3g}, umax={d.umax:.3g}, vmax={d.vmax:.3g}, CFL={current_cfl:.3g}")
        print(f"[Step {iteration}] Initial divergence norm: {d.div_before:.3g}")
        scheme = "(disabled)" if getattr(config, "disable_advection", False) else getattr(config, "advection_scheme", "upwind")
        print(f"[Step {iteration}] Advection scheme: {scheme}")

    lin_tol = getattr(config, "lin_tol", 1e-10)
    rp_iters = d.rp_info if d.rp_info > 0 else 0
    if rp_iters == 0 and d.rp_residual > lin_tol:
        rp_iters = 1
    pressure_diag = {
        "Rp_iterations": rp_iters,
        "Rp_residual": d.rp_residual,
        "Rp_converged": bool(d.rp_converged),
        "Rp_method": "cg",
        "divergence_norm": d.div_norm,
    }
    if diagnostics_enabled:
        print(f"[Step {iteration}] Pressure solve: iterations={rp_iters}, residual={d.rp_residual:.3g}, converged={bool(d.rp_converged)}")
    state.meta["prev_div_norm"] = d.div_norm
    Ru, Rv, Rp = d.ru, d.rv, 0.0
    continuity = d.continuity
    nan_detected = bool(d.nan_detected)
    if diagnostics_enabled:
        print(f"[Step {iteration}] Residuals: Ru={Ru:.3g}, Rv={Rv:.3g}, continuity={continuity:.3g}")
    tracker.add("Ru", Ru)
    tracker.add("Rv", Rv)
    tracker.add("Rp", Rp)
    tracker.add("continuity", continuity)
    tracker.add("Rp_iterations", rp_iters)
    tracker.add("Rp_residual", d.rp_residual)
    residuals = {"Ru": Ru, "Rv": Rv, "Rp": Rp, "continuity": continuity}
    series = getattr(tracker, "series", {})
    drop = 0.0
    if "Ru" in series and series["Ru"].values and residuals["Ru"] > 0 and series["Ru"].values[0] > 0:
        drop = math.log10(series["Ru"].values[0]) - math.log10(residuals["Ru"])
    diagnostics = {
        "iteration": iteration,
        "dt": dt,
        "CFL": current_cfl,
        "plateau_Ru": tracker.plateau_detect("Ru", window=10, threshold=-0.005),
        "max_residual": max(residuals.values()),
        "residual_drop_Ru_orders": max(0.0, drop),
        "nu": nu,
        "Rp_iterations": rp_iters,
        "Rp_residual": d.rp_residual,
        "Rp_converged": bool(d.rp_converged),
        "Rp_method": "cg",
        "divergence_norm": d.div_norm,
        "nan_detected": nan_detected,
        "Rp_true_iterations": int(d.rp_iterations),
    }
    if getattr(config, "seed", None) is not None:
        diagnostics["seed"] = config.seed
    if getattr(config, "assert_invariants", False):
        prev_div = pressure_diag.get("divergence_norm")
        if prev_div is None:
            raise AssertionError("Invariant failed: divergence_norm missing from pressure diagnostics")
        assert continuity <= max(1e3, prev_div * 10)
        mirror.ensure_host()
        assert abs(state._raw("p")[1:-1, 1:-1].flat[0]) < 1e-10, "Invariant failed: reference pressure cell not ~0"
    log_path = getattr(config, "log_path", None)
    log_stream = getattr(config, "log_stream", None)
    if log_path or log_stream:
        logger = getattr(config, "_logger_instance", None)
        if logger is None:
            logger = JsonlLogger(path=log_path, stream=log_stream)
            config._logger_instance = logger
        logger.log({"type": "step", "it": iteration, "dt": dt, "CFL": current_cfl, "Ru": Ru, "Rv": Rv,
                    "continuity": continuity, "Rp_it": rp_iters, "nan": nan_detected})
        if nan_detected:
            logger.log({"type": "error", "it": iteration, "reason": "nan_detected"})
            try:
                from ..io.checkpoint import save_checkpoint
                ck_path = getattr(config, "emergency_checkpoint_path", None)
                if ck_path:
                    base, ext = (ck_path.rsplit(".", 1) + ["npz"])[:2]
                    fail_path = f"{base}_FAIL_NAN.{ext}"
                else:
                    fail_path = f"emergency_FAIL_NAN_it{iteration}.npz"
                save_checkpoint(fail_path, state, iteration, diagnostics.get("time", diagnostics.get("wall_time", 0.0)), config)
                logger.log({"type": "emergency_checkpoint", "path": fail_path, "it": iteration})
            except Exception:
                logger.log({"type": "emergency_checkpoint", "status": "failed", "it": iteration})
    return state, residuals, diagnostics
