"""Seeded synthetic workloads of the batched and served configurations.

* ``batch_params``: independent lid-driven cavities with square sizes
  ``n_i = 32 * U{1..16}`` and Reynolds numbers log-uniform in [100, 5000]
  (256 cavities with seed 0 is the batched configuration), QUICK, PCG tol
  1e-6, lin_maxiter 10000, started from rest, dx = dy = 1.
* ``serve_requests``: cavity requests with ``n ~ round(exp(U(ln 64, ln 2048)))``,
  Re log-uniform in [100, 1e4] and ``U{100..2000}`` steps (times a scale),
  and ``poisson_arrivals``: exponential inter-arrival times at a given rate.

The reference runs one cavity per process from its CLI
(``drivers/simulation_driver.py:58-110``); every cavity here follows that
step (``solvers/solver.py:15-193``) with the CLI's grid convention lx = nx - 1.
"""
from __future__ import annotations

import numpy as np

from ._device import DEFAULT_INNER_RTOL, cavity_params

__all__ = ["batch_sizes", "batch_params", "serve_requests", "poisson_arrivals"]


def batch_sizes(count: int = 256, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """(n_i, Re_i) of the batched configuration."""
    rng = np.random.default_rng(seed)
    ns = 32 * rng.integers(1, 17, size=count)
    res = np.exp(rng.uniform(np.log(100.0), np.log(5000.0), size=count))
    return ns.astype(np.int64), res


def batch_params(count: int = 256, seed: int = 0, inner_rtol: float = DEFAULT_INNER_RTOL):
    """Device parameters of the batched configuration and the sizes n_i."""
    ns, res = batch_sizes(count, seed)
    params = [cavity_params(int(n), int(n), dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0, Re=float(re),
                            cfl_target=0.5, cfl_growth=1.05, lin_tol=1e-6, lin_maxiter=10000, jacobi=True,
                            inner_rtol=inner_rtol) for n, re in zip(ns, res)]
    return params, ns


def serve_requests(count: int, seed: int = 0, step_scale: float = 1.0):
    """(n, Re, steps) arrays of the served configuration; steps are
    ``U{100..2000}`` times ``step_scale`` (at least 1)."""
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(np.log(64.0), np.log(2048.0), size=count))).astype(int)
    res = np.exp(rng.uniform(np.log(100.0), np.log(1e4), size=count))
    steps = np.maximum(1, np.rint(rng.integers(100, 2001, size=count) * step_scale)).astype(int)
    return ns, res, steps


def poisson_arrivals(count: int, rate_per_s: float, seed: int = 1) -> np.ndarray:
    """Arrival times (s) of a Poisson process of the given rate."""
    return np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate_per_s, size=count))
