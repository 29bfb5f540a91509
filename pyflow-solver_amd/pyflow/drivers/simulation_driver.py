"""Simulation driver (reference: drivers/simulation_driver.py:38-110).

``SimulationDriver.run`` is the same lazy generator: one ``step`` per resume,
yielding the same ``State`` object with ``(residuals, diagnostics)``,
``u_centerline`` and ``wall_time`` added, and a checkpoint whenever
``iteration % checkpoint_interval == 0``. Fields stay on the GPU between
steps; the centerline is gathered from the device (ny values) instead of
downloading the whole field.
"""
from __future__ import annotations

import time
from collections.abc import Callable, Generator
from typing import Any

from ..core.ghost_fields import State, allocate_state
from ..io.checkpoint import save_checkpoint
from ..residuals.manager import ResidualManager
from ..solvers.solver import step

__all__ = ["SimulationDriver"]


class SimulationDriver:
    def __init__(self, config: Any, state: State | None = None, tracker: ResidualManager | None = None, *,
                 allocate: Callable[[int, int], State] = allocate_state):
        self.config = config
        self.state = state
        self.tracker = tracker or ResidualManager()
        self.allocate_fn = allocate
        self._t0 = None
        self.iteration = 0

    def ensure_state(self, nx: int, ny: int) -> State:
        if self.state is None:
            self.state = self.allocate_fn(nx, ny)
        return self.state

    @staticmethod
    def _centerline(state: State):
        dev = getattr(state, "_dev", None)
        if dev is not None and not dev.host_valid:
            return dev.cav.centerline()
        u = state.fields["u"][1:-1, 1:-1]
        return u[:, u.shape[1] // 2]

    def run(self, max_steps: int | None = None, *, start_iteration: int = 0, progress: bool = False,
            checkpoint_path: str | None = None,
            checkpoint_interval: int | None = None) -> Generator[tuple[State, dict[str, float], dict[str, Any]], None, None]:
        """Run the simulation, yielding after every solver step."""
        self.iteration = start_iteration
        self._t0 = time.time()
        while True:
            if max_steps is not None and self.iteration >= max_steps:
                return
            if self.state is None:
                nx = getattr(self.config, "nx", None)
                ny = getattr(self.config, "ny", None)
                if nx is None or ny is None:
                    raise ValueError("SimulationDriver requires either an initial State or config.nx & config.ny for allocation")
                self.state = self.allocate_fn(nx, ny)
            s = self.state
            state, residuals, diagnostics = step(self.config, s, self.tracker, self.iteration)
            try:
                u_center = self._centerline(state)
                nyi = len(u_center)
                ly_val = getattr(self.config, "ly", float(nyi - 1))
                y_coords = [j * (ly_val / (nyi - 1)) for j in range(nyi)] if nyi > 1 else [0.0]
                diagnostics["u_centerline"] = [y_coords, [float(x) for x in u_center]]
            except Exception:  # pragma: no cover
                pass
            diagnostics["wall_time"] = time.time() - self._t0
            if progress:
                print(f"iter={self.iteration} dt={diagnostics.get('dt'):.3g} CFL={diagnostics.get('CFL'):.3g} continuity={residuals.get('continuity'):.3e}")
            yield state, residuals, diagnostics
            if checkpoint_path and checkpoint_interval and checkpoint_interval > 0:
                if (self.iteration % checkpoint_interval) == 0:
                    try:
                        save_checkpoint(checkpoint_path, state, self.iteration,
                                        diagnostics.get("time", diagnostics.get("wall_time", 0.0)), self.config)
                        diagnostics["checkpoint_written"] = True
                    except Exception:  # pragma: no cover
                        diagnostics["checkpoint_written"] = False
            self.iteration += 1
