"""Continuous-batching server for lid-driven-cavity requests.

Requests arrive at any time. The server keeps one device batch running and,
between chunks of rounds:

* admits waiting requests while a slot and device memory are free: a
  request's footprint is its workspace (``pf_batch_workspace_bytes``: every
  field, PCG vector and partial-sum buffer of the cavity), and the footprints
  of the admitted cavities never exceed the server's HBM budget. Waiting
  requests are considered in arrival order; one that does not fit waits while
  a later, smaller one may start (first fit);
* arms each admitted cavity for its own number of steps next to the cavities
  already running (``pf_batch_add`` + ``pf_batch_start``); their fields stay
  resident on the device for the whole request;
* advances every running cavity by one chunk of rounds (``pf_batch_advance``),
  so the work of all admitted requests overlaps on the device;
* retires the cavities that finished: their fields and diagnostics are
  copied to the host and their workspace is released.

A cavity's results do not depend on what else runs in the batch: every
reduction folds its own workgroups' partials in a fixed order, so a request
gives the same bits alone and in any mix.

The reference has no server: its ``SimulationDriver.run``
(``drivers/simulation_driver.py:58-110``) advances one cavity on the host. A
request here is that driver's loop for ``steps`` steps from rest, on the
device, with the reference's step semantics (``solvers/solver.py:15-193``).
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _hip as H
from ._device import DEFAULT_INNER_RTOL, DeviceBatch, cavity_params, precision_code

__all__ = ["CavityRequest", "CavityResult", "CavityServer", "footprint", "first_fit"]


@dataclass
class CavityRequest:
    """One lid-driven cavity to advance ``steps`` projection steps from rest
    (or from ``fields``, ghost-layout (ny+2, nx+2) float64 arrays u, v, p)."""
    nx: int
    ny: int
    steps: int
    Re: float = 100.0
    lid_velocity: float = 1.0
    scheme: str = "quick"
    dx: float = 1.0
    dy: float = 1.0
    cfl_target: float = 0.5
    cfl_growth: float = 1.05
    lin_tol: float = 1e-6
    lin_maxiter: int = 10000
    fields: tuple | None = None

    def params(self, inner_rtol: float = DEFAULT_INNER_RTOL) -> H.CavityParams:
        return cavity_params(self.nx, self.ny, dx=self.dx, dy=self.dy, scheme=self.scheme,
                             lid_velocity=self.lid_velocity, Re=self.Re, cfl_target=self.cfl_target,
                             cfl_growth=self.cfl_growth, lin_tol=self.lin_tol, lin_maxiter=self.lin_maxiter,
                             jacobi=True, inner_rtol=inner_rtol)


@dataclass
class CavityResult:
    ticket: int
    request: CavityRequest
    u: np.ndarray
    v: np.ndarray
    p: np.ndarray
    diag: H.StepDiag
    t_submit: float
    t_admit: float
    t_done: float

    @property
    def latency(self) -> float:
        """Seconds from submission to completion."""
        return self.t_done - self.t_submit


def footprint(req: CavityRequest, precision: str | int = "mixed", inner_rtol: float = DEFAULT_INNER_RTOL) -> int:
    """Device bytes a request holds while admitted: its workspace plus the
    256-byte alignment slack of its allocation. Needs no GPU."""
    import ctypes as C
    p = req.params(inner_rtol)
    n = C.c_size_t()
    H.check(H.lib().pf_batch_workspace_bytes(precision_code(precision), 1, C.byref(p), C.byref(n)),
            "pf_batch_workspace_bytes")
    return int(n.value) + 256


def first_fit(waiting: Sequence[int], free_bytes: int, free_slots: int) -> list[int]:
    """Indices into ``waiting`` (footprints in arrival order) admitted now:
    each in turn if it fits what is still free."""
    out = []
    for k, b in enumerate(waiting):
        if free_slots <= 0:
            break
        if b <= free_bytes:
            out.append(k)
            free_bytes -= b
            free_slots -= 1
    return out


@dataclass
class _Entry:
    ticket: int
    request: CavityRequest
    bytes: int
    t_submit: float
    t_admit: float = 0.0
    slot: int = -1


class CavityServer:
    """Continuous batching of cavity requests on one GPU.

    ``hbm_budget``: device bytes the admitted cavities may hold together
    (default: 80% of the free device memory at creation). ``max_cavities``:
    slots of the batch. ``chunk_rounds``: rounds queued per scheduling step.
    """

    def __init__(self, precision: str | int = "mixed", *, hbm_budget: int | None = None, max_cavities: int = 64,
                 chunk_rounds: int = 64, inner_rtol: float = DEFAULT_INNER_RTOL, clock=time.perf_counter):
        torch = H.require_gpu()
        self.prec = precision_code(precision)
        self.inner_rtol = inner_rtol
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info()
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.max_cavities = int(max_cavities)
        self.chunk_rounds = int(chunk_rounds)
        self.clock = clock
        self.batch = DeviceBatch(None, self.prec, max_cav=self.max_cavities, max_tiles=1 << 22)
        self._next_ticket = 0
        self._waiting: list[_Entry] = []
        self._running: dict[int, _Entry] = {}  # slot -> entry
        self.in_use_bytes = 0
        self.peak_bytes = 0
        self.peak_running = 0

    # -- queue ----------------------------------------------------------------
    def submit(self, req: CavityRequest, t_submit: float | None = None) -> int:
        """Queue a request; returns its ticket. A request whose footprint
        exceeds the whole budget could never run and is refused."""
        b = footprint(req, self.prec, self.inner_rtol)
        if b > self.hbm_budget:
            raise ValueError(f"request {req.nx}x{req.ny} needs {b} device bytes, more than the budget "
                             f"{self.hbm_budget}")
        if req.steps <= 0:
            raise ValueError("steps must be > 0")
        t = self._next_ticket
        self._next_ticket += 1
        self._waiting.append(_Entry(t, req, b, self.clock() if t_submit is None else t_submit))
        return t

    @property
    def waiting(self) -> int:
        return len(self._waiting)

    @property
    def running(self) -> int:
        return len(self._running)

    def idle(self) -> bool:
        return not self._waiting and not self._running

    # -- scheduling -------------------------------------------------------------
    def _admit(self) -> None:
        free_slots = self.max_cavities - len(self._running)
        picks = first_fit([e.bytes for e in self._waiting], self.hbm_budget - self.in_use_bytes, free_slots)
        if not picks:
            return
        now = self.clock()
        chosen = [self._waiting[k] for k in picks]
        taken = set(picks)
        self._waiting = [e for k, e in enumerate(self._waiting) if k not in taken]
        for e in chosen:
            slot = self.batch.add(e.request.params(self.inner_rtol))
            if e.request.fields is not None:
                u, v, p = e.request.fields
                self.batch.upload(slot, u, v, p)
            self.batch.start(slot, int(e.request.steps))
            e.slot, e.t_admit = slot, now
            self._running[slot] = e
            self.in_use_bytes += e.bytes
        self.peak_bytes = max(self.peak_bytes, self.in_use_bytes)
        self.peak_running = max(self.peak_running, len(self._running))

    def _retire(self) -> list[CavityResult]:
        out = []
        finished = self.batch.finished()
        if not finished:
            return out
        diags = self.batch.diag()
        now = self.clock()
        for slot in finished:
            e = self._running.pop(slot)
            r = e.request
            u = np.zeros((r.ny + 2, r.nx + 2))
            v = np.zeros_like(u)
            p = np.zeros_like(u)
            self.batch.download(slot, u, v, p)
            d = diags[slot]
            self.batch.remove(slot)
            self.in_use_bytes -= e.bytes
            out.append(CavityResult(e.ticket, r, u, v, p, d, e.t_submit, e.t_admit, now))
        return out

    def pump(self) -> list[CavityResult]:
        """One scheduling step: admit what fits, advance the running cavities
        by one chunk of rounds, retire those that finished."""
        self._admit()
        if not self._running:
            return []
        self.batch.advance(self.chunk_rounds)
        while not self.batch.poll():
            time.sleep(0.0001)
        return self._retire()

    def run_until_idle(self) -> list[CavityResult]:
        out = []
        while not self.idle():
            out.extend(self.pump())
        return out

    def close(self) -> None:
        self.batch.close()
