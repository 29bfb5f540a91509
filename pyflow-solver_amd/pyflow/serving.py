"""Continuous-batching server for lid-driven-cavity requests.

Requests arrive at any time. The server keeps one device batch running and,
between chunks of rounds:

* admits waiting requests while a slot and device memory are free: a
  request's footprint is its workspace (``pf_batch_workspace_bytes``: every
  field, PCG vector and partial-sum buffer of the cavity). The workspaces are
  carved from one device allocation of exactly the server's HBM budget (an
  arena with first-fit placement), so the server never holds more device
  memory than its budget, whatever the caching allocator keeps. Waiting
  requests are considered in arrival order; one that does not fit waits while
  a later, smaller one may start (first fit);
* arms each admitted cavity for its own number of steps next to the cavities
  already running (``pf_batch_add`` + ``pf_batch_start``); their fields stay
  resident on the device for the whole request;
* advances every running cavity by one chunk of rounds (``pf_batch_advance``),
  so the work of all admitted requests overlaps on the device;
* retires the cavities that finished: their fields and diagnostics are
  copied to the host and their workspace is released;
* evicts a cavity whose fields turned non-finite (the step's NaN flag,
  ``solvers/solver.py:116-121``, or a NaN/Inf velocity maximum at a step's
  start) and reports it with ``status == "nan_detected"``: in the reference's
  semantics such a cavity would keep running every PCG solve to lin_maxiter.
  Its neighbours are not disturbed.

A cavity's results do not depend on what else runs in the batch: every
reduction folds its own workgroups' partials in a fixed order, so a request
gives the same bits alone and in any mix.

The reference has no server: its ``SimulationDriver.run``
(``drivers/simulation_driver.py:58-110``) advances one cavity on the host. A
request here is that driver's loop for ``steps`` steps from rest, on the
device, with the reference's step semantics (``solvers/solver.py:15-193``).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
import numpy as np

from . import _hip as H
from ._device import DEFAULT_INNER_RTOL, DeviceBatch, cavity_params, precision_code

__all__ = ["Arena", "CavityRequest", "CavityResult", "CavityServer", "footprint"]


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
    """``status``: "done" (every step ran) or "nan_detected" (evicted with
    non-finite fields after ``diag.steps_done`` steps)."""
    ticket: int
    request: CavityRequest
    u: np.ndarray
    v: np.ndarray
    p: np.ndarray
    diag: H.StepDiag
    t_submit: float
    t_admit: float
    t_done: float
    status: str = "done"

    @property
    def latency(self) -> float:
        """Seconds from submission to completion."""
        return self.t_done - self.t_submit


ALIGN = 256


def footprint(req: CavityRequest, precision: str | int = "mixed", inner_rtol: float = DEFAULT_INNER_RTOL) -> int:
    """Device bytes a request holds while admitted: its workspace, rounded up
    to the arena's 256-byte placement. Needs no GPU."""
    import ctypes as C
    p = req.params(inner_rtol)
    n = C.c_size_t()
    H.check(H.lib().pf_batch_workspace_bytes(precision_code(precision), 1, C.byref(p), C.byref(n)),
            "pf_batch_workspace_bytes")
    return -(-int(n.value) // ALIGN) * ALIGN


class Arena:
    """First-fit placement of workspaces inside one region of ``size`` bytes
    (offsets and lengths in multiples of 256; freed neighbours merge). Pure
    host bookkeeping: the server backs it with one device allocation."""

    def __init__(self, size: int):
        self.size = int(size) // ALIGN * ALIGN
        self._free: list[tuple[int, int]] = [(0, self.size)] if self.size > 0 else []  # (offset, length), sorted

    def alloc(self, nbytes: int) -> int | None:
        n = -(-int(nbytes) // ALIGN) * ALIGN
        for k, (off, length) in enumerate(self._free):
            if length >= n:
                if length == n:
                    del self._free[k]
                else:
                    self._free[k] = (off + n, length - n)
                return off
        return None

    def release(self, off: int, nbytes: int) -> None:
        n = -(-int(nbytes) // ALIGN) * ALIGN
        blocks = sorted(self._free + [(off, n)])
        merged: list[tuple[int, int]] = []
        for o, length in blocks:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + length)
            else:
                if merged and merged[-1][0] + merged[-1][1] > o:
                    raise ValueError("arena: block released twice or overlapping")
                merged.append((o, length))
        self._free = merged

    @property
    def free_bytes(self) -> int:
        return sum(length for _, length in self._free)

    @property
    def largest_free(self) -> int:
        return max((length for _, length in self._free), default=0)


@dataclass
class _Entry:
    ticket: int
    request: CavityRequest
    bytes: int
    t_submit: float
    t_admit: float = 0.0
    slot: int = -1
    offset: int = -1


class CavityServer:
    """Continuous batching of cavity requests on one GPU.

    ``hbm_budget``: device bytes the admitted cavities may hold together
    (default: 80% of the free device memory at creation); the server
    allocates exactly this much once, as the arena of every workspace.
    ``max_cavities``: slots of the batch. ``chunk_rounds``: rounds queued per
    scheduling step.
    """

    def __init__(self, precision: str | int = "mixed", *, hbm_budget: int | None = None, max_cavities: int = 64,
                 chunk_rounds: int = 64, inner_rtol: float = DEFAULT_INNER_RTOL, clock=time.perf_counter):
        torch = H.require_gpu()
        self.prec = precision_code(precision)
        self.inner_rtol = inner_rtol
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info()
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget) // ALIGN * ALIGN
        device = torch.device("cuda", torch.cuda.current_device())
        raw = torch.empty(self.hbm_budget + ALIGN, dtype=torch.uint8, device=device)
        skip = (-raw.data_ptr()) % ALIGN
        self._arena_mem = raw[skip:skip + self.hbm_budget]
        self.arena = Arena(self.hbm_budget)
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
        self.evicted = 0

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
        """First fit in arrival order: each waiting request in turn is placed
        in the arena if a free block holds it and a slot is free."""
        if not self._waiting or len(self._running) >= self.max_cavities:
            return
        now = self.clock()
        still = []
        for e in self._waiting:
            off = self.arena.alloc(e.bytes) if len(self._running) < self.max_cavities else None
            if off is None:
                still.append(e)
                continue
            ws = self._arena_mem[off:off + e.bytes]
            slot = self.batch.add(e.request.params(self.inner_rtol), workspace=ws)
            if e.request.fields is not None:
                u, v, p = e.request.fields
                self.batch.upload(slot, u, v, p)
            self.batch.start(slot, int(e.request.steps))
            e.slot, e.offset, e.t_admit = slot, off, now
            self._running[slot] = e
            self.in_use_bytes += e.bytes
        self._waiting = still
        self.peak_bytes = max(self.peak_bytes, self.in_use_bytes)
        self.peak_running = max(self.peak_running, len(self._running))

    @staticmethod
    def _nonfinite(d: H.StepDiag) -> bool:
        return bool(d.nan_detected) or not (math.isfinite(d.umax) and math.isfinite(d.vmax))

    def _retire(self) -> list[CavityResult]:
        out = []
        finished = set(self.batch.finished())  # also refreshes every slot's diagnostics
        diags = self.batch.diag()
        evict = {s for s in self._running if s not in finished and self._nonfinite(diags[s])}
        if not finished and not evict:
            return out
        now = self.clock()
        for slot in sorted(finished | evict):
            e = self._running.pop(slot)
            r = e.request
            u = np.zeros((r.ny + 2, r.nx + 2))
            v = np.zeros_like(u)
            p = np.zeros_like(u)
            self.batch.download(slot, u, v, p)
            d = diags[slot]
            self.batch.remove(slot)
            self.arena.release(e.offset, e.bytes)
            self.in_use_bytes -= e.bytes
            status = "done" if slot in finished and not d.nan_detected else "nan_detected"
            if slot in evict:
                self.evicted += 1
            out.append(CavityResult(e.ticket, r, u, v, p, d, e.t_submit, e.t_admit, now, status))
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
        self._arena_mem = None
