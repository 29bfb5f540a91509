"""Device-resident cavities on top of the C ABI.

``DeviceBatch`` owns one ``pf_batch`` plus the torch byte tensors that back
each cavity's workspace (torch is used only as the device allocator and for
the current stream). ``DeviceCavity`` is a batch of one, used behind the
reference-compatible ``State``/``step``/``solve`` API.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Any, Sequence

import numpy as np

from . import _hip as H

__all__ = ["precision_code", "cavity_params", "DeviceBatch", "DeviceCavity", "default_precision"]


def default_precision() -> str:
    return os.environ.get("PYFLOW_PRECISION", "fp64").lower()


def precision_code(precision: str | int | None) -> int:
    if precision is None:
        precision = default_precision()
    if isinstance(precision, int):
        return precision
    p = str(precision).lower()
    if p in ("fp64", "float64", "double", "f64"):
        return H.PREC_FP64
    if p in ("mixed", "fp32", "mixed32", "f32"):
        return H.PREC_MIXED
    raise ValueError(f"unknown precision {precision!r} (use 'fp64' or 'mixed')")


def _inv_diag(dx: float, dy: float) -> float:
    # preconditioners.py:18-21 on the diagonal assembled at pressure_solver.py:35
    D = 2.0 / dx ** 2 + 2.0 / dy ** 2
    return 1.0 / D if D != 0 else 0.0


# Mixed precision: an FP32 inner PCG pass stops once its residual fell by this
# factor below the last FP64 true residual. Each pass restarts the Krylov
# space; 1e-2 keeps the passes short enough that the FP32 residual still
# tracks the true one (three passes for a 1e-6 solve).
DEFAULT_INNER_RTOL = 1e-2


def cavity_params(nx: int, ny: int, *, dx: float, dy: float, scheme: str | int = "upwind",
                  disable_advection: bool = False, lid_velocity: float | None = None,
                  Re: float = 100.0, cfl_target: float = 0.5, cfl_growth: float = 1.05,
                  lin_tol: float = 1e-10, lin_maxiter: int = 400, jacobi: bool = True,
                  op: int = H.OP_NEG_LAP_PINNED, mode: int = H.MODE_STEP,
                  inner_rtol: float = DEFAULT_INNER_RTOL, deflate: int = 0) -> H.CavityParams:
    p = H.CavityParams()
    p.nx, p.ny = int(nx), int(ny)
    if disable_advection:
        p.scheme = H.SCHEME_NONE
    elif isinstance(scheme, int):
        p.scheme = scheme
    else:
        p.scheme = H.SCHEME_QUICK if scheme == "quick" else H.SCHEME_UPWIND
    p.lid_enabled = 0 if lid_velocity is None else 1
    p.lid_velocity = 0.0 if lid_velocity is None else float(lid_velocity)
    p.jacobi = 1 if jacobi else 0
    p.lin_maxiter = int(lin_maxiter)
    p.op = int(op)
    p.mode = int(mode)
    p.dx, p.dy = float(dx), float(dy)
    p.nu = 1.0 / float(Re)
    p.cfl_target, p.cfl_growth = float(cfl_target), float(cfl_growth)
    p.lin_tol = float(lin_tol)
    p.inv_diag = _inv_diag(float(dx), float(dy))
    p.inner_rtol = float(inner_rtol)
    p.deflate = int(deflate)
    return p


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)


class DeviceBatch:
    """A set of cavities advanced together on one GPU (one launch per round)."""

    def __init__(self, params: Sequence[H.CavityParams] | None = None, precision: str | int | None = None,
                 *, max_cav: int | None = None, max_tiles: int | None = None, stream=None):
        torch = H.require_gpu()
        self.torch = torch
        self.L = H.lib()
        self.prec = precision_code(precision)
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        params = list(params or [])
        if max_cav is None:
            max_cav = max(1, len(params))
        if max_tiles is None:
            max_tiles = max(1, sum(self.tiles(p.nx, p.ny) for p in params))
        h = C.c_void_p()
        H.check(self.L.pf_batch_create_empty(self.prec, int(max_cav), int(max_tiles),
                                             C.c_void_p(self.stream.cuda_stream), C.byref(h)),
                "pf_batch_create_empty")
        self.handle = h
        self.max_cav = int(max_cav)
        self.slots: dict[int, dict[str, Any]] = {}
        for p in params:
            self.add(p)

    @staticmethod
    def tiles(nx: int, ny: int) -> int:
        return ((nx + 63) // 64) * ((ny + 15) // 16)

    def workspace_bytes(self, p: H.CavityParams) -> int:
        n = C.c_size_t()
        H.check(self.L.pf_batch_workspace_bytes(self.prec, 1, C.byref(p), C.byref(n)),
                "pf_batch_workspace_bytes")
        return int(n.value)

    def add(self, p: H.CavityParams, workspace=None) -> int:
        nbytes = self.workspace_bytes(p)
        if workspace is None:
            workspace = self.torch.empty(nbytes + 256, dtype=self.torch.uint8, device=self.device)
        base = workspace.data_ptr()
        off = (-base) % 256
        slot = C.c_int()
        H.check(self.L.pf_batch_add(self.handle, C.byref(p), C.c_void_p(base + off), C.c_size_t(nbytes),
                                    C.byref(slot)), "pf_batch_add")
        s = int(slot.value)
        self.slots[s] = {"params": p, "ws": workspace, "bytes": nbytes}
        return s

    def remove(self, slot: int) -> None:
        H.check(self.L.pf_batch_remove(self.handle, int(slot)), "pf_batch_remove")
        self.slots.pop(slot, None)

    def set_params(self, slot: int, p: H.CavityParams) -> None:
        H.check(self.L.pf_batch_set_params(self.handle, int(slot), C.byref(p)), "pf_batch_set_params")
        self.slots[slot]["params"] = p

    def params(self, slot: int) -> H.CavityParams:
        return self.slots[slot]["params"]

    def resolved_params(self, slot: int) -> H.CavityParams:
        """The slot's parameters with the library's defaults resolved (tiling,
        iterations per round, FP64 iterate, inner tolerance)."""
        out = H.CavityParams()
        H.check(self.L.pf_batch_resolved_params(self.handle, int(slot), C.byref(out)), "pf_batch_resolved_params")
        return out

    # -- fields -----------------------------------------------------------
    def upload(self, slot: int, u=None, v=None, p=None, ghost: bool = True) -> None:
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, p)]
        ld = next((a.shape[1] for a in arrs if a is not None), 0)
        for a in arrs:
            if a is not None and a.shape[1] != ld:
                raise ValueError("u, v, p must share a row stride")
        H.check(self.L.pf_batch_upload(self.handle, int(slot), _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]),
                                       int(ld), 1 if ghost else 0), "pf_batch_upload")

    def download(self, slot: int, u=None, v=None, p=None, ghost: bool = True) -> None:
        arrs = [u, v, p]
        for a in arrs:
            if a is not None and (a.dtype != np.float64 or not a.flags.c_contiguous):
                raise ValueError("download targets must be C-contiguous float64 arrays")
        ld = next((a.shape[1] for a in arrs if a is not None), 0)
        H.check(self.L.pf_batch_download(self.handle, int(slot), _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]),
                                         int(ld), 1 if ghost else 0), "pf_batch_download")

    def zero_fields(self) -> None:
        H.check(self.L.pf_batch_zero_fields(self.handle), "pf_batch_zero_fields")

    def set_meta(self, slot: int, dt_prev: float, hist: Sequence[float]) -> None:
        h = np.ascontiguousarray(np.asarray(list(hist), dtype=np.float64))
        H.check(self.L.pf_batch_set_meta(self.handle, int(slot), float(dt_prev),
                                         _ptr(h) if h.size else None, int(h.size)), "pf_batch_set_meta")

    def set_dt(self, slot: int, dt: float) -> None:
        H.check(self.L.pf_batch_set_dt(self.handle, int(slot), float(dt)), "pf_batch_set_dt")

    def set_rhs(self, slot: int, rhs: np.ndarray) -> None:
        nx = self.slots[slot]["params"].nx
        r = np.ascontiguousarray(np.asarray(rhs, dtype=np.float64).reshape(-1, nx))
        H.check(self.L.pf_batch_set_rhs(self.handle, int(slot), _ptr(r), int(nx)), "pf_batch_set_rhs")

    def solution(self, slot: int) -> np.ndarray:
        p = self.slots[slot]["params"]
        x = np.empty((p.ny, p.nx), dtype=np.float64)
        H.check(self.L.pf_batch_get_solution(self.handle, int(slot), _ptr(x), int(p.nx)),
                "pf_batch_get_solution")
        return x

    def centerline(self, slot: int) -> np.ndarray:
        p = self.slots[slot]["params"]
        out = np.empty(p.ny, dtype=np.float64)
        H.check(self.L.pf_batch_centerline(self.handle, int(slot), _ptr(out)), "pf_batch_centerline")
        return out

    # -- execution ----------------------------------------------------------
    def _active(self, slots):
        if slots is None:
            return None
        a = np.zeros(self.max_cav, dtype=np.int32)
        for s in slots:
            a[s] = 1
        return a

    def run(self, nsteps: int = 1, slots=None) -> dict[int, H.StepDiag]:
        diags = (H.StepDiag * self.max_cav)()
        act = self._active(slots)
        H.check(self.L.pf_batch_run(self.handle, int(nsteps), _ptr(act), C.cast(diags, C.c_void_p)),
                "pf_batch_run")
        return {s: diags[s] for s in self.slots}

    def launch(self, nsteps: int = 1, slots=None) -> None:
        act = self._active(slots)
        H.check(self.L.pf_batch_launch(self.handle, int(nsteps), _ptr(act)), "pf_batch_launch")

    def poll(self) -> bool:
        d = C.c_int()
        H.check(self.L.pf_batch_poll(self.handle, C.byref(d)), "pf_batch_poll")
        return bool(d.value)

    # Continuous batching: arm one cavity while the others keep running,
    # queue a chunk of rounds, and collect the cavities that completed.
    def start(self, slot: int, nsteps: int = 1) -> None:
        H.check(self.L.pf_batch_start(self.handle, int(slot), int(nsteps)), "pf_batch_start")

    def advance(self, rounds: int) -> None:
        H.check(self.L.pf_batch_advance(self.handle, int(rounds)), "pf_batch_advance")

    def finished(self) -> list[int]:
        f = np.zeros(self.max_cav, dtype=np.int32)
        H.check(self.L.pf_batch_finished(self.handle, _ptr(f)), "pf_batch_finished")
        return [int(s) for s in np.nonzero(f)[0] if int(s) in self.slots]

    def diag(self) -> dict[int, H.StepDiag]:
        diags = (H.StepDiag * self.max_cav)()
        H.check(self.L.pf_batch_diag(self.handle, C.cast(diags, C.c_void_p)), "pf_batch_diag")
        return {s: diags[s] for s in self.slots}

    def rounds(self) -> int:
        r = C.c_int64()
        H.check(self.L.pf_batch_rounds(self.handle, C.byref(r)), "pf_batch_rounds")
        return int(r.value)

    def set_graph(self, rounds: int) -> None:
        """Rounds per replayed hipGraph (0: one launch per round)."""
        H.check(self.L.pf_batch_set_graph(self.handle, int(rounds)), "pf_batch_set_graph")

    def set_timing(self, enable: bool = True) -> None:
        H.check(self.L.pf_batch_set_timing(self.handle, 1 if enable else 0), "pf_batch_set_timing")

    def timing(self) -> tuple[float, int]:
        ms, r = C.c_double(), C.c_int64()
        H.check(self.L.pf_batch_timing(self.handle, C.byref(ms), C.byref(r)), "pf_batch_timing")
        return float(ms.value), int(r.value)

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.L.pf_batch_destroy(self.handle)
            self.handle = None
            self.slots.clear()

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass


class DeviceCavity:
    """One cavity on the device (a batch of one)."""

    def __init__(self, params: H.CavityParams, precision: str | int | None = None):
        self.batch = DeviceBatch([params], precision)
        self.slot = next(iter(self.batch.slots))
        self.nx, self.ny = params.nx, params.ny
        self.prec = self.batch.prec

    @property
    def params(self) -> H.CavityParams:
        return self.batch.params(self.slot)

    def set_params(self, p: H.CavityParams) -> None:
        self.batch.set_params(self.slot, p)

    def resolved_params(self) -> H.CavityParams:
        return self.batch.resolved_params(self.slot)

    def upload(self, u=None, v=None, p=None, ghost=True):
        self.batch.upload(self.slot, u, v, p, ghost)

    def download(self, u=None, v=None, p=None, ghost=True):
        self.batch.download(self.slot, u, v, p, ghost)

    def set_meta(self, dt_prev, hist):
        self.batch.set_meta(self.slot, dt_prev, hist)

    def set_dt(self, dt):
        self.batch.set_dt(self.slot, dt)

    def set_rhs(self, rhs):
        self.batch.set_rhs(self.slot, rhs)

    def solution(self):
        return self.batch.solution(self.slot)

    def centerline(self):
        return self.batch.centerline(self.slot)

    def run(self, nsteps: int = 1) -> H.StepDiag:
        return self.batch.run(nsteps)[self.slot]

    def rounds(self) -> int:
        return self.batch.rounds()

    def close(self):
        self.batch.close()


def isfinite_or_nan(x: float) -> float:
    return x if math.isfinite(x) else float("nan")
