"""Autotuner for the tile shape and schedule of the PCG rounds.

The PCG round dominates every workload (about 98% of the launches of a
1024x1024 step). Three knobs shape it, all carried per cavity in
``pf_cavity_params``:

* ``tile_rows``: region height in 16-row tiles (regions are 64 cells wide);
* ``max_workgroups``: cap on the cavity's workgroups; regions are grouped so
  that each workgroup walks ``rpw`` consecutive regions;
* ``pcg_per_round``: MIXED, 1, 2 or 3 PCG iterations per round (s-step PCG,
  s = 2 or 3), with an FP32 or (on large grids, s <= 2) an FP64 iterate.
  Three iterations per round are a candidate only when ``PYFLOW_TUNE_S3=1``:
  on MI355X its round costs more per iteration than the two-iteration round
  (21.3 us for three against 13.2 us for two at 1024x1024), so it can only
  win a timing by noise, and a 200-step 1024x1024 run that picked it was 9%
  slower.

They change where cells are computed and the order in which FP64 partial sums
are folded, never the arithmetic of a cell. ``tune`` times each distinct
candidate on the device: a pinned-Neumann Poisson solve (the operator of
``solvers/pressure_solver.py:97-106`` in the reference) with a seeded N(0,1)
right-hand side runs a few warm-up rounds, then a timed span of rounds, and the
candidate's cost is GPU time per PCG iteration executed in that span. The
tolerance is set so the solve cannot finish inside the span. The candidates
within 10% of the fastest are timed twice more, in turn, and each keeps its
best time, so one slow or fast span does not decide the pick. ``tune_batch``
does the same for the region height and per-cavity workgroup cap of a whole
batch.

Results are cached per (nx, ny, precision, device) in-process and, when
``PYFLOW_TUNE_CACHE`` names a file, in that JSON file. ``pyflow.autotune`` has
no counterpart in the reference, whose step runs on the host.
"""
from __future__ import annotations

import json
import os
import threading
from dataclasses import asdict, dataclass
from typing import Iterable, Sequence

import numpy as np

from . import _hip as H
from ._device import DeviceBatch, cavity_params, precision_code

__all__ = ["Tile", "blocking", "candidates", "tune", "tune_batch", "apply", "cached"]

TX, TY = 64, 16  # cells per tile (one 64-lane wave per tile row)
_WG_CAPS = (32, 64, 128, 256, 512, 1024)
_ROWS = (1, 2, 4, 8)


@dataclass(frozen=True)
class Tile:
    tile_rows: int
    max_workgroups: int
    pcg_per_round: int
    us_per_iteration: float = float("nan")


def blocking(nx: int, ny: int, tile_rows: int, max_workgroups: int) -> tuple[int, int, int, int]:
    """(tpb, regions, regions per workgroup, workgroups) the library derives
    from a cavity's tile fields; mirrors ``blocking_of`` in pf_core.hip."""
    tiles_x, tiles_y = -(-nx // TX), -(-ny // TY)
    m = 1
    while m * 2 <= tile_rows and m < 8:
        m *= 2
    while m > 1 and m // 2 >= tiles_y:
        m //= 2
    regions = tiles_x * (-(-tiles_y // m))
    k = 1
    while -(-regions // k) > max(1, max_workgroups):
        k *= 2
    return m, regions, k, -(-regions // k)


def candidates(nx: int, ny: int, precision: str | int = "mixed",
               caps: Iterable[int] = _WG_CAPS, rows: Iterable[int] = _ROWS) -> list[Tile]:
    """Distinct (tile_rows, max_workgroups, pcg_per_round) settings for a
    cavity: settings that give the library the same blocking are listed once."""
    prec = precision_code(precision)
    mixed = prec == H.PREC_MIXED
    # The library keeps the PCG iterate in FP64 from 1536 cells a side (or as
    # PYFLOW_XD says); three iterations per round need an FP32 iterate.
    xd_env = os.environ.get("PYFLOW_XD")
    xd = (xd_env not in ("0", "")) if xd_env is not None else max(nx, ny) >= 1536
    s3 = os.environ.get("PYFLOW_TUNE_S3") == "1"
    seen, out = set(), []
    for r in rows:
        for cap in caps:
            tpb, _, rpw, nblk = blocking(nx, ny, r, cap)
            for s in (3, 2, 1):
                if s == 3 and not (s3 and mixed and tpb in (2, 4) and not xd):
                    continue
                if s == 2 and not (mixed and tpb <= 4):
                    continue
                key = (tpb, rpw, nblk, s if mixed else 1)
                if key in seen:
                    continue
                seen.add(key)
                out.append(Tile(tpb, cap, s if mixed else 1))
    return out


def apply(p: H.CavityParams, tile: Tile) -> H.CavityParams:
    """Write a tile's fields into cavity parameters (in place; returns p)."""
    p.tile_rows = int(tile.tile_rows)
    p.max_workgroups = int(tile.max_workgroups)
    p.pcg_per_round = int(tile.pcg_per_round)
    return p


_lock = threading.Lock()
_cache: dict[str, Tile] = {}


def _device_name() -> str:
    torch = H.require_gpu()
    return torch.cuda.get_device_name(torch.cuda.current_device())


def _key(nx: int, ny: int, prec: int) -> str:
    return f"{nx}x{ny}/{'mixed' if prec == H.PREC_MIXED else 'fp64'}/{_device_name()}"


def _load_file() -> dict:
    path = os.environ.get("PYFLOW_TUNE_CACHE")
    if not path or not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _store_file(key: str, tile: Tile) -> None:
    path = os.environ.get("PYFLOW_TUNE_CACHE")
    if not path:
        return
    data = _load_file()
    data[key] = asdict(tile)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def cached(nx: int, ny: int, precision: str | int = "mixed") -> Tile | None:
    prec = precision_code(precision)
    key = _key(nx, ny, prec)
    with _lock:
        if key in _cache:
            return _cache[key]
    rec = _load_file().get(key)
    if rec is None:
        return None
    tile = Tile(**rec)
    with _lock:
        _cache[key] = tile
    return tile


def _solve_params(nx: int, ny: int, prec: int, tile: Tile, maxiter: int) -> H.CavityParams:
    p = cavity_params(nx, ny, dx=1.0, dy=1.0, lin_tol=1e-30, lin_maxiter=maxiter, op=H.OP_NEG_LAP_PINNED,
                      mode=H.MODE_SOLVE)
    return apply(p, tile)


def _rhs(nx: int, ny: int, seed: int) -> np.ndarray:
    # The standalone Poisson workload's right-hand side: N(0,1), mean removed, pinned cell 0.
    b = np.random.default_rng(seed).standard_normal((ny, nx))
    b -= b.mean()
    b.flat[0] = 0.0
    return b


def _time_rounds(batch: DeviceBatch, slots: Sequence[int], warmup: int, rounds: int) -> tuple[float, int]:
    """GPU ms of `rounds` rounds after `warmup` rounds, and the PCG iterations
    the batch's cavities executed in them."""
    for s in slots:
        batch.start(s, 1)
    batch.advance(warmup)
    while not batch.poll():
        pass
    batch.finished()  # pulls the device state
    before = sum(int(d.rp_iterations_run) for d in batch.diag().values())
    batch.set_timing(True)
    batch.advance(rounds)
    while not batch.poll():
        pass
    ms, _ = batch.timing()
    batch.set_timing(False)
    batch.finished()
    after = sum(int(d.rp_iterations_run) for d in batch.diag().values())
    return ms, after - before


def measure(nx: int, ny: int, precision: str | int, tile: Tile, *, warmup: int = 16, rounds: int = 64,
            seed: int = 0) -> Tile:
    """Time one candidate; returns it with ``us_per_iteration`` filled in."""
    prec = precision_code(precision)
    maxiter = 1 << 30
    p = _solve_params(nx, ny, prec, tile, maxiter)
    b = DeviceBatch([p], prec)
    try:
        b.set_rhs(0, _rhs(nx, ny, seed))
        ms, its = _time_rounds(b, [0], warmup, rounds)
    finally:
        b.close()
    us = 1e3 * ms / its if its > 0 else float("inf")
    return Tile(tile.tile_rows, tile.max_workgroups, tile.pcg_per_round, us)


def tune(nx: int, ny: int, precision: str | int = "mixed", *, warmup: int = 16, rounds: int = 64,
         force: bool = False, log=None) -> Tile:
    """Fastest candidate for one cavity of nx x ny cells (cached)."""
    prec = precision_code(precision)
    if not force:
        hit = cached(nx, ny, prec)
        if hit is not None:
            return hit
    def timed(cand: Tile) -> Tile:
        t = measure(nx, ny, prec, cand, warmup=warmup, rounds=rounds)
        if log:
            log(f"[autotune] {nx}x{ny} rows={t.tile_rows} maxwg={t.max_workgroups} s={t.pcg_per_round}: "
                f"{t.us_per_iteration:.2f} us/iteration")
        return t

    times = [timed(c) for c in candidates(nx, ny, prec)]
    finite = [t.us_per_iteration for t in times if np.isfinite(t.us_per_iteration)]
    if finite:
        close = [k for k, t in enumerate(times) if t.us_per_iteration <= 1.1 * min(finite)]
        if len(close) > 1:
            for _ in range(2):
                for k in close:
                    t = timed(times[k])
                    if t.us_per_iteration < times[k].us_per_iteration:
                        times[k] = t
    best = min(times, key=lambda t: t.us_per_iteration) if times else None
    if best is None or not np.isfinite(best.us_per_iteration):
        raise RuntimeError(f"autotune found no working candidate for {nx}x{ny}")
    key = _key(nx, ny, prec)
    with _lock:
        _cache[key] = best
    _store_file(key, best)
    return best


def tune_batch(params: Sequence[H.CavityParams], precision: str | int = "mixed", *,
               tilings: Iterable[tuple[int, int]] = ((4, 16), (4, 64), (4, 256), (2, 64), (2, 1024)),
               warmup: int = 8, rounds: int = 24, log=None) -> tuple[tuple[int, int], dict[tuple[int, int], float]]:
    """Region height and per-cavity workgroup cap that make a whole batch's
    PCG rounds fastest.

    Every cavity of ``params`` (their sizes; other fields are ignored) runs
    the pinned Poisson solve together in one batch per tiling
    ``(tile_rows, max_workgroups)`` of ``tilings``. Regions of at most two
    tiles with one region per workgroup let the library run two workgroups
    per CU. Returns the best tiling and the GPU picoseconds per
    cell-iteration of each."""
    prec = precision_code(precision)
    sizes = [(int(p.nx), int(p.ny)) for p in params]
    results: dict[tuple[int, int], float] = {}
    for rows, cap in tilings:
        ps = [_solve_params(nx, ny, prec, Tile(rows, cap, 0), 1 << 30) for nx, ny in sizes]
        b = DeviceBatch(ps, prec)
        try:
            for s, (nx, ny) in enumerate(sizes):
                b.set_rhs(s, _rhs(nx, ny, s))
            ms, its = _time_rounds(b, list(range(len(sizes))), warmup, rounds)
        finally:
            b.close()
        # cells advanced per iteration differ by cavity: weight by the mean cavity size
        cells = sum(nx * ny for nx, ny in sizes) / len(sizes)
        results[(rows, cap)] = 1e9 * ms / (its * cells) if its > 0 else float("inf")
        if log:
            log(f"[autotune] batch of {len(sizes)} tile_rows={rows} cap={cap}: "
                f"{results[(rows, cap)]:.3f} ps/cell-iteration")
    best = min(results, key=results.get)
    return best, results
