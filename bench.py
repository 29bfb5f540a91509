#!/usr/bin/env python3
"""Benchmark of PyFlow's projection time step on one MI355X.

Default workload (the headline metric's configuration): one lid-driven cavity
on a 1024x1024 grid, Re=1000, lid U=1, QUICK advection, CFL 0.5, growth 1.05,
PCG tol 1e-6, lin_maxiter 10000, started from rest, dx = dy = 1 (the CLI
convention lx = nx-1). A "step" is one projection time step of that cavity.
Arithmetic: fields, stencils, SpMV and axpy in float32; dot products, norms,
the Poisson right-hand side and an FP64 iterative-refinement residual in
float64 ("mixed"). Inputs are resident in HBM when the timed region starts.

Prints ONE JSON line with the metric, the dominant kernel's roofline figures,
the field error of the timed run against the float64 path, and a CPU baseline
(the FP64 NumPy oracle timed on this host on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "pyflow-solver_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Mcell-steps/s at 1024²; PCG GB/s as % HBM peak; p50/p99 served-solve latency"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024, help="grid size (n x n)")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--inner-rtol", type=float, default=1e-2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the float64 comparison runs")
    ap.add_argument("--oracle-steps", type=int, default=2,
                    help="steps from rest also compared with the FP64 NumPy oracle on the host, outside the "
                         "timed region (default 2: step 0 from rest has a zero right-hand side, step 1 is the "
                         "first full projection; about 90 s of host time at 1024x1024; 0: skip)")
    ap.add_argument("--no-drop-in", action="store_true",
                    help="skip timing the drop-in path (SimulationDriver.run) and the C1 CPU baselines")
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain FP64 GPU form")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default tiles instead of timing the candidates first")
    ap.add_argument("--tile", default=None, metavar="ROWS,CAP",
                    help="batch: region height (tiles) and per-cavity workgroup cap for every cavity, instead "
                         "of the autotuner's pick")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="run only a plain 256 MiB device copy (10x) to calibrate PMC byte counters")
    ap.add_argument("--requests", type=int, default=64, help="serve: number of requests")
    ap.add_argument("--step-scale", type=float, default=1.0,
                    help="serve: scale of the per-request step count U{100..2000} (default 1: the configuration)")
    ap.add_argument("--hbm-budget-gib", type=float, default=4.0,
                    help="serve: device memory the admitted requests may hold together (GiB)")
    ap.add_argument("--workload", choices=["cavity", "poisson", "batch", "serve"], default="cavity",
                    help="cavity: one lid-driven cavity (default, the headline metric); poisson: one 4096x4096 "
                         "pinned-Neumann solve to 1e-10 (FP32 PCG + FP64 refinement); batch: 256 cavities of "
                         "32..512 cells a side advanced together; serve: Poisson arrivals of cavity requests "
                         "admitted into a running batch (continuous batching)")
    return ap.parse_args()


def log(args, msg):
    if not args.quiet:
        print(msg, file=sys.stderr, flush=True)


def c2_params(n, prec_inner_rtol):
    from pyflow._device import cavity_params
    return cavity_params(n, n, dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0, Re=1000.0, cfl_target=0.5,
                         cfl_growth=1.05, lin_tol=1e-6, lin_maxiter=10000, jacobi=True,
                         inner_rtol=prec_inner_rtol)


# Compulsory (algorithmic) HBM bytes per cell, s = bytes per field element.
# PCG iteration of the reference recurrence with two global syncs: d = z +
# beta d (read r, d_old; write d), then x += a d, r -= a A d (read x, r, d;
# write x, r): 8 vectors. The single-reduction kernel moves 6 (r, p, x read and
# written); the metric still grades it against 8.
def pcg_iter_bytes(s):
    return 8 * s


# Rest of a step (predictor reads u, v, writes the RHS; PCG init reads b;
# correction reads u*, v*, p... ): the 9-vector figure of the metric contract.
def step_other_bytes(s):
    return 9 * s


# One FP64 refinement pass (MIXED): x64 += e (r/w x64, r/w e) and
# r = b - A x64 (read x64, b; write r): 24 + 20 bytes per cell.
REFINE_BYTES = 44


# rocprofv3 FETCH_SIZE / WRITE_SIZE summaries (calibrated on a plain copy) per
# workload: the grid, and the tiling they were taken with (the library's
# default, also the autotuner's pick: regions of 4 tiles, at most 256
# workgroups, two PCG iterations a round).
TRAFFIC_PROFILES = {
    1024: os.path.join(ROOT, "profiles", "k_round_c2_1024_pmc_traffic.md"),
    4096: os.path.join(ROOT, "profiles", "k_round_poisson_4096_pmc_traffic.md"),
}
PROFILED_TILE = (4, 256, 2)
# The 256-cavity batch's summaries, taken with 2-tile regions and a cap of
# 1024 workgroups per cavity (one region per workgroup).
BATCH_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_round_occ2_batch256_pmc_traffic.md")
PROFILED_BATCH_TILE = (2, 1024)


def traffic_line(path):
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":", 1)[1])
    return None


def batch_traffic(precision, rows, cap):
    """Median HBM bytes per round launch of the 256-cavity batch, measured
    with the profiled tiling; None for other tilings or precisions."""
    if precision != "mixed" or (rows, cap) != PROFILED_BATCH_TILE:
        return None
    return traffic_line(BATCH_TRAFFIC_PROFILE)


def measured_traffic(n, precision, tile=None):
    """HBM bytes per k_round launch measured for this grid's mixed-precision
    workload with the profiled tiling; None for other grids or tilings, or
    when the summary is absent."""
    path = TRAFFIC_PROFILES.get(n)
    if path is None or precision != "mixed" or not os.path.exists(path):
        return None
    if tile is not None and (tile.tile_rows, tile.max_workgroups, tile.pcg_per_round) != PROFILED_TILE:
        return None
    return traffic_line(path)


def _constant_part(p, pref):
    """Share of rel_l2_dp carried by the mean of the interior difference (the
    pinned operator's slowest mode is nearly constant): |mean| sqrt(N) / ||p||."""
    import numpy as np
    d = p[1:-1, 1:-1] - pref[1:-1, 1:-1]
    return float(abs(d.mean()) * np.sqrt(d.size) / max(np.linalg.norm(pref[1:-1, 1:-1]), 1e-300))


def _grad_rel(p, p_ref):
    """Relative L2 difference of the face gradients of two ghost-layout
    pressure fields (dx = dy = 1): the part of p the velocity sees. A constant
    offset of p (the slowest mode of the pinned operator) does not enter."""
    import numpy as np
    a, b = p[1:-1, 1:-1], p_ref[1:-1, 1:-1]
    num = np.sum(np.diff(a - b, axis=0) ** 2) + np.sum(np.diff(a - b, axis=1) ** 2)
    den = np.sum(np.diff(b, axis=0) ** 2) + np.sum(np.diff(b, axis=1) ** 2)
    return float(np.sqrt(num / max(den, 1e-300)))


def oracle_accuracy(n, params, args):
    """The first steps of the run (from rest) on the GPU in mixed precision
    and with the FP64 NumPy oracle on the same inputs: field and continuity
    differences. Step 0 from rest has a zero right-hand side; step 1 is the
    first full projection."""
    import numpy as np
    from types import SimpleNamespace
    from oracle import pyflow_oracle as O
    from pyflow._device import DeviceBatch
    k = args.oracle_steps
    log(args, f"[bench] FP64 NumPy oracle, first {k} steps at {n}x{n}")
    cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05,
                          Re=1000.0, lid_velocity=1.0, lin_tol=1e-6, lin_maxiter=10000)
    ref = O.allocate_state(n, n)
    t0 = time.perf_counter()
    for it in range(k):
        res, _ = O.step(cfg, ref, it)
    t_oracle = time.perf_counter() - t0
    b = DeviceBatch([params], "mixed")
    d = b.run(k)[0]
    u = np.zeros((n + 2, n + 2)); v = np.zeros_like(u); p = np.zeros_like(u)
    b.download(0, u, v, p)
    b.close()
    po = ref.fields["p"]
    cont_o = float(res["continuity"])
    return {
        "steps": k,
        "max_abs_du_over_U": float(np.max(np.abs(u - ref.fields["u"]))),
        "max_abs_dv_over_U": float(np.max(np.abs(v - ref.fields["v"]))),
        "rel_l2_dp": float(np.linalg.norm(p - po) / max(np.linalg.norm(po), 1e-300)),
        "rel_l2_dp_constant_part": _constant_part(p, po),
        "rel_l2_dgradp": _grad_rel(p, po),
        "continuity": float(d.continuity),
        "continuity_oracle": cont_o,
        "continuity_diff": float(abs(d.continuity - cont_o)),
        "oracle_seconds": round(t_oracle, 2),
    }


def _oracle_step_costs(n, threads):
    """FP64 NumPy oracle: one n x n QUICK step (step 1 from rest) with the
    PCG capped at 8 and at 40 iterations; t(m) = a + b m gives the cost a
    outside the PCG and the cost b per iteration. ``threads``: BLAS threads
    (None: the library's default, every core)."""
    from types import SimpleNamespace
    from oracle import pyflow_oracle as O
    limiter = None
    if threads is not None:
        try:
            from threadpoolctl import threadpool_limits
            limiter = threadpool_limits(limits=threads)
        except Exception:  # pragma: no cover
            limiter = None
    cfg0 = dict(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05, Re=1000.0,
                lid_velocity=1.0, lin_tol=1e-6)
    st = O.allocate_state(n, n)
    O.step(SimpleNamespace(lin_maxiter=10, **cfg0), st, 0)  # from rest: zero RHS, lid row set
    times = {}
    for m in (8, 40):
        s2 = O.OracleState(n, n, {k: v.copy() for k, v in st.fields.items()}, dict(st.meta))
        t0 = time.perf_counter()
        O.step(SimpleNamespace(lin_maxiter=m, **cfg0), s2, 1)
        times[m] = time.perf_counter() - t0
    if limiter is not None:
        limiter.unregister()
    b = (times[40] - times[8]) / 32.0
    a = max(times[8] - 8 * b, 0.0)
    return a, b


def cpu_baseline(n, iters_per_step, args, iters_source):
    """Time the FP64 NumPy oracle on this host (one BLAS thread, and again
    with every core) and extrapolate to ``iters_per_step`` PCG iterations."""
    a, b = _oracle_step_costs(n, 1)
    t_step = a + b * iters_per_step
    a_all, b_all = _oracle_step_costs(n, None)
    t_all = a_all + b_all * iters_per_step
    ncpu = os.cpu_count()
    return {
        "value": round(n * n / t_step / 1e6, 6),
        "unit": "Mcell-steps/s",
        "cores": 1,
        "kind": "port",
        "host_cpu_count": ncpu,
        "all_cores": {"value": round(n * n / t_all / 1e6, 6), "cores": ncpu,
                      "note": "same oracle with BLAS free to use every core; only its dot products and norms "
                              "are threaded, the stencils are single-threaded NumPy"},
        "sample": (f"FP64 NumPy oracle, one {n}x{n} QUICK step (step 1 from rest) timed with the PCG capped at 8 and "
                   f"40 iterations: {a:.3f} s outside PCG + {b * 1e3:.2f} ms per PCG iteration, extrapolated to "
                   f"{iters_per_step:.0f} iterations/step ({iters_source}); BLAS limited to 1 thread"),
    }


def c1_baselines(args):
    """The 64^2 cavity run (Re 100, lid 1, 50 steps, lin_tol 1e-8, lin_maxiter
    300, QUICK, FP64): the drop-in SimulationDriver.run on the GPU, and the FP64
    oracle on this host in its vectorised form (i) and in the reference's
    loop-faithful form (ii: per-cell Python advection and a per-step LIL
    Jacobi assembly)."""
    import numpy as np
    from types import SimpleNamespace
    import torch
    from oracle import pyflow_oracle as O
    from pyflow.core.ghost_fields import allocate_state
    from pyflow.drivers.simulation_driver import SimulationDriver
    from pyflow.residuals.manager import ResidualManager
    n, steps = 64, 50
    kw = dict(Re=100.0, lid_velocity=1.0, cfl_target=0.5, cfl_growth=1.05, advection_scheme="quick",
              disable_advection=False, lin_tol=1e-8, lin_maxiter=300)
    cfg = SimpleNamespace(nx=n, ny=n, diagnostics=False, force_quiet=True, precision="fp64", **kw)
    for _ in SimulationDriver(cfg, allocate_state(n, n), ResidualManager()).run(max_steps=2):
        pass  # warm: first device allocation and code load
    state = allocate_state(n, n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in SimulationDriver(cfg, state, ResidualManager()).run(max_steps=steps):
        pass
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    u_gpu = state.fields["u"].copy()
    ocfg = SimpleNamespace(**kw)
    out = {}
    for name, loops in (("vectorised", False), ("loop_faithful", True)):
        st = O.allocate_state(n, n)
        t0 = time.perf_counter()
        for it in range(steps):
            O.step(ocfg, st, it, loops=loops)
        out[name] = time.perf_counter() - t0
    err = float(np.max(np.abs(u_gpu - st.fields["u"])))
    cells = n * n * steps
    return {
        "workload": "64x64 lid-driven cavity, Re 100, lid 1, 50 steps, lin_tol 1e-8, lin_maxiter 300, QUICK, FP64",
        "gpu_drop_in": {"seconds": round(t_gpu, 4), "ms_per_step": round(t_gpu / steps * 1e3, 3),
                        "mcell_steps_s": round(cells / t_gpu / 1e6, 4),
                        "path": "SimulationDriver.run -> step -> libpyflow_hip (one device step per yield)"},
        "cpu_vectorised": {"seconds": round(out["vectorised"], 3), "mcell_steps_s": round(cells / out["vectorised"] / 1e6, 4),
                           "cores": 1, "kind": "port"},
        "cpu_loop_faithful": {"seconds": round(out["loop_faithful"], 3),
                              "mcell_steps_s": round(cells / out["loop_faithful"] / 1e6, 4), "cores": 1,
                              "kind": "port"},
        "max_abs_du_gpu_vs_oracle": err,
        "speedup_vs_vectorised": round(out["vectorised"] / t_gpu, 2),
        "speedup_vs_loop_faithful": round(out["loop_faithful"] / t_gpu, 2),
    }


def drop_in_c2(args, steps, tile):
    """The 1024^2 cavity from rest for ``steps`` steps three times: through
    the C-ABI batch run (pf_batch_run of all steps), through one
    pf_batch_run per step, and through the drop-in SimulationDriver.run (one
    device step per yield). The driver's first yield also creates the device
    copy of the state and uploads it; later yields add parameter arming, the
    scalar state download, the diagnostics dictionaries and the centerline,
    measured against the per-step runs."""
    from types import SimpleNamespace
    import torch
    from pyflow import autotune
    from pyflow._device import DeviceBatch
    from pyflow.core.ghost_fields import allocate_state
    from pyflow.drivers.simulation_driver import SimulationDriver
    from pyflow.residuals.manager import ResidualManager
    n = args.n
    params = c2_params(n, args.inner_rtol)
    if tile is not None:
        autotune.apply(params, tile)
    batch = DeviceBatch([params], args.precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.run(steps)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    batch.close()
    batch = DeviceBatch([params], args.precision)
    per_run = []
    for _ in range(steps):
        t0 = time.perf_counter()
        batch.run(1)
        per_run.append(time.perf_counter() - t0)
    batch.close()
    cfg = SimpleNamespace(nx=n, ny=n, Re=1000.0, lid_velocity=1.0, cfl_target=0.5, cfl_growth=1.05,
                          advection_scheme="quick", disable_advection=False, lin_tol=1e-6, lin_maxiter=10000,
                          diagnostics=False, force_quiet=True, precision=args.precision,
                          inner_rtol=args.inner_rtol)
    if tile is not None:  # the drop-in cavity runs with the batch's tiling
        cfg.tile_rows, cfg.max_workgroups, cfg.pcg_per_round = tile.tile_rows, tile.max_workgroups, tile.pcg_per_round
    state = allocate_state(n, n)
    per_yield = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tl = t0
    for _ in SimulationDriver(cfg, state, ResidualManager()).run(max_steps=steps):
        t = time.perf_counter()
        per_yield.append(t - tl)
        tl = t
    torch.cuda.synchronize()
    t_drv = time.perf_counter() - t0
    later = (sum(per_yield[1:]) - sum(per_run[1:])) / max(1, steps - 1)
    return {"steps_from_rest": steps, "driver_ms_per_step": round(t_drv / steps * 1e3, 3),
            "batch_run_ms_per_step": round(t_batch / steps * 1e3, 3),
            "one_run_per_step_ms_per_step": round(sum(per_run) / steps * 1e3, 3),
            "driver_first_yield_ms": round(per_yield[0] * 1e3, 3),
            "one_run_first_step_ms": round(per_run[0] * 1e3, 3),
            "host_overhead_ms_per_later_step": round(later * 1e3, 3),
            "driver_mcell_steps_s": round(n * n * steps / t_drv / 1e6, 3)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}); there is no multi-GPU path. "
              "Independent cavities would shard by request, one process per GPU; a single large cavity by "
              "row slabs with a two-row halo exchange per PCG round.", file=sys.stderr)
        return 2
    import torch
    torch.cuda.set_device(0)
    if args.copy_calibrate:
        n_bytes = 256 << 20
        src = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibrate": {"bytes_read_per_copy": n_bytes, "bytes_written_per_copy": n_bytes,
                                             "copies": 10}}))
        return 0
    if args.workload == "poisson":
        return run_poisson(args)
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "serve":
        return run_serve(args)
    return run_cavity(args)


# ---------------------------------------------------------------------------
# "Plain FP64 GPU form": the reference's step written as one PyTorch kernel per
# NumPy operation, float64, no fusion, no graph, one cavity -- the comparison
# form of the >= 10x target. It is a baseline timed here, not part of the
# library; the library never uses PyTorch for arithmetic.
# ---------------------------------------------------------------------------
def _plain_ops(torch):
    F = torch.nn.functional

    def grad_c(p):
        gx = torch.zeros_like(p)
        gy = torch.zeros_like(p)
        gx[:, :-1] = (p[:, 1:] - p[:, :-1]) / 1.0
        gy[:-1, :] = (p[1:, :] - p[:-1, :]) / 1.0
        return gx, gy

    def div_c(u, v):
        xp = torch.empty_like(u)
        xp[:, 1:] = (u[:, 1:] - u[:, :-1]) / 1.0
        xp[:, 0] = u[:, 0] / 1.0
        yp = torch.empty_like(v)
        yp[1:, :] = (v[1:, :] - v[:-1, :]) / 1.0
        yp[0, :] = v[0, :] / 1.0
        return (torch.zeros_like(u) + xp) + yp

    def div_central(u, v):
        d = torch.zeros_like(u)
        d[1:-1, 1:-1] = (u[1:-1, 2:] - u[1:-1, :-2]) / 2.0 + (v[2:, 1:-1] - v[:-2, 1:-1]) / 2.0
        d[1:-1, 0] = (u[1:-1, 1] - u[1:-1, 0]) + (v[2:, 0] - v[:-2, 0]) / 2.0
        d[0, 1:-1] = (u[0, 2:] - u[0, :-2]) / 2.0 + (v[1, 1:-1] - v[0, 1:-1])
        d[-1, 1:-1] = (u[-1, 2:] - u[-1, :-2]) / 2.0 + (v[-1, 1:-1] - v[-2, 1:-1])
        return d

    def quick(u, v, f):
        ph = F.pad(f[None, None], (2, 2, 2, 2), mode="replicate")[0, 0]
        P = ph[2:-2, 2:-2]
        E, W, WW = ph[2:-2, 3:-1], ph[2:-2, 1:-3], ph[2:-2, 0:-4]
        N, S, SS = ph[3:-1, 2:-2], ph[1:-3, 2:-2], ph[0:-4, 2:-2]
        pos_u, pos_v = u >= 0, v >= 0
        fe = torch.where(pos_u, (0.375 * E + 0.75 * P) - 0.125 * W, (0.375 * W + 0.75 * P) - 0.125 * E)
        fw = torch.where(pos_u, (0.375 * P + 0.75 * W) - 0.125 * WW, (0.375 * WW + 0.75 * W) - 0.125 * P)
        fn = torch.where(pos_v, (0.375 * N + 0.75 * P) - 0.125 * S, (0.375 * S + 0.75 * P) - 0.125 * N)
        fs = torch.where(pos_v, (0.375 * P + 0.75 * S) - 0.125 * SS, (0.375 * SS + 0.75 * S) - 0.125 * P)
        return (u * fe - u * fw) / 1.0 + (v * fn - v * fs) / 1.0

    def matvec(x):
        gx, gy = grad_c(x)
        out = -div_c(gx, gy)
        out[0, 0] = x[0, 0]
        return out

    return grad_c, div_c, div_central, quick, matvec


def plain_fp64_step(torch, ops, u, v, meta, maxiter, tol=1e-6, cfl=0.5, growth=1.05, Re=1000.0, lid=1.0):
    grad_c, div_c, div_central, quick, matvec = ops
    ny, nx = u.shape
    umax = float(torch.max(torch.abs(u))) + 1e-12
    vmax = float(torch.max(torch.abs(v))) + 1e-12
    dt_cand = min(min(cfl / umax, cfl / vmax), 0.25 / max(1.0 / Re, 1e-12))
    cap = meta["dt_prev"] * growth
    dt = cap if dt_cand > cap else dt_cand
    hist = (meta["hist"] + [float(torch.linalg.norm(div_central(u, v)))])[-5:]
    if len(hist) >= 3 and hist[-3] > 0 and hist[-2] > hist[-3] * 1.01 and hist[-1] > hist[-2] * 1.01:
        dt *= 0.5
    meta["dt_prev"], meta["hist"] = dt, hist
    u_old, v_old = u.clone(), v.clone()
    u = u_old - dt * quick(u_old, v_old, u_old)
    v = v_old - dt * quick(u_old, v_old, v_old)
    rhs = -div_c(u, v) / max(dt, 1e-14)
    rhs = rhs - torch.mean(rhs)
    rhs[0, 0] = 0.0
    minv = torch.full_like(rhs, 1.0 / 4.0)
    minv[0, :] = 1.0
    minv[-1, :] = 1.0
    minv[:, 0] = 1.0
    minv[:, -1] = 1.0
    x = torch.zeros_like(rhs)
    r = rhs.clone()
    atol = max(0.0, tol * float(torch.linalg.norm(rhs)))
    rho_prev, p = None, None
    for _ in range(maxiter):
        if float(torch.linalg.norm(r)) < atol:
            break
        z = minv * r
        rho = torch.sum(r * z)
        p = z if p is None else z + (rho / rho_prev) * p
        q = matvec(p)
        alpha = rho / torch.sum(p * q)
        x = x + alpha * p
        r = r - alpha * q
        rho_prev = rho
    x = x - x[0, 0]
    gx, gy = grad_c(x)
    u = u - dt * gx
    v = v - dt * gy
    u[-1, :] = lid
    cont = float(torch.linalg.norm(div_central(u, v))) / (nx * ny)
    return u, v, cont


def plain_fp64_baseline(n, iters_per_step):
    """Time the plain FP64 GPU form on one n x n step (step 1 from rest) with
    the PCG capped at 8 and at 40 iterations: t(m) = a + b m, extrapolated to
    the fused run's mean iterations per step."""
    import torch
    ops = _plain_ops(torch)
    dev = torch.device("cuda")
    times = {}
    for m in (8, 40, 8, 40):
        u = torch.zeros((n, n), dtype=torch.float64, device=dev)
        v = torch.zeros_like(u)
        u[-1, :] = 1.0
        meta = {"dt_prev": 1e-2, "hist": []}
        u, v, _ = plain_fp64_step(torch, ops, u, v, meta, 4)  # step 0 from rest
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plain_fp64_step(torch, ops, u, v, meta, m)
        torch.cuda.synchronize()
        times[m] = time.perf_counter() - t0  # second pass of each cap is kept (warm)
    b = (times[40] - times[8]) / 32.0
    a = max(times[8] - 8 * b, 0.0)
    t_step = a + b * iters_per_step
    return {"value": round(n * n / t_step / 1e6, 4), "unit": "Mcell-steps/s",
            "definition": "one PyTorch float64 kernel per reference NumPy operation, no fusion, no graph, one cavity "
                          "per launch sequence; SciPy-style PCG with a host norm check per iteration",
            "sample": f"{n}x{n} step 1 from rest, PCG capped at 8 and 40 iterations: {a * 1e3:.1f} ms outside PCG + "
                      f"{b * 1e3:.3f} ms per iteration, extrapolated to {iters_per_step:.0f} iterations/step"}


def tune_tile(args, nx, ny):
    """The autotuner's tile for one cavity (timed on the device before the
    measured run; None with --no-autotune)."""
    if args.no_autotune:
        return None
    from pyflow import autotune
    t0 = time.perf_counter()
    tile = autotune.tune(nx, ny, args.precision, log=lambda m: log(args, m))
    log(args, f"[bench] autotune {nx}x{ny}: tile_rows={tile.tile_rows} max_workgroups={tile.max_workgroups} "
              f"pcg_per_round={tile.pcg_per_round} ({tile.us_per_iteration:.2f} us/iteration, "
              f"{time.perf_counter() - t0:.1f} s)")
    return tile


def tile_config(tile):
    if tile is None:
        return "library default"
    return {"tile_rows": tile.tile_rows, "max_workgroups": tile.max_workgroups,
            "pcg_per_round": tile.pcg_per_round, "tuned_us_per_iteration": round(tile.us_per_iteration, 3)}


def run_cavity(args):
    import numpy as np
    import torch
    from pyflow import autotune
    from pyflow._device import DeviceBatch
    n = args.n
    N = n * n
    s = 4 if args.precision == "mixed" else 8
    params = c2_params(n, args.inner_rtol)
    tile = tune_tile(args, n, n)
    batch = DeviceBatch([params if tile is None else autotune.apply(params, tile)], args.precision)
    if args.warmup > 0:
        log(args, f"[bench] warmup {args.warmup} steps at {n}x{n}")
        batch.run(args.warmup)
    batch.zero_fields()
    batch.set_meta(0, 1e-2, [])
    batch.set_timing(True)
    log(args, f"[bench] timed run: {args.steps} steps at {n}x{n} ({args.precision})")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d = batch.run(args.steps)[0]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    gpu_ms, rounds = batch.timing()
    its = int(d.rp_iterations_run)
    outer = int(d.outer_passes_run)
    log(args, f"[bench] {wall:.3f} s wall, {gpu_ms / 1e3:.3f} s GPU in {rounds} rounds, {its} PCG iterations, "
              f"{outer} refinement passes, final continuity {d.continuity:.3e}")
    if d.steps_done != args.steps:
        raise RuntimeError(f"expected {args.steps} steps, device reports {d.steps_done}")
    alg_bytes = N * (args.steps * step_other_bytes(s) + its * pcg_iter_bytes(s) + outer * REFINE_BYTES)
    pcg_bytes = N * its * pcg_iter_bytes(s)
    avg_launch_s = (gpu_ms / 1e3) / max(rounds, 1)
    achieved = alg_bytes / max(rounds, 1) / avg_launch_s / 1e9
    # PCG share of the GPU time: the mixed path runs two PCG iterations per
    # round (one with pcg_per_round = 1), the float64 path one iteration in two
    # rounds (SciPy's recurrence)
    if args.precision == "mixed":
        two = (tile.pcg_per_round == 2) if tile is not None else os.environ.get("PYFLOW_CA", "1") != "0"
        rounds_per_iter = 0.5 if two else 1.0
    else:
        rounds_per_iter = 2.0
    pcg_time_s = (gpu_ms / 1e3) * min(1.0, rounds_per_iter * its / max(rounds, 1))
    pcg_gbs = pcg_bytes / max(pcg_time_s, 1e-30) / 1e9
    value = N * args.steps / wall / 1e6
    traffic = measured_traffic(n, args.precision, tile)

    accuracy = None
    if not args.no_accuracy and args.precision == "mixed":
        log(args, "[bench] float64 comparison run (same steps, FP64 fused path)")
        u32 = np.zeros((n + 2, n + 2)); v32 = np.zeros_like(u32); p32 = np.zeros_like(u32)
        batch.download(0, u32, v32, p32)
        ref = DeviceBatch([params], "fp64")
        d64 = ref.run(args.steps)[0]
        u64 = np.zeros_like(u32); v64 = np.zeros_like(u32); p64 = np.zeros_like(u32)
        ref.download(0, u64, v64, p64)
        ref.close()
        accuracy = {
            "vs": "float64 fused GPU path (parity-tested against the FP64 NumPy oracle)",
            "max_abs_du_over_U": float(np.max(np.abs(u32 - u64))),
            "max_abs_dv_over_U": float(np.max(np.abs(v32 - v64))),
            "rel_l2_dp": float(np.linalg.norm(p32 - p64) / max(np.linalg.norm(p64), 1e-300)),
            "rel_l2_dp_constant_part": _constant_part(p32, p64),
            "rel_l2_dgradp": _grad_rel(p32, p64),
            "continuity": float(d.continuity),
            "continuity_fp64": float(d64.continuity),
            "continuity_diff": float(abs(d.continuity - d64.continuity)),
            "pcg_iterations_fp64": int(d64.rp_iterations_run),
        }
        log(args, f"[bench] accuracy {accuracy}")
    if not args.no_accuracy and args.oracle_steps > 0:
        accuracy = accuracy or {}
        accuracy["vs_oracle"] = oracle_accuracy(n, params, args)
        log(args, f"[bench] accuracy vs the FP64 NumPy oracle {accuracy['vs_oracle']}")

    # The baselines run the FP64 recurrence (the reference's), so they are
    # extrapolated to the FP64 path's own iteration count when it was measured.
    if accuracy and accuracy.get("pcg_iterations_fp64"):
        base_iters = accuracy["pcg_iterations_fp64"] / args.steps
        iters_source = "the FP64 fused path's mean over the same steps"
    else:
        base_iters = its / args.steps
        iters_source = "the timed run's mean"
    cpu = None
    if not args.no_cpu_baseline:
        log(args, "[bench] CPU baseline (FP64 NumPy oracle, bounded sample)")
        cpu = cpu_baseline(n, base_iters, args, iters_source)
        cpu["speedup"] = round(value / cpu["value"], 1)
    plain = None
    if not args.no_plain:
        log(args, "[bench] plain FP64 GPU form (one PyTorch kernel per NumPy op), bounded sample")
        plain = plain_fp64_baseline(n, base_iters)
        plain["speedup"] = round(value / plain["value"], 2)
    drop_in = None
    if not args.no_drop_in:
        log(args, "[bench] drop-in path: SimulationDriver.run at 1024^2 and the 64^2 run with its CPU baselines")
        drop_in = {"c2": drop_in_c2(args, 20, tile), "c1": c1_baselines(args)}
        log(args, f"[bench] drop-in {drop_in}")

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic: lid-driven cavity from rest (zero u, v, p)",
        "config": {"workload": f"lid-driven cavity {n}x{n}, Re=1000, U=1, QUICK, {args.steps} projection steps, "
                               "PCG tol 1e-6, lin_maxiter 10000, dx=dy=1",
                   "grid": [n, n], "precision": args.precision, "inner_rtol": args.inner_rtol,
                   "tile": tile_config(tile)},
        "roofline": {
            "bound": "hbm",
            "kernel": "pf::k_round (one launch = one phase round of the cavity; PCG rounds dominate)",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "actual_frac": None if traffic is None else round(traffic / avg_launch_s / 1e9 / HBM_PEAK_GBS, 4),
            "avg_launch_us": round(avg_launch_s * 1e6, 3),
            "alg_bytes_per_launch": int(alg_bytes / max(rounds, 1)),
            "launches": rounds,
            "launches_per_step": round(rounds / args.steps, 1),
        },
        "pcg": {"iterations": its, "refinement_passes": outer, "gbs": round(pcg_gbs, 2),
                "hbm_frac": round(pcg_gbs / HBM_PEAK_GBS, 4), "iters_per_step": round(its / args.steps, 1)},
        "accuracy": accuracy,
        "cpu_baseline": cpu,
        "plain_fp64_gpu": plain,
        "drop_in": drop_in,
    }
    print(json.dumps(out), flush=True)
    batch.close()
    return 0


def run_logged(args, batch, nsteps, what):
    """batch.run(nsteps) through launch/poll, logging the rounds executed every
    10 s (the chunks double up to 256 rounds, so host polling gaps are a small
    part of the wall time; GPU time comes from the chunks' events)."""
    batch.launch(nsteps)
    t0 = last = time.perf_counter()
    while not batch.poll():
        time.sleep(0.0002)
        if time.perf_counter() - last > 10.0:
            last = time.perf_counter()
            log(args, f"[bench] {what}: {last - t0:.0f} s, {batch.rounds()} rounds")
    return batch.diag()


def gpu_roofline(kernel, alg_bytes, gpu_ms, rounds, traffic=None):
    avg_launch_s = (gpu_ms / 1e3) / max(rounds, 1)
    achieved = alg_bytes / max(rounds, 1) / max(avg_launch_s, 1e-30) / 1e9
    return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "avg_launch_us": round(avg_launch_s * 1e6, 3), "alg_bytes_per_launch": int(alg_bytes / max(rounds, 1)),
            "launches": rounds}


def cpu_pcg_iteration_seconds(n, iters):
    """FP64 NumPy oracle: seconds per Jacobi-PCG iteration on the n x n pinned
    operator, from `iters` iterations on one BLAS thread."""
    import numpy as np
    from oracle import pyflow_oracle as O
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(limits=1)
    except Exception:  # pragma: no cover
        limiter = None
    b = np.random.default_rng(1).standard_normal(n * n)
    b -= b.mean()
    b[0] = 0.0
    minv = O.jacobi_inv_diag(n, n, 1.0, 1.0)
    t0 = time.perf_counter()
    O.cg(lambda z: O.neg_laplacian_pinned(z, n, n, 1.0, 1.0), b, 1e-30, iters, lambda r: minv * r)
    dt = (time.perf_counter() - t0) / iters
    if limiter is not None:
        limiter.unregister()
    return dt


def run_poisson(args):
    """One pinned-Neumann Poisson solve on a 4096 x 4096 grid: right-hand side
    standard normal (seed 0), mean removed, cell 0 pinned to 0, x0 = 0; FP32
    Jacobi-PCG passes with FP64 refinement until ||b - A x|| <= 1e-10 ||b||."""
    import numpy as np
    import torch
    from pyflow import _hip as H
    from pyflow import autotune
    from pyflow._device import DeviceBatch, cavity_params
    n = 4096 if args.n == 1024 else args.n
    N = n * n
    tol = 1e-10
    rhs = np.random.default_rng(0).standard_normal((n, n))
    rhs -= rhs.mean()
    rhs[0, 0] = 0.0
    params = cavity_params(n, n, dx=1.0, dy=1.0, lin_tol=tol, lin_maxiter=1_000_000, jacobi=True,
                           op=H.OP_NEG_LAP_PINNED, mode=H.MODE_SOLVE, inner_rtol=args.inner_rtol)
    tile = tune_tile(args, n, n)
    if tile is not None:
        autotune.apply(params, tile)
    batch = DeviceBatch([params], "mixed")
    batch.set_rhs(0, rhs)
    if args.warmup > 0:
        log(args, f"[bench] warmup solve at {n}x{n}")
        batch.run(1)
    batch.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d = run_logged(args, batch, 1, "solve")[0]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms, rounds = batch.timing()
    its, outer = int(d.rp_iterations), int(d.outer_passes)
    rel = float(d.rp_residual) / float(d.bnorm)
    log(args, f"[bench] {wall:.3f} s, {its} PCG iterations, {outer} refinement passes, ||r||/||b|| = {rel:.3e}")
    if not (d.rp_converged and rel <= tol):
        raise RuntimeError(f"solve did not reach {tol}: rel residual {rel:.3e}, info {d.rp_info}")
    s = 4
    alg_bytes = N * (its * pcg_iter_bytes(s) + outer * REFINE_BYTES)
    pcg_gbs = N * its * pcg_iter_bytes(s) / (gpu_ms / 1e3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        m = 10
        log(args, f"[bench] CPU baseline: {m} oracle PCG iterations at {n}x{n}")
        t_it = cpu_pcg_iteration_seconds(n, m)
        cpu = {"value": round(N * 8 * 8 / t_it / 1e9, 3), "unit": "GB/s (FP64 PCG, 8 vectors x 8 B per cell)",
               "cores": 1, "kind": "port",
               "sample": f"FP64 NumPy oracle PCG, {m} iterations at {n}x{n}: {t_it * 1e3:.1f} ms per iteration "
                         f"(one BLAS thread); the GPU solve's {its} iterations would take {t_it * its:.0f} s"}
    out = {
        "metric": "PCG GB/s as % HBM peak (4096x4096 Poisson, FP32 PCG + FP64 refinement to 1e-10)",
        "value": round(pcg_gbs, 2), "unit": "GB/s", "n_gpus": 1, "steps": 1, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic: standard-normal right-hand side, seed 0",
        "config": {"workload": f"pinned-Neumann Poisson {n}x{n}, Jacobi-PCG, tol 1e-10, x0 = 0", "grid": [n, n],
                   "inner_rtol": args.inner_rtol, "tile": tile_config(tile)},
        "roofline": gpu_roofline("pf::k_round (one launch = one PCG round of the solve: two iterations)",
                                 alg_bytes, gpu_ms, rounds, measured_traffic(n, "mixed", tile)),
        "pcg": {"iterations": its, "refinement_passes": outer, "rel_residual": rel,
                "hbm_frac": round(pcg_gbs / HBM_PEAK_GBS, 4), "solve_s": round(wall, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    batch.close()
    return 0


# The 256-cavity batch: the autotuner picks the region height and the
# per-cavity workgroup cap for the whole batch (two-tile regions, one per
# workgroup, run two workgroups per CU; several regions per workgroup make a
# round hold fewer workgroups, each paying its state round trip and tail once).
# Without it: the default regions and at most 16 workgroups per cavity.
# Served requests keep the default (256): few cavities run at a time there, and
# a large one capped at 16 workgroups would leave most CUs idle.
MANY_CAVITY_MAXWG = 16


def batch_round_kernel(ns, rows, cap, precision):
    """Round kernel the library launches for a mixed-precision batch of n x n
    cavities below 1536 cells a side: the two-workgroups-per-CU instance when
    every cavity's regions hold at most two tiles, one region per workgroup."""
    from pyflow import autotune
    blk = [autotune.blocking(int(n), int(n), rows or 4, cap) for n in ns]
    if precision == "mixed" and all(tpb <= 2 and rpw == 1 for tpb, _, rpw, _ in blk):
        return "pf::k_round_occ2<false>"
    return "pf::k_round<float, U, false>"


def run_batch(args):
    import numpy as np
    import torch
    from pyflow import autotune
    from pyflow._device import DeviceBatch
    steps = args.steps
    from pyflow.workloads import batch_params
    params, ns = batch_params(256, 0, args.inner_rtol)
    cells = int(np.sum(ns.astype(np.int64) ** 2))
    rows, cap, tuned = 0, MANY_CAVITY_MAXWG, None
    if args.tile:
        rows, cap = (int(v) for v in args.tile.split(","))
    elif not args.no_autotune:
        t0 = time.perf_counter()
        (rows, cap), tuned = autotune.tune_batch(params, args.precision, log=lambda m: log(args, m))
        log(args, f"[bench] autotune batch: tile_rows={rows} max_workgroups={cap} "
                  f"({time.perf_counter() - t0:.1f} s)")
    for p in params:
        p.tile_rows = rows
        p.max_workgroups = cap
    batch = DeviceBatch(params, args.precision)
    if args.warmup > 0:
        log(args, f"[bench] warmup {args.warmup} steps, 256 cavities")
        batch.run(args.warmup)
        batch.zero_fields()
        for k in range(len(params)):
            batch.set_meta(k, 1e-2, [])
    batch.set_timing(True)
    log(args, f"[bench] timed run: {steps} steps of 256 cavities ({cells} cells)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    diags = run_logged(args, batch, steps, "batch")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms, rounds = batch.timing()
    s = 4 if args.precision == "mixed" else 8
    its = np.array([diags[k].rp_iterations_run for k in range(len(params))], dtype=np.int64)
    outer = np.array([diags[k].outer_passes_run for k in range(len(params))], dtype=np.int64)
    done = np.array([diags[k].steps_done for k in range(len(params))])
    if not np.all(done == steps):
        raise RuntimeError("not every cavity completed its steps")
    n2 = ns.astype(np.int64) ** 2
    alg_bytes = float(np.sum(n2 * (steps * step_other_bytes(s) + its * pcg_iter_bytes(s) + outer * REFINE_BYTES)))
    value = cells * steps / wall / 1e6
    nan = [k for k in range(len(params)) if diags[k].nan_detected]
    log(args, f"[bench] {wall:.3f} s wall, {rounds} rounds, PCG iterations per cavity-step mean "
              f"{its.mean() / steps:.0f}, max {its.max() / steps:.0f}")
    out = {
        "metric": "Mcell-steps/s, 256 independent cavities of 32..512 cells a side advanced together",
        "value": round(value, 3), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(wall / steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic: 256 lid-driven cavities from rest; n_i = 32*U{1..16}, Re_i log-uniform [100, 5000], "
                "seed 0",
        "config": {"workload": f"256 cavities, QUICK, tol 1e-6, lin_maxiter 10000, {steps} steps", "cells": cells,
                   "precision": args.precision, "inner_rtol": args.inner_rtol,
                   "tile_rows": rows, "max_workgroups_per_cavity": cap,
                   "autotune_ps_per_cell_iteration": None if tuned is None else
                   {f"rows{k[0]}_cap{k[1]}": round(v, 4) for k, v in tuned.items()}},
        "roofline": gpu_roofline(f"{batch_round_kernel(ns, rows, cap, args.precision)} "
                                 "(all cavities' regions in one launch)", alg_bytes, gpu_ms, rounds,
                                 batch_traffic(args.precision, rows, cap)),
        "pcg": {"iterations_total": int(its.sum()), "mean_per_cavity_step": round(float(its.mean()) / steps, 1)},
        "nan_cavities": nan,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    batch.close()
    return 0


def serve(reqs, arrivals, precision, inner_rtol, hbm_budget, chunk_rounds=64, max_cav=64, log_fn=None):
    """Continuous batching through pyflow.serving.CavityServer: requests are
    submitted at their arrival times, admitted by HBM footprint into the
    running batch, advanced in chunks and retired when done. Returns
    per-request completion times (seconds from the start), the makespan and
    the server's peak concurrency and device bytes."""
    import collections
    from pyflow.serving import CavityRequest, CavityServer
    ns, res, steps = reqs
    srv = CavityServer(precision, hbm_budget=hbm_budget, max_cavities=max_cav, chunk_rounds=chunk_rounds,
                       inner_rtol=inner_rtol)
    order = collections.deque(sorted(range(len(ns)), key=lambda k: arrivals[k]))
    ticket_of, done = {}, {}
    t0 = time.perf_counter()
    last_log = t0
    while order or not srv.idle():
        now = time.perf_counter() - t0
        while order and arrivals[order[0]] <= now:
            k = order.popleft()
            req = CavityRequest(int(ns[k]), int(ns[k]), int(steps[k]), Re=float(res[k]), lid_velocity=1.0,
                                scheme="quick", lin_tol=1e-6, lin_maxiter=10000)
            ticket_of[srv.submit(req, t0 + arrivals[k])] = k
        if srv.idle():
            time.sleep(max(0.0, arrivals[order[0]] - (time.perf_counter() - t0)))
            continue
        for r in srv.pump():
            done[ticket_of[r.ticket]] = r.t_done - t0
        if log_fn and time.perf_counter() - last_log > 10.0:
            last_log = time.perf_counter()
            log_fn(f"[bench] serving: {len(done)} done, {srv.running} running, {srv.waiting + len(order)} waiting")
    makespan = time.perf_counter() - t0
    peak = (srv.peak_running, srv.peak_bytes, srv.evicted)
    srv.close()
    return done, makespan, peak


def run_serve(args):
    import numpy as np
    import torch
    from pyflow.workloads import poisson_arrivals, serve_requests
    count = args.requests
    budget = int(args.hbm_budget_gib * (1 << 30))
    reqs = serve_requests(count, 0, args.step_scale)
    ns, res, steps = reqs
    work = ns.astype(np.float64) ** 2 * steps
    # capacity: the first requests served back to back (all arriving at t = 0)
    m = min(8, count)
    log(args, f"[bench] capacity probe: {m} requests at once")
    cal = (ns[:m], res[:m], steps[:m])
    _, span, _ = serve(cal, np.zeros(m), args.precision, args.inner_rtol, budget, log_fn=lambda m_: log(args, m_))
    capacity = float(np.sum(work[:m])) / span  # cell-steps/s
    lam = 0.7 * capacity / float(np.mean(work))  # requests/s at 70% of capacity
    arrivals = poisson_arrivals(count, lam, seed=1)
    log(args, f"[bench] capacity {capacity / 1e6:.1f} Mcell-steps/s; arrival rate {lam:.3f}/s; {count} requests")
    torch.cuda.reset_peak_memory_stats()
    done, makespan, (peak_running, peak_bytes, evicted) = serve(reqs, arrivals, args.precision, args.inner_rtol,
                                                                budget, log_fn=lambda m_: log(args, m_))
    reserved = torch.cuda.max_memory_reserved()
    lat = np.array([done[k] - arrivals[k] for k in range(count)])
    out = {
        "metric": "p50/p99 served-solve latency (continuous batching of lid-driven cavity requests)",
        "value": round(float(np.percentile(lat, 50)), 4), "unit": "s (p50 latency)", "n_gpus": 1,
        "steps": int(np.sum(steps)), "warmup": 0, "ms_per_step": None, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": ("synthetic: Poisson arrivals (seed 1) at 70% of the measured capacity; n ~ round(exp(U(ln 64, "
                 f"ln 2048))), Re log-uniform [100, 1e4], steps ~ U{{100..2000}} x {args.step_scale} (seed 0)"),
        "config": {"workload": f"{count} cavity requests, QUICK, tol 1e-6, lin_maxiter 10000", "requests": count,
                   "arrival_rate_per_s": round(lam, 4), "precision": args.precision,
                   "admission": "HBM footprint, first fit in arrival order, in one arena of the budget "
                                "(pyflow.serving.CavityServer)",
                   "hbm_budget_bytes": budget, "step_scale": args.step_scale,
                   "peak_running": peak_running, "peak_workspace_bytes": int(peak_bytes),
                   "torch_max_reserved_bytes": int(reserved), "evicted_nan": evicted},
        "latency_s": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4),
                      "mean": round(float(np.mean(lat)), 4), "max": round(float(np.max(lat)), 4)},
        "throughput_mcell_steps_s": round(float(np.sum(work)) / makespan / 1e6, 3),
        "capacity_mcell_steps_s": round(capacity / 1e6, 3),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
