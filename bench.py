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
    ap.add_argument("--no-accuracy", action="store_true", help="skip the float64 comparison run")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="run only a plain 256 MiB device copy (10x) to calibrate PMC byte counters")
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


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_round_c2_1024_pmc_traffic.md")


def measured_traffic(n, precision):
    """HBM bytes per k_round launch measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (calibrated on a plain copy) for the 1024x1024 mixed-precision
    workload; None for other workloads or when the summary is absent."""
    if n != 1024 or precision != "mixed" or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":", 1)[1])
    return None


def cpu_baseline(n, iters_per_step, args):
    """Time the FP64 NumPy oracle on this host: one step at n x n with the PCG
    capped at m1 and at m2 iterations; t(m) = a + b m gives the non-PCG cost a
    and the per-iteration cost b, extrapolated to the GPU's mean iterations."""
    import numpy as np
    from types import SimpleNamespace
    from oracle import pyflow_oracle as O
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(limits=1)
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
    t_step = a + b * iters_per_step
    return {
        "value": round(n * n / t_step / 1e6, 6),
        "unit": "Mcell-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"FP64 NumPy oracle, one {n}x{n} QUICK step (step 1 from rest) timed with the PCG capped at 8 and "
                   f"40 iterations: {a:.3f} s outside PCG + {b * 1e3:.2f} ms per PCG iteration, extrapolated to "
                   f"the GPU run's mean {iters_per_step:.0f} iterations/step (BLAS limited to 1 thread)"),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}); this build has no multi-GPU path. "
              "Independent cavities shard by request across GPUs; see DESIGN.md.", file=sys.stderr)
        return 2
    import numpy as np
    import torch
    from pyflow._device import DeviceBatch
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
    n = args.n
    N = n * n
    s = 4 if args.precision == "mixed" else 8
    params = c2_params(n, args.inner_rtol)
    batch = DeviceBatch([params], args.precision)
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
    # PCG share of the GPU time: the mixed path runs one single-reduction round
    # per PCG iteration, the float64 path two (SciPy's recurrence)
    rounds_per_iter = 1.0 if args.precision == "mixed" else 2.0
    pcg_time_s = (gpu_ms / 1e3) * min(1.0, rounds_per_iter * its / max(rounds, 1))
    pcg_gbs = pcg_bytes / max(pcg_time_s, 1e-30) / 1e9
    value = N * args.steps / wall / 1e6

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
            "continuity": float(d.continuity),
            "continuity_fp64": float(d64.continuity),
            "continuity_diff": float(abs(d.continuity - d64.continuity)),
            "pcg_iterations_fp64": int(d64.rp_iterations_run),
        }
        log(args, f"[bench] accuracy {accuracy}")

    cpu = None
    if not args.no_cpu_baseline:
        log(args, "[bench] CPU baseline (FP64 NumPy oracle, bounded sample)")
        cpu = cpu_baseline(n, its / args.steps, args)

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
                   "grid": [n, n], "precision": args.precision, "inner_rtol": args.inner_rtol},
        "roofline": {
            "bound": "hbm",
            "kernel": "pf::k_round (one launch = one phase round of the cavity; PCG rounds dominate)",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(n, args.precision),
            "avg_launch_us": round(avg_launch_s * 1e6, 3),
            "alg_bytes_per_launch": int(alg_bytes / max(rounds, 1)),
            "launches": rounds,
        },
        "pcg": {"iterations": its, "refinement_passes": outer, "gbs": round(pcg_gbs, 2),
                "hbm_frac": round(pcg_gbs / HBM_PEAK_GBS, 4), "iters_per_step": round(its / args.steps, 1)},
        "accuracy": accuracy,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    batch.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
