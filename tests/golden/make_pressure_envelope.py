"""How much pressure the reference's own stopping rule leaves free, on the
1024x1024 cavity's configuration.

    python tests/golden/make_pressure_envelope.py [n ...]   (default 128 256 512 1024)

The lid-driven cavity of the headline workload (Re 1000, lid U = 1, QUICK,
CFL 0.5, growth 1.05, lin_tol 1e-6, lin_maxiter 10000, dx = dy = 1, from
rest) runs 3 steps in the FP64 oracle (bitwise the reference). At every step,
from the same predicted velocity, the projection is also solved

* to lin_tol 1e-13 (the exact solution p* of that step's system), and
* to the same lin_tol 1e-6 without the Jacobi preconditioner (another solve
  the reference accepts: ``enable_jacobi_pc=False``).

A solve is only held to |b - A p| <= lin_tol |b|; the pressure it leaves can
sit anywhere in that ball, and its smooth modes move by up to about
lin_tol * cond(A), cond(A) ~ (2n/pi)^2. Per grid and step this records the
relative L2 distance from p* of the reference's pressure and of the
unpreconditioned one, and the same for the velocity correction dt * grad p
and for u, v after the step: the tolerance envelope the reference itself
has. A mixed-precision solve meets the same rule at another iterate, so its
pressure is held to a small multiple of this envelope (1024^2 test).

Along a trajectory the pressure is also sensitive to the velocity it is
computed from: a step's solve leaves u, v free by up to dt * grad of the
pressure's tolerance ball, and the next projection amplifies the smooth part
of that velocity change by up to cond(A) -- most of all in the slowest mode
of the pinned operator, which is nearly constant away from the pinned cell.
So the script also runs two other trajectories the reference's contract
accepts, every step solved without the Jacobi preconditioner, and every step
solved to 1e-13, and records per step the relative L2 change of p (and of
dt * grad p) and the largest change of u, v against the reference
trajectory: the trajectory envelope. A mixed-precision trajectory is held to
a small multiple of it. For grids up to 512 the same is recorded for the
reference trajectory with u, v rounded to float32 after every step (the
mixed path's storage), which moves p far less.

Also recorded: how far the reference's pressure moves when only the
summation order of its reductions changes (running sums first to last and
last to first), which is far smaller: two FP64 solves in different orders
stop at the same iterate, within one.

Writes tests/golden/pressure_envelope.json. CPU only; no reference import.
"""
from __future__ import annotations

import json
import os
import sys
import time
from types import SimpleNamespace

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from oracle import pyflow_oracle as O  # noqa: E402
from make_envelope import ORDERS, _numpy_with  # noqa: E402

STEPS = 3
CFG = dict(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05, Re=1000.0,
           lid_velocity=1.0, lin_tol=1e-6, lin_maxiter=10000)
REORDERS = ("seq", "rev")


def _snapshot(st):
    return O.OracleState(st.nx, st.ny, {k: v.copy() for k, v in st.fields.items()}, json.loads(json.dumps(st.meta)))


def _after(st, diag):
    p = st.fields["p"][1:-1, 1:-1].copy()
    gx, gy = O.gradient_coherent(p, 1.0, 1.0)
    return {"p": p, "corr": np.concatenate([(diag["dt"] * gx).ravel(), (diag["dt"] * gy).ravel()]),
            "u": st.fields["u"][1:-1, 1:-1].copy(), "v": st.fields["v"][1:-1, 1:-1].copy(),
            "iters": int(diag["Rp_true_iterations"])}


def _step(st, it, **over):
    cfg = SimpleNamespace(**{**CFG, **over})
    _, diag = O.step(cfg, st, it)
    return _after(st, diag)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def tolerance_envelope(n):
    st = O.allocate_state(n, n)
    rows = []
    for it in range(STEPS):
        before = _snapshot(st)
        ref = _step(st, it)
        exact = _step(_snapshot(before), it, lin_tol=1e-13, lin_maxiter=200000)
        nojac = _step(_snapshot(before), it, enable_jacobi_pc=False)
        rows.append({
            "p_rel_l2_ref_vs_exact": _rel(ref["p"], exact["p"]),
            "p_rel_l2_nojacobi_vs_exact": _rel(nojac["p"], exact["p"]),
            "p_rel_l2_ref_vs_nojacobi": _rel(ref["p"], nojac["p"]),
            "corr_rel_l2_ref_vs_exact": _rel(ref["corr"], exact["corr"]),
            "corr_rel_l2_nojacobi_vs_exact": _rel(nojac["corr"], exact["corr"]),
            "uv_max_abs_ref_vs_exact": max(float(np.max(np.abs(ref["u"] - exact["u"]))),
                                           float(np.max(np.abs(ref["v"] - exact["v"])))),
            "pcg_iterations": {"reference": ref["iters"], "exact": exact["iters"], "no_jacobi": nojac["iters"]},
        })
    return rows


def trajectory_envelope(n, float32_storage: bool):
    """Per step, the change of p, dt * grad p and u, v of other trajectories
    the reference accepts against the reference trajectory."""
    runs = {"reference": ({}, None), "no_jacobi": ({"enable_jacobi_pc": False}, None),
            "exact": ({"lin_tol": 1e-13, "lin_maxiter": 200000}, None)}
    if float32_storage:
        runs["float32_velocity"] = ({}, np.float32)
    states = {k: O.allocate_state(n, n) for k in runs}
    rows = []
    for it in range(STEPS):
        out = {k: _step(states[k], it, **over) for k, (over, _) in runs.items()}
        ref = out["reference"]
        row = {}
        for k, res in out.items():
            if k == "reference":
                continue
            row[k] = {"p_rel_l2": _rel(res["p"], ref["p"]), "corr_rel_l2": _rel(res["corr"], ref["corr"]),
                      "uv_max_abs": max(float(np.max(np.abs(res["u"] - ref["u"]))),
                                        float(np.max(np.abs(res["v"] - ref["v"])))),
                      "pcg_iterations": res["iters"]}
        row["p_norm"] = float(np.linalg.norm(ref["p"]))
        rows.append(row)
        for k, (_, store) in runs.items():
            if store is not None:
                for f in ("u", "v"):
                    states[k].fields[f][:] = states[k].fields[f].astype(store).astype(np.float64)
    return rows


def order_envelope(n):
    real = O.np
    base = O.allocate_state(n, n)
    ref = [_step(base, it) for it in range(STEPS)]
    worst = [0.0] * STEPS
    for name in REORDERS:
        O.np = _numpy_with(ORDERS[name])
        try:
            st = O.allocate_state(n, n)
            alt = [_step(st, it) for it in range(STEPS)]
        finally:
            O.np = real
        worst = [max(w, _rel(a["p"], r["p"])) for w, a, r in zip(worst, alt, ref)]
    return worst


def main(sizes):
    path = os.path.join(HERE, "pressure_envelope.json")
    out = {"config": CFG, "steps": STEPS, "grids": {}}
    if os.path.exists(path):
        with open(path) as f:
            out = json.load(f)
    for n in sizes:
        t0 = time.perf_counter()
        g = {"tolerance": tolerance_envelope(n), "trajectory": trajectory_envelope(n, n <= 512)}
        if n <= 256:
            g["order_p_rel_l2"] = order_envelope(n)
        g["seconds"] = round(time.perf_counter() - t0, 1)
        out["grids"][str(n)] = g
        print(n, json.dumps(g), flush=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main([int(a) for a in sys.argv[1:]] or [128, 256, 512, 1024])
