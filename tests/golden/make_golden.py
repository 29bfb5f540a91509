"""Write golden input/output vectors by running the reference implementation.

Run in the development container only (the reference tree is not present on
the GPU machines):

    python tests/golden/make_golden.py /root/reference/src

It imports the reference's pure-Python package from the given ``src``
directory, with bytecode caching redirected to a temporary directory so no
``__pycache__`` file shipped with the reference is ever read, and evaluates:

* ``advect_quick`` / ``advect_upwind`` on seeded random fields (the only
  specification of the non-standard QUICK faces is the reference source),
* the coherent operators and the central divergence,
* ``solve_pressure_poisson`` in both modes,
* short ``step`` trajectories (QUICK/upwind/disabled advection, lid on/off,
  NaN propagation), and
* the ``pyflow run --json-stream`` line stream of the 64x64 Re=100 case,
* ``SimulationConfig.config_hash`` of a few configurations.

Outputs: ``tests/golden/golden_ops.npz``, ``golden_steps.npz`` and
``golden_cli_c1.jsonl`` (wall times included but never compared),
``golden_config.json``.
"""
from __future__ import annotations

import contextlib
import io
import json
import os
import sys
import tempfile
from types import SimpleNamespace

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _import_reference(src: str):
    sys.dont_write_bytecode = True
    sys.pycache_prefix = tempfile.mkdtemp(prefix="pf_golden_pyc_")
    sys.path.insert(0, src)
    import pyflow  # noqa: F401
    from pyflow.numerics.operators import advection
    from pyflow.numerics import core_ops_coherent, fluid_ops
    from pyflow.solvers import pressure_solver, solver
    from pyflow.core import ghost_fields
    from pyflow.residuals.manager import ResidualManager
    from pyflow import cli
    assert os.path.abspath(pyflow.__file__).startswith(os.path.abspath(src)), pyflow.__file__
    return SimpleNamespace(adv=advection, coh=core_ops_coherent, fl=fluid_ops,
                           ps=pressure_solver, sv=solver, gf=ghost_fields,
                           RM=ResidualManager, cli=cli)


def _ops(ref, out):
    rng = np.random.default_rng(1234)
    shapes = [(3, 3), (5, 9), (8, 8), (16, 12), (7, 4)]
    for k, (ny, nx) in enumerate(shapes):
        u = rng.standard_normal((ny, nx))
        v = rng.standard_normal((ny, nx))
        f = rng.standard_normal((ny, nx))
        # exact zeros exercise the >= (QUICK) vs > (upwind) branch split
        u[rng.random((ny, nx)) < 0.2] = 0.0
        v[rng.random((ny, nx)) < 0.2] = 0.0
        dx, dy = 0.37, 1.3
        out[f"ops{k}_u"] = u
        out[f"ops{k}_v"] = v
        out[f"ops{k}_f"] = f
        out[f"ops{k}_dxdy"] = np.array([dx, dy])
        out[f"ops{k}_quick"] = ref.adv.advect_quick(u, v, f, dx, dy)
        out[f"ops{k}_upwind"] = ref.adv.advect_upwind(u, v, f, dx, dy)
        out[f"ops{k}_cdiv"] = ref.fl.divergence(u, v, dx, dy)
        gx, gy = ref.coh.gradient_coherent(f, dx, dy)
        out[f"ops{k}_gx"] = gx
        out[f"ops{k}_gy"] = gy
        out[f"ops{k}_hdiv"] = ref.coh.divergence_coherent(u, v, dx, dy)
        out[f"ops{k}_lap"] = ref.coh.laplacian_coherent(f, dx, dy)
    out["n_ops"] = np.array(len(shapes))


def _pressure(ref, out):
    # bump (tests/test_preconditioner_effect.py:9-33 shape), Jacobi on/off
    nx, ny = 24, 20
    for tag, pc in (("pc", True), ("nopc", False)):
        st = ref.gf.allocate_state(nx, ny)
        ref.gf.interior_view(st.fields["u"])[ny // 3, nx // 3] = 1.0
        cfg = SimpleNamespace(nx=nx, ny=ny, lin_tol=1e-10, lin_maxiter=300,
                              diagnostics=False, enable_jacobi_pc=pc)
        p, d = ref.ps.solve_pressure_poisson(st, 0.01, 1.0, 1.0, cfg)
        out[f"ps_bump_{tag}_p"] = p
        out[f"ps_bump_{tag}_u"] = st.fields["u"]
        out[f"ps_bump_{tag}_v"] = st.fields["v"]
        out[f"ps_bump_{tag}_diag"] = np.array(
            [d["Rp_iterations"], d["Rp_residual"], float(d["Rp_converged"]),
             d["divergence_norm"], d["lin_residual_norm"], d["divergence_norm_initial"]])
    # random velocities, non-unit spacing
    rng = np.random.default_rng(99)
    nx, ny = 13, 11
    st = ref.gf.allocate_state(nx, ny)
    st.fields["u"][1:-1, 1:-1] = rng.standard_normal((ny, nx))
    st.fields["v"][1:-1, 1:-1] = rng.standard_normal((ny, nx))
    out["ps_rand_u0"] = st.fields["u"].copy()
    out["ps_rand_v0"] = st.fields["v"].copy()
    cfg = SimpleNamespace(lin_tol=1e-12, lin_maxiter=400, diagnostics=False)
    p, d = ref.ps.solve_pressure_poisson(st, 0.05, 0.1, 0.2, cfg)
    out["ps_rand_p"] = p
    out["ps_rand_u"] = st.fields["u"]
    out["ps_rand_v"] = st.fields["v"]
    out["ps_rand_diag"] = np.array(
        [d["Rp_iterations"], d["Rp_residual"], float(d["Rp_converged"]),
         d["divergence_norm"], d["lin_residual_norm"], d["divergence_norm_initial"]])
    # manufactured Dirichlet mode (tests/test_projection_method_contract.py:408-466)
    nx = ny = 8
    dx = dy = 1.0 / (nx - 1)
    X, Y = np.meshgrid(np.linspace(0, 1, nx), np.linspace(0, 1, ny))
    pa = np.sin(np.pi * X) * np.sin(np.pi * Y)
    rhs = np.zeros((ny, nx))
    rhs[1:-1, 1:-1] = -2.0 * np.pi ** 2 * pa[1:-1, 1:-1]
    rhs = rhs.reshape(-1)
    st = ref.gf.allocate_state(nx, ny)
    cfg = SimpleNamespace(lin_tol=1e-10, lin_maxiter=500, diagnostics=False)
    p, d = ref.ps.solve_pressure_poisson(st, 1.0, dx, dy, cfg, rhs_override=rhs)
    out["ps_mms_rhs"] = rhs
    out["ps_mms_p"] = p
    out["ps_mms_diag"] = np.array([d["Rp_iterations"], d["Rp_residual"],
                                   float(d["Rp_converged"]), d["divergence_norm"]])


STEP_KEYS = ["dt", "CFL", "Ru", "Rv", "continuity", "Rp_iterations", "Rp_residual",
             "Rp_converged", "divergence_norm", "nan_detected", "max_residual",
             "residual_drop_Ru_orders", "nu"]

CASES = {
    # name: (nx, ny, cfg-attrs, init, nsteps)
    "quick8": (8, 8, dict(disable_advection=False, advection_scheme="quick", cfl_target=0.5,
                          cfl_growth=1.1, Re=100.0, lid_velocity=1.0, lin_tol=1e-10,
                          lin_maxiter=200, diagnostics=False), "zero", 6),
    "upwind16": (16, 16, dict(disable_advection=False, advection_scheme="upwind", cfl_target=0.45,
                              cfl_growth=1.08, Re=100.0, lid_velocity=1.0, lin_tol=1e-10,
                              lin_maxiter=320, diagnostics=False, lx=1.0, ly=1.0), "patch", 8),
    "quick12x10": (12, 10, dict(disable_advection=False, advection_scheme="quick", cfl_target=0.5,
                                cfl_growth=1.05, Re=400.0, lid_velocity=0.7, lin_tol=1e-9,
                                lin_maxiter=300, diagnostics=False), "random", 6),
    "safeguard8": (8, 8, dict(disable_advection=False, advection_scheme="quick", cfl_target=0.9,
                              cfl_growth=1.2, Re=100.0, lid_velocity=1.0, lin_tol=1e-10,
                              lin_maxiter=50, diagnostics=False), "alternating", 6),
    "noadv12x10": (12, 10, dict(disable_advection=True, advection_scheme="upwind", cfl_target=0.5,
                                cfl_growth=1.05, Re=100.0, lid_velocity=0.0, lin_tol=1e-10,
                                lin_maxiter=60, diagnostics=False, lx=11, ly=9), "random", 4),
    "nolid9x7": (9, 7, dict(disable_advection=False, advection_scheme="quick", cfl_target=0.5,
                            cfl_growth=1.05, Re=50.0, lin_tol=1e-10, lin_maxiter=100,
                            diagnostics=False, enable_jacobi_pc=False), "random", 4),
    "nan10": (10, 10, dict(disable_advection=True, advection_scheme="upwind", cfl_target=0.5,
                           cfl_growth=1.05, Re=100.0, lid_velocity=0.0, lin_tol=1e-10,
                           lin_maxiter=30, diagnostics=False, lx=9, ly=9), "nan", 1),
    "c1_64": (64, 64, dict(disable_advection=False, advection_scheme="quick", cfl_target=0.5,
                           cfl_growth=1.05, Re=100.0, lid_velocity=1.0, lin_tol=1e-8,
                           lin_maxiter=300, diagnostics=False), "zero", 50),
}


def _init_fields(kind, nx, ny):
    u = np.zeros((ny + 2, nx + 2))
    v = np.zeros((ny + 2, nx + 2))
    if kind == "patch":
        u[5:9, 5:9] = 1.0
    elif kind == "random":
        rng = np.random.default_rng(nx * 1000 + ny)
        u[1:-1, 1:-1] = 0.3 * rng.standard_normal((ny, nx))
        v[1:-1, 1:-1] = 0.3 * rng.standard_normal((ny, nx))
    elif kind == "alternating":
        for j in range(ny):
            for i in range(nx):
                u[j + 1, i + 1] = ((-1) ** (i + j)) * (i + j) * 0.05
    elif kind == "nan":
        u[3, 3] = np.nan
    return u, v


def _steps(ref, out):
    for name, (nx, ny, attrs, init, nsteps) in CASES.items():
        cfg = SimpleNamespace(nx=nx, ny=ny, **attrs)
        st = ref.gf.allocate_state(nx, ny)
        u0, v0 = _init_fields(init, nx, ny)
        st.fields["u"][:] = u0
        st.fields["v"][:] = v0
        out[f"st_{name}_u0"] = u0
        out[f"st_{name}_v0"] = v0
        out[f"st_{name}_cfg"] = np.array(json.dumps(attrs))
        tr = ref.RM()
        rows = []
        snaps_u, snaps_v, snaps_p = [], [], []
        for it in range(nsteps):
            with contextlib.redirect_stdout(io.StringIO()):
                st, res, diag = ref.sv.step(cfg, st, tr, it)
            rows.append([float(diag[k]) if k in diag else float(res[k]) for k in STEP_KEYS])
            if nsteps <= 10 or it in (0, 1, nsteps - 1):
                snaps_u.append(st.fields["u"].copy())
                snaps_v.append(st.fields["v"].copy())
                snaps_p.append(st.fields["p"].copy())
        out[f"st_{name}_diag"] = np.array(rows)
        out[f"st_{name}_u"] = np.array(snaps_u)
        out[f"st_{name}_v"] = np.array(snaps_v)
        out[f"st_{name}_p"] = np.array(snaps_p)
        out[f"st_{name}_meta"] = np.array(json.dumps({
            "dt_prev": st.meta["dt_prev"], "divergence_history": st.meta["divergence_history"]}))
    out["step_keys"] = np.array(json.dumps(STEP_KEYS))
    out["step_cases"] = np.array(json.dumps(list(CASES)))


def _cli(ref, path):
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        rc = ref.cli.main(["run", "--nx", "64", "--ny", "64", "--re", "100", "--lid-velocity", "1",
                           "--steps", "50", "--lin-tol", "1e-8", "--cfl", "0.5", "--json-stream"])
    assert rc == 0
    with open(path, "w") as f:
        f.write(buf.getvalue())


CONFIGS = [
    dict(nx=64, ny=64),
    dict(nx=12, ny=10, Re=150.0),
    dict(nx=1024, ny=1024, Re=1000.0, lin_tol=1e-6, lin_maxiter=10000, advection_scheme="quick",
         diagnostics=False),
    dict(nx=33, ny=17, seed=7, checkpoint_interval=5, log_path="x.jsonl", cfl_growth=1.1),
]


def _configs(ref, path):
    from pyflow.config.model import SimulationConfig
    out = []
    for kw in CONFIGS:
        c = SimulationConfig(**kw)
        out.append({"kwargs": kw, "config_hash": c.config_hash, "brief": c.brief(),
                    "dump": json.loads(json.dumps(c.model_dump(), default=str))})
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src"
    ref = _import_reference(src)
    ops, steps = {}, {}
    _ops(ref, ops)
    _pressure(ref, ops)
    np.savez_compressed(os.path.join(HERE, "golden_ops.npz"), **ops)
    _steps(ref, steps)
    np.savez_compressed(os.path.join(HERE, "golden_steps.npz"), **steps)
    _cli(ref, os.path.join(HERE, "golden_cli_c1.jsonl"))
    _configs(ref, os.path.join(HERE, "golden_config.json"))
    print("wrote golden fixtures to", HERE)


if __name__ == "__main__":
    main()
