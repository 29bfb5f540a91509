"""The FP64 oracle reproduces the reference bit for bit on the golden vectors
that tests/golden/make_golden.py wrote by running the reference itself."""
import json
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import pyflow_oracle as O


def test_operators_bitwise(golden_ops):
    g = golden_ops
    for k in range(int(g["n_ops"])):
        u, v, f = g[f"ops{k}_u"], g[f"ops{k}_v"], g[f"ops{k}_f"]
        dx, dy = g[f"ops{k}_dxdy"]
        assert np.array_equal(O.advect_quick(u, v, f, dx, dy), g[f"ops{k}_quick"])
        assert np.array_equal(O.advect_upwind(u, v, f, dx, dy), g[f"ops{k}_upwind"])
        assert np.array_equal(O.divergence_central(u, v, dx, dy), g[f"ops{k}_cdiv"])
        gx, gy = O.gradient_coherent(f, dx, dy)
        assert np.array_equal(gx, g[f"ops{k}_gx"]) and np.array_equal(gy, g[f"ops{k}_gy"])
        assert np.array_equal(O.divergence_coherent(u, v, dx, dy), g[f"ops{k}_hdiv"])
        assert np.array_equal(O.laplacian_coherent(f, dx, dy), g[f"ops{k}_lap"])


def _ps_diag_row(d):
    return np.array([d["Rp_iterations"], d["Rp_residual"], float(d["Rp_converged"]), d["divergence_norm"],
                     d["lin_residual_norm"], d["divergence_norm_initial"]])


@pytest.mark.parametrize("tag,pc", [("pc", True), ("nopc", False)])
def test_pressure_bump_bitwise(golden_ops, tag, pc):
    g = golden_ops
    nx, ny = 24, 20
    u = np.zeros((ny, nx)); v = np.zeros((ny, nx)); p = np.zeros((ny, nx))
    u[ny // 3, nx // 3] = 1.0
    cfg = SimpleNamespace(lin_tol=1e-10, lin_maxiter=300, enable_jacobi_pc=pc)
    psol, d = O.solve_pressure_poisson(u, v, p, 0.01, 1.0, 1.0, cfg)
    assert np.array_equal(psol, g[f"ps_bump_{tag}_p"])
    assert np.array_equal(u, g[f"ps_bump_{tag}_u"][1:-1, 1:-1])
    assert np.array_equal(_ps_diag_row(d), g[f"ps_bump_{tag}_diag"])


def test_pressure_random_and_manufactured_bitwise(golden_ops):
    g = golden_ops
    u = g["ps_rand_u0"][1:-1, 1:-1].copy(); v = g["ps_rand_v0"][1:-1, 1:-1].copy()
    p = np.zeros_like(u)
    cfg = SimpleNamespace(lin_tol=1e-12, lin_maxiter=400)
    psol, d = O.solve_pressure_poisson(u, v, p, 0.05, 0.1, 0.2, cfg)
    assert np.array_equal(psol, g["ps_rand_p"])
    assert np.array_equal(u, g["ps_rand_u"][1:-1, 1:-1]) and np.array_equal(v, g["ps_rand_v"][1:-1, 1:-1])
    assert np.array_equal(_ps_diag_row(d), g["ps_rand_diag"])
    nx = ny = 8
    u = np.zeros((ny, nx)); v = np.zeros((ny, nx)); p = np.zeros((ny, nx))
    cfg = SimpleNamespace(lin_tol=1e-10, lin_maxiter=500)
    psol, d = O.solve_pressure_poisson(u, v, p, 1.0, 1 / 7, 1 / 7, cfg, rhs_override=g["ps_mms_rhs"])
    assert np.array_equal(psol, g["ps_mms_p"])
    assert np.array_equal(np.array([d["Rp_iterations"], d["Rp_residual"], float(d["Rp_converged"]),
                                    d["divergence_norm"]]), g["ps_mms_diag"])


def _cases(golden_steps):
    return json.loads(str(golden_steps["step_cases"]))


def test_step_trajectories_bitwise(golden_steps):
    s = golden_steps
    keys = json.loads(str(s["step_keys"]))
    for name in _cases(s):
        cfg = SimpleNamespace(**json.loads(str(s[f"st_{name}_cfg"])))
        u0 = s[f"st_{name}_u0"]
        ny, nx = u0.shape[0] - 2, u0.shape[1] - 2
        st = O.allocate_state(nx, ny)
        st.fields["u"][:] = u0
        st.fields["v"][:] = s[f"st_{name}_v0"]
        D = s[f"st_{name}_diag"]
        rows = []
        for it in range(D.shape[0]):
            res, diag = O.step(cfg, st, it)
            rows.append([float(diag[k]) if k in diag else float(res[k]) for k in keys])
        assert np.array_equal(np.array(rows), D, equal_nan=True), name
        for f in ("u", "v", "p"):
            assert np.array_equal(st.fields[f], s[f"st_{name}_{f}"][-1], equal_nan=True), (name, f)
        meta = json.loads(str(s[f"st_{name}_meta"]))
        assert st.meta["dt_prev"] == meta["dt_prev"] or (np.isnan(st.meta["dt_prev"]) and meta["dt_prev"] is None
                                                         or np.isnan(meta["dt_prev"]))


def test_c1_cli_stream_matches_oracle(golden_steps):
    """The reference CLI's JSON stream for the 64x64 Re=100 run (config C1)
    equals the oracle's step diagnostics value for value."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "golden_cli_c1.jsonl")
    lines = [json.loads(l) for l in open(path) if l.strip()]
    assert len(lines) == 50
    cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05,
                          Re=100.0, lid_velocity=1.0, lin_tol=1e-8, lin_maxiter=300)
    st = O.allocate_state(64, 64)
    for it, line in enumerate(lines):
        res, diag = O.step(cfg, st, it)
        assert line["iteration"] == it
        assert line["dt"] == diag["dt"] and line["CFL"] == diag["CFL"]
        assert line["residuals"] == {"Ru": res["Ru"], "Rv": res["Rv"], "Rp": 0.0, "continuity": res["continuity"]}
        dg = line["diagnostics"]
        assert dg["Rp_iterations"] == diag["Rp_iterations"] == (0 if it == 0 else 300)
        assert dg["divergence_norm"] == diag["divergence_norm"]
        assert dg["Rp_residual"] == diag["Rp_residual"]
