"""The 64^2 lid-driven cavity run (Re 100, lid 1, 50 steps, lin_tol 1e-8, CFL
0.5, QUICK, lin_maxiter 300) through both drop-in boundaries on the GPU,
compared line by line with the JSON stream the reference's own CLI wrote for
it (tests/golden/golden_cli_c1.jsonl, reference cli.py:147-173 over
drivers/simulation_driver.py:58-110):

* ``SimulationDriver.run`` yields, serialised the way the CLI does;
* ``python -m pyflow run ... --json-stream`` as a subprocess.

Every key of every line is checked. Integer and flag values (iteration,
Rp_iterations, Rp_converged, nan_detected, plateau_Ru, Rp_method, nu, the
centerline's y coordinates) must be equal; floating-point values must lie
within 4x the reference's own summation-order envelope for this run (the
``c1_64`` case of golden_envelope.json, the same trajectory); ``wall_time``
is a clock reading and only has to be increasing. The build adds one key,
``Rp_true_iterations`` (the PCG iterations actually run; the reference's
``Rp_iterations`` is 0, 1 or lin_maxiter).
"""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FACTOR = 4.0
ADDED_KEYS = {"Rp_true_iterations"}


def _golden():
    with open(os.path.join(HERE, "golden", "golden_cli_c1.jsonl")) as f:
        lines = [json.loads(l) for l in f if l.strip()]
    with open(os.path.join(HERE, "golden", "golden_envelope.json")) as f:
        env = json.load(f)["c1_64"]
    return lines, env


def _rel(a, b):
    return abs(a - b) / abs(b) if b != 0 else abs(a - b)


def _compare_stream(got_lines, label):
    exp_lines, env = _golden()
    assert len(got_lines) == len(exp_lines) == 50, (label, len(got_lines))
    worst = {}
    last_wall = -1.0
    for got, exp in zip(got_lines, exp_lines):
        it = exp["iteration"]
        assert set(got) == set(exp), (label, it, set(got) ^ set(exp))
        assert got["type"] == "step" and got["iteration"] == it
        assert got["wall_time"] > last_wall
        last_wall = got["wall_time"]
        for k in ("dt", "CFL", "continuity"):
            worst[k] = max(worst.get(k, 0.0), _rel(got[k], exp[k]))
        assert set(got["residuals"]) == set(exp["residuals"])
        for k in ("Ru", "Rv", "continuity"):
            worst[k] = max(worst.get(k, 0.0), _rel(got["residuals"][k], exp["residuals"][k]))
        assert got["residuals"]["Rp"] == exp["residuals"]["Rp"] == 0.0
        gd, ed = got["diagnostics"], exp["diagnostics"]
        assert set(gd) == set(ed) | ADDED_KEYS, (label, it, set(gd) ^ (set(ed) | ADDED_KEYS))
        for k in ("Rp_iterations", "Rp_converged", "nan_detected", "plateau_Ru", "Rp_method", "nu"):
            assert gd[k] == ed[k], (label, it, k, gd[k], ed[k])
        for k in ("max_residual", "residual_drop_Ru_orders", "divergence_norm"):
            worst[k] = max(worst.get(k, 0.0), _rel(gd[k], ed[k]))
        if ed["Rp_converged"]:
            bound = env["rp_abs_bound"][it]
            assert gd["Rp_residual"] <= bound * (1 + 1e-12) and ed["Rp_residual"] <= bound * (1 + 1e-12)
        else:
            worst["Rp_residual"] = max(worst.get("Rp_residual", 0.0), _rel(gd["Rp_residual"], ed["Rp_residual"]))
        (gy, gu), (ey, eu) = gd["u_centerline"], ed["u_centerline"]
        assert gy == ey and len(gu) == len(eu) == 64
        scale = max(max(abs(x) for x in eu), 1e-300)
        worst["centerline_maxrel"] = max(worst.get("centerline_maxrel", 0.0),
                                         max(abs(a - b) for a, b in zip(gu, eu)) / scale)
        assert gd["Rp_true_iterations"] >= 0
    for k, w in sorted(worst.items()):
        bound = max(1e-12, FACTOR * env[k])
        print(f"[c1 {label}] {k}: rel err {w:.2e} (envelope {env[k]:.2e}, ratio {w / max(env[k], 1e-12):.2f})",
              flush=True)
        assert w <= bound, (label, k, w, env[k])


def test_c1_through_simulation_driver_matches_reference_stream():
    from pyflow.cli import stream_record
    from pyflow.config.model import SimulationConfig
    from pyflow.core.ghost_fields import allocate_state
    from pyflow.drivers.simulation_driver import SimulationDriver
    from pyflow.residuals.manager import ResidualManager
    cfg = SimulationConfig(nx=64, ny=64, Re=100.0, lid_velocity=1.0, cfl_target=0.5, cfl_growth=1.05,
                           advection_scheme="quick", disable_advection=False, lin_tol=1e-8, lin_maxiter=300,
                           diagnostics=False, checkpoint_interval=None, log_path=None, seed=None,
                           turbulence_model=None)
    cfg.force_quiet = True
    state = allocate_state(64, 64)
    got = []
    for st, residuals, diag in SimulationDriver(cfg, state, ResidualManager()).run(max_steps=50):
        assert st is state
        got.append(json.loads(json.dumps(stream_record(residuals, diag, cfg.seed))))
    _compare_stream(got, "driver")


def test_c1_through_cli_json_stream_matches_reference_stream(tmp_path):
    env = dict(os.environ)
    env["PYTHONPATH"] = os.path.join(ROOT, "pyflow-solver_amd")
    env.setdefault("PYFLOW_PRECISION", "fp64")
    cmd = [sys.executable, "-m", "pyflow", "run", "--nx", "64", "--ny", "64", "--re", "100", "--lid-velocity", "1",
           "--steps", "50", "--lin-tol", "1e-8", "--cfl", "0.5", "--json-stream"]
    print("[c1 cli] " + " ".join(cmd[2:]), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True, env=env, cwd=str(tmp_path), timeout=300)
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr[-2000:]
    got = [json.loads(l) for l in proc.stdout.splitlines() if l.startswith("{")]
    assert len([l for l in proc.stdout.splitlines() if l.strip()]) == 50, proc.stdout[-2000:]
    _compare_stream(got, "cli")
