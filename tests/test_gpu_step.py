"""GPU parity of the full projection step against the reference's own outputs
(golden trajectories written by running the reference) and the reference's
behavioural tests restated against the GPU drop-in."""
import io
import json
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import pyflow_oracle as O

pytestmark = pytest.mark.gpu

# FP64 parity bound. The stencils are bit-identical to the reference; only the
# reduction order of dot products, norms and the RHS mean differs, and CG
# amplifies last-bit differences. tests/golden/golden_envelope.json holds, per
# golden case and quantity, the largest change of the reference trajectory
# itself over six other summation orders (exactly rounded, running sums
# forwards and backwards, 64-element blocks, two random orders). The GPU must
# stay within ENVELOPE_FACTOR times that, and u, v also within the 1e-5
# relative field bound (or twice the case's own envelope where the reference
# itself moves by more than 5e-6).
# Rp_residual, the true residual |b - A x| at the end of the solve, is set by
# where CG stops: for a converged solve it lies anywhere below lin_tol*|b|
# (its envelope is 30%-140% relative), so there both the reference's and the
# GPU's values must lie in [0, lin_tol*|b|] -- their difference is bounded by
# the stopping threshold, not by a relative envelope. A solve that ran to
# lin_maxiter is held to the relative envelope like every other quantity.
ENVELOPE_FACTOR = 4.0
FIELD_UV_CAP = 1e-5
FLOOR_REL = 1e-12

STEP_KEYS_FLOAT = ["dt", "CFL", "Ru", "Rv", "continuity", "Rp_residual", "divergence_norm", "max_residual",
                   "residual_drop_Ru_orders", "nu"]


def _envelope():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "golden_envelope.json")
    with open(path) as f:
        return json.load(f)


def _run_case(golden_steps, name, precision="fp64"):
    from pyflow.core.ghost_fields import allocate_state
    from pyflow.residuals.manager import ResidualManager
    from pyflow.solvers.solver import step
    s = golden_steps
    cfg = SimpleNamespace(precision=precision, **json.loads(str(s[f"st_{name}_cfg"])))
    u0 = s[f"st_{name}_u0"]
    ny, nx = u0.shape[0] - 2, u0.shape[1] - 2
    st = allocate_state(nx, ny)
    st.fields["u"][:] = u0
    st.fields["v"][:] = s[f"st_{name}_v0"]
    tr = ResidualManager()
    rows = []
    for it in range(s[f"st_{name}_diag"].shape[0]):
        st, res, diag = step(cfg, st, tr, it)
        rows.append({**res, **diag})
    return st, rows


def _rel(a, b):
    scale = max(float(np.nanmax(np.abs(b))), 1e-300)
    return float(np.nanmax(np.abs(a - b))) / scale


@pytest.mark.parametrize("name", ["quick8", "upwind16", "quick12x10", "safeguard8", "noadv12x10", "nolid9x7",
                                  "c1_64"])
def test_step_trajectory_matches_reference_fp64(golden_steps, name):
    s = golden_steps
    keys = json.loads(str(s["step_keys"]))
    env = _envelope()[name]
    print(f"[step] case={name}", flush=True)
    st, rows = _run_case(golden_steps, name)
    D = s[f"st_{name}_diag"]
    worst = {}
    rp_abs = 0.0  # largest |delta Rp_residual| / (lin_tol |b|) over converged steps
    for it, row in enumerate(rows):
        exp = dict(zip(keys, D[it]))
        assert row["Rp_iterations"] == int(exp["Rp_iterations"]), (it, row["Rp_iterations"], exp["Rp_iterations"])
        assert row["Rp_converged"] == bool(exp["Rp_converged"])
        assert row["nan_detected"] == bool(exp["nan_detected"])
        for k in STEP_KEYS_FLOAT:
            e = exp[k]
            if k == "Rp_residual" and bool(exp["Rp_converged"]):
                bound = env["rp_abs_bound"][it]
                assert 0.0 <= row[k] <= bound * (1 + 1e-12) and e <= bound * (1 + 1e-12), (name, it, row[k], e, bound)
                if bound > 0:
                    rp_abs = max(rp_abs, abs(row[k] - e) / bound)
                continue
            rel = abs(row[k] - e) / abs(e) if e != 0 else abs(row[k] - e)
            worst[k] = max(worst.get(k, 0.0), rel)
    print(f"[step] {name} Rp_residual (converged steps): |delta| / (lin_tol |b|) {rp_abs:.2e} (bound 1)", flush=True)
    for k, w in worst.items():
        tol = max(FLOOR_REL, ENVELOPE_FACTOR * env[k])
        print(f"[step] {name} {k}: rel err {w:.2e} (envelope {env[k]:.2e}, ratio {w / max(env[k], FLOOR_REL):.2f})",
              flush=True)
        assert w <= tol, (name, k, w, env[k])
    for f in ("u", "v", "p"):
        ref = s[f"st_{name}_{f}"][-1]
        got = st.fields[f]
        assert got.shape == ref.shape
        assert np.array_equal(got[0, :], ref[0, :]) and np.array_equal(got[:, 0], ref[:, 0])  # ghost ring
        maxrel = _rel(got, ref)
        l2rel = float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))
        print(f"[step] {name} field {f}: max rel {maxrel:.2e} (envelope {env[f'field_{f}_maxrel']:.2e}, ratio "
              f"{maxrel / max(env[f'field_{f}_maxrel'], FLOOR_REL):.2f}), l2 rel {l2rel:.2e} (envelope "
              f"{env[f'field_{f}_l2rel']:.2e}, ratio {l2rel / max(env[f'field_{f}_l2rel'], FLOOR_REL):.2f})",
              flush=True)
        assert maxrel <= max(FLOOR_REL, ENVELOPE_FACTOR * env[f"field_{f}_maxrel"]), (name, f, maxrel)
        assert l2rel <= max(FLOOR_REL, ENVELOPE_FACTOR * env[f"field_{f}_l2rel"]), (name, f, l2rel)
        if f in ("u", "v"):
            assert maxrel <= max(FIELD_UV_CAP, 2.0 * env[f"field_{f}_maxrel"]), (name, f, maxrel)


def test_nan_propagation_matches_reference_fixture(golden_steps):
    """NaN in u spreads exactly like the reference's emergency checkpoint:
    all interior u except the lid row, all of v and p."""
    st, rows = _run_case(golden_steps, "nan10")
    assert rows[0]["nan_detected"] is True
    assert rows[0]["Rp_iterations"] == 30
    for f, exp in (("u", 90), ("v", 100), ("p", 100)):
        assert int(np.isnan(st.fields[f]).sum()) == exp
    assert np.array_equal(np.isnan(st.fields["u"]), np.isnan(golden_steps["st_nan10_u"][-1]))
    assert np.all(st.fields["u"][-2, 1:-1] == 0.0)


def test_nan_detection_logging_and_emergency_checkpoint(tmp_path, monkeypatch):
    """tests/test_nan_detection_logging.py:10-22 restated."""
    from pyflow.core.ghost_fields import allocate_state, interior_view
    from pyflow.drivers.simulation_driver import SimulationDriver
    from pyflow.residuals.manager import ResidualManager
    monkeypatch.chdir(tmp_path)
    buf = io.StringIO()
    cfg = SimpleNamespace(disable_advection=True, advection_scheme="upwind", cfl_target=0.5, cfl_growth=1.05,
                          Re=100.0, lid_velocity=0.0, lin_tol=1e-10, lin_maxiter=30, nx=10, ny=10, lx=9, ly=9,
                          diagnostics=False, log_stream=buf)
    state = allocate_state(cfg.nx, cfg.ny)
    interior_view(state.fields["u"])[2, 2] = np.nan
    for _st, _res, diag in SimulationDriver(cfg, state, ResidualManager()).run(max_steps=1):
        assert diag.get("nan_detected") is True
    lines = [l for l in buf.getvalue().splitlines() if l.strip()]
    assert any('"reason":"nan_detected"' in l for l in lines)
    assert (tmp_path / "emergency_FAIL_NAN_it0.npz").exists()


def test_driver_yields_same_state_and_centerline():
    from pyflow.core.ghost_fields import allocate_state
    from pyflow.drivers.simulation_driver import SimulationDriver
    from pyflow.residuals.manager import ResidualManager

    class Cfg:
        disable_advection = False
        advection_scheme = "quick"
        cfl_target = 0.5
        cfl_growth = 1.05
        Re = 100.0
        lid_velocity = 1.0
        lin_tol = 1e-10
        lin_maxiter = 50
        nx = 8
        ny = 8
        lx = nx - 1
        ly = ny - 1
        diagnostics = False

    cfg = Cfg()
    state = allocate_state(cfg.nx, cfg.ny)
    last_dt = None
    n = 0
    for st, residuals, diag in SimulationDriver(cfg, state, ResidualManager()).run(max_steps=5):
        assert st is state
        if last_dt is not None:
            assert diag["dt"] <= last_dt * 1.5
        last_dt = diag["dt"]
        yc, uc = diag["u_centerline"]
        assert len(yc) == len(uc) == 8
        assert uc == [float(x) for x in state.fields["u"][1:-1, 1:-1][:, 4]]
        n += 1
    assert n == 5


def test_lid_bc_and_divergence_reduced():
    """tests/test_projection_method_contract.py:186-268 restated."""
    from pyflow.core.ghost_fields import allocate_state, interior_view
    from pyflow.numerics.fluid_ops import divergence
    from pyflow.residuals.manager import ResidualManager
    from pyflow.solvers.solver import step

    class Cfg:
        disable_advection = False
        advection_scheme = "quick"
        cfl_target = 0.5
        cfl_growth = 1.1
        Re = 100.0
        lid_velocity = 1.0
        lin_tol = 1e-10
        lin_maxiter = 200
        diagnostics = False

    state = allocate_state(8, 8)
    u = np.zeros((10, 10))
    u[3:6, 3:6] = 1.0
    state.fields["u"][:] = u
    before = np.linalg.norm(divergence(interior_view(state.fields["u"]), interior_view(state.fields["v"]), 1.0, 1.0))
    state, residuals, diagnostics = step(Cfg(), state, ResidualManager(), 0)
    after = np.linalg.norm(divergence(interior_view(state.fields["u"]), interior_view(state.fields["v"]), 1.0, 1.0))
    assert after < before
    assert np.allclose(state.fields["u"][-2, 1:-1], 1.0)


def test_safeguard_and_seed_reproducibility_bitwise():
    from pyflow.core.ghost_fields import allocate_state
    from pyflow.residuals.manager import ResidualManager
    from pyflow.solvers.solver import step

    def run():
        cfg = SimpleNamespace(nx=16, ny=16, Re=100.0, lid_velocity=1.0, cfl_target=0.5, cfl_growth=1.05,
                              advection_scheme="upwind", disable_advection=False, lin_tol=1e-10, lin_maxiter=50,
                              seed=42, diagnostics=False)
        st = allocate_state(16, 16)
        tr = ResidualManager()
        for it in range(5):
            st, _r, d = step(cfg, st, tr, it)
        return st.fields["u"].copy(), st.fields["v"].copy(), st.fields["p"].copy(), d

    a, b = run(), run()
    for k in range(3):
        assert np.array_equal(a[k], b[k])
    assert a[3]["seed"] == 42


def test_restart_is_bitwise(tmp_path):
    """Checkpoint after N steps, restart, run M more: identical to N+M straight
    (the dt controller's memory is persisted in the checkpoint meta)."""
    from pyflow.core.ghost_fields import allocate_state
    from pyflow.drivers.simulation_driver import SimulationDriver
    from pyflow.io.checkpoint import load_checkpoint, save_checkpoint
    from pyflow.residuals.manager import ResidualManager
    for precision in ("fp64", "mixed"):
        cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05,
                              Re=100.0, lid_velocity=1.0, lin_tol=1e-9, lin_maxiter=200, nx=20, ny=18,
                              diagnostics=False, precision=precision)
        full = allocate_state(20, 18)
        for _ in SimulationDriver(cfg, full, ResidualManager()).run(max_steps=9):
            pass
        part = allocate_state(20, 18)
        for _ in SimulationDriver(cfg, part, ResidualManager()).run(max_steps=5):
            pass
        ck = str(tmp_path / f"ck_{precision}.npz")
        save_checkpoint(ck, part, 4, 0.0, cfg)
        loaded, meta = load_checkpoint(ck)
        for _ in SimulationDriver(cfg, loaded, ResidualManager()).run(max_steps=9, start_iteration=5):
            pass
        for f in ("u", "v", "p"):
            assert np.array_equal(loaded.fields[f], full.fields[f]), (precision, f)
        assert loaded.meta["dt_prev"] == full.meta["dt_prev"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_restart_is_bitwise_under_another_environment(tmp_path, monkeypatch, precision):
    """The checkpoint pins the tiling: resumed in a process whose PYFLOW_*
    settings pick another region height, workgroup count and (MIXED) one PCG
    iteration per round, the run still repeats the uninterrupted one bit for
    bit. Without the pinned config the same resume gives other bits."""
    from pyflow.core.ghost_fields import allocate_state
    from pyflow.drivers.simulation_driver import SimulationDriver
    from pyflow.io.checkpoint import load_checkpoint, save_checkpoint
    from pyflow.residuals.manager import ResidualManager
    nx, ny = 200, 150
    cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05,
                          Re=100.0, lid_velocity=1.0, lin_tol=1e-9, lin_maxiter=2000, nx=nx, ny=ny,
                          diagnostics=False, precision=precision)
    for k in ("PYFLOW_TPB", "PYFLOW_MAXWG", "PYFLOW_CA", "PYFLOW_XD"):
        monkeypatch.delenv(k, raising=False)
    full = allocate_state(nx, ny)
    for _ in SimulationDriver(cfg, full, ResidualManager()).run(max_steps=6):
        pass
    part = allocate_state(nx, ny)
    for _ in SimulationDriver(cfg, part, ResidualManager()).run(max_steps=3):
        pass
    kc = dict(part.meta["kernel_config"])
    assert kc["precision"] == precision and kc["tile_rows"] == 4 and kc["max_workgroups"] == 12
    ck = str(tmp_path / "ck.npz")
    save_checkpoint(ck, part, 2, 0.0, cfg)
    monkeypatch.setenv("PYFLOW_TPB", "1")
    monkeypatch.setenv("PYFLOW_MAXWG", "7")
    monkeypatch.setenv("PYFLOW_CA", "0")
    monkeypatch.setenv("PYFLOW_XD", "1")
    loaded, meta = load_checkpoint(ck)
    assert meta["kernel_config"] == kc and loaded.meta["kernel_config"] == kc
    for _ in SimulationDriver(cfg, loaded, ResidualManager()).run(max_steps=6, start_iteration=3):
        pass
    assert loaded.meta["kernel_config"] == kc
    for f in ("u", "v", "p"):
        assert np.array_equal(loaded.fields[f], full.fields[f]), (precision, f)
    assert loaded.meta["dt_prev"] == full.meta["dt_prev"]
    unpinned, _ = load_checkpoint(ck)
    del unpinned.meta["kernel_config"]
    for _ in SimulationDriver(cfg, unpinned, ResidualManager()).run(max_steps=6, start_iteration=3):
        pass
    assert unpinned.meta["kernel_config"]["tile_rows"] == 1
    same = all(np.array_equal(unpinned.fields[f], full.fields[f]) for f in ("u", "v", "p"))
    print(f"[restart] {precision}: unpinned resume bitwise equal = {same}", flush=True)
    if precision == "mixed":
        assert not same


def test_preconditioner_effect_and_manufactured_projection():
    """tests/test_preconditioner_effect.py:9-33 and test_projection_coherent.py:11-45 restated."""
    from pyflow.core.ghost_fields import allocate_state, interior_view
    from pyflow.numerics.core_ops_coherent import divergence_coherent
    from pyflow.solvers.pressure_solver import solve_pressure_poisson
    nx, ny = 24, 20
    out = {}
    for pc in (True, False):
        st = allocate_state(nx, ny)
        interior_view(st.fields["u"])[ny // 3, nx // 3] = 1.0
        cfg = SimpleNamespace(nx=nx, ny=ny, lin_tol=1e-10, lin_maxiter=300, diagnostics=False, enable_jacobi_pc=pc)
        out[pc] = solve_pressure_poisson(st, 0.01, 1.0, 1.0, cfg)
    assert np.max(np.abs(out[True][0] - out[False][0])) / np.max(np.abs(out[True][0])) < 1e-9
    # The reference's Jacobi puts 1.0 on boundary cells instead of A's diagonal,
    # which here makes it slower than no preconditioning: the oracle takes 189 vs
    # 143 iterations and the device reproduces both counts (the reference test's
    # "fewer iterations" check only sees Rp_iterations, which is 1 for both).
    assert out[True][1]["Rp_true_iterations"] == 189 and out[False][1]["Rp_true_iterations"] == 143
    assert out[True][1]["Rp_iterations"] == out[False][1]["Rp_iterations"] == 1
    assert np.linalg.norm(out[True][0]) > 0
    for nx, ny in ((8, 8), (12, 10)):
        dx, dy = 1.0 / (nx - 1), 1.0 / (ny - 1)
        X, Y = np.meshgrid(np.linspace(0, 1, nx), np.linspace(0, 1, ny))
        dt = 0.01
        st = allocate_state(nx, ny)
        interior_view(st.fields["u"])[:] = dt * (-np.pi * np.sin(np.pi * X) * np.cos(np.pi * Y))
        interior_view(st.fields["v"])[:] = dt * (-np.pi * np.cos(np.pi * X) * np.sin(np.pi * Y))
        n0 = np.linalg.norm(divergence_coherent(interior_view(st.fields["u"]), interior_view(st.fields["v"]), dx, dy))
        cfg = SimpleNamespace(lin_tol=1e-12, lin_maxiter=500, diagnostics=False)
        p, diag = solve_pressure_poisson(st, dt, dx, dy, cfg)
        n1 = np.linalg.norm(divergence_coherent(interior_view(st.fields["u"]), interior_view(st.fields["v"]), dx, dy))
        assert n1 < max(1e-10, n0 * 5e-6)
        pe = np.cos(np.pi * X) * np.cos(np.pi * Y)
        pe -= pe.flat[0]
        assert np.linalg.norm(p - pe) / np.linalg.norm(pe) < 0.25


def test_pressure_solver_matches_reference_golden(golden_ops):
    from pyflow.core.ghost_fields import allocate_state
    from pyflow.solvers.pressure_solver import solve_pressure_poisson
    g = golden_ops
    st = allocate_state(13, 11)
    st.fields["u"][:] = g["ps_rand_u0"]
    st.fields["v"][:] = g["ps_rand_v0"]
    cfg = SimpleNamespace(lin_tol=1e-12, lin_maxiter=400, diagnostics=False)
    p, d = solve_pressure_poisson(st, 0.05, 0.1, 0.2, cfg)
    assert _rel(p, g["ps_rand_p"]) < 1e-9
    assert _rel(st.fields["u"], g["ps_rand_u"]) < 1e-9
    exp = g["ps_rand_diag"]
    assert d["Rp_iterations"] == int(exp[0]) and d["Rp_converged"] == bool(exp[2])
    assert abs(d["divergence_norm_initial"] - exp[5]) <= 1e-12 * exp[5]
    st = allocate_state(8, 8)
    cfg = SimpleNamespace(lin_tol=1e-10, lin_maxiter=500, diagnostics=False)
    p, d = solve_pressure_poisson(st, 1.0, 1 / 7, 1 / 7, cfg, rhs_override=g["ps_mms_rhs"])
    assert _rel(p, g["ps_mms_p"]) < 1e-9
    assert d["divergence_norm_initial"] is None and d["divergence_reduction_factor"] is None
    X, Y = np.meshgrid(np.linspace(0, 1, 8), np.linspace(0, 1, 8))
    assert np.allclose(p[1:-1, 1:-1], (np.sin(np.pi * X) * np.sin(np.pi * Y))[1:-1, 1:-1], atol=5e-2)


def test_mixed_precision_step_against_fp64_oracle():
    """Mixed precision (FP32 fields/stencils/SpMV, FP64 reductions + refinement)
    against the FP64 oracle: 12 QUICK steps on 96x96, tight tolerance."""
    from pyflow.core.ghost_fields import allocate_state
    from pyflow.residuals.manager import ResidualManager
    from pyflow.solvers.solver import step
    n = 96
    cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05,
                          Re=1000.0, lid_velocity=1.0, lin_tol=1e-8, lin_maxiter=5000, diagnostics=False,
                          precision="mixed")
    st = allocate_state(n, n)
    ref = O.allocate_state(n, n)
    tr = ResidualManager()
    for it in range(12):
        st, res, diag = step(cfg, st, tr, it)
        rres, rdiag = O.step(cfg, ref, it)
        assert diag["Rp_converged"] == rdiag["Rp_converged"]
        assert abs(diag["dt"] - rdiag["dt"]) <= 1e-6 * rdiag["dt"]
    du = float(np.max(np.abs(st.fields["u"] - ref.fields["u"])))
    dv = float(np.max(np.abs(st.fields["v"] - ref.fields["v"])))
    dp = float(np.linalg.norm(st.fields["p"] - ref.fields["p"]) / np.linalg.norm(ref.fields["p"]))
    print(f"[mixed] max|du|={du:.3e} max|dv|={dv:.3e} rel|dp|={dp:.3e} "
          f"cont={res['continuity']:.6e} oracle={rres['continuity']:.6e}", flush=True)
    assert du <= 1e-5 and dv <= 1e-5 and dp <= 1e-4
    assert abs(res["continuity"] - rres["continuity"]) <= 1e-6


def test_batch_alone_equals_in_mix():
    """Per-cavity reductions depend only on that cavity: a cavity advanced
    alone and inside a mixed batch of other sizes gives identical bits."""
    from pyflow._device import DeviceBatch, cavity_params
    mk = lambda n, re: cavity_params(n, n, dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0, Re=re,  # noqa: E731
                                     lin_tol=1e-6, lin_maxiter=3000)
    for prec in ("fp64", "mixed"):
        alone = DeviceBatch([mk(48, 300.0)], prec)
        da = alone.run(6)
        ua = np.zeros((50, 50)); alone.download(0, ua, None, None)
        mix = DeviceBatch([mk(32, 100.0), mk(48, 300.0), mk(64, 5000.0), mk(96, 1000.0)], prec)
        dm = mix.run(6)
        um = np.zeros((50, 50)); mix.download(1, um, None, None)
        assert np.array_equal(ua, um), prec
        assert da[0].rp_iterations == dm[1].rp_iterations and da[0].continuity == dm[1].continuity
        assert all(d.steps_done == 6 for d in dm.values())
        alone.close()
        mix.close()


def test_continuous_batching_matches_standalone():
    """Cavities admitted into a running batch, advanced in chunks next to
    others at different points of their steps, retired when done and re-armed
    for more steps give the same bits as each cavity run alone."""
    from pyflow._device import DeviceBatch, cavity_params
    mk = lambda n, re: cavity_params(n, n, dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0, Re=re,  # noqa: E731
                                     lin_tol=1e-6, lin_maxiter=3000)
    pa, pb = mk(40, 400.0), mk(24, 100.0)
    for prec in ("fp64", "mixed"):
        ref_a = DeviceBatch([pa], prec)
        ref_a.run(7)
        ua_ref = np.zeros((42, 42)); ref_a.download(0, ua_ref, None, None)
        ref_b = DeviceBatch([pb], prec)
        db_ref = ref_b.run(3)[0]
        ub_ref = np.zeros((26, 26)); ref_b.download(0, ub_ref, None, None)
        srv = DeviceBatch(None, prec, max_cav=4, max_tiles=64)
        sa = srv.add(pa)
        srv.start(sa, 5)
        done, sb, chunks, retired_b = set(), None, 0, None
        while chunks < 100000:
            srv.advance(7)  # an odd chunk length: admissions land on both round parities
            while not srv.poll():
                pass
            chunks += 1
            if chunks == 3:
                sb = srv.add(pb)
                srv.start(sb, 3)
            for s in srv.finished():
                if s == sa and "a5" not in done:
                    done.add("a5")
                    srv.start(sa, 2)  # continue the same cavity for two more steps
                elif s == sa and "a5" in done:
                    done.add("a7")
                elif s == sb and retired_b is None:
                    retired_b = srv.diag()[sb]
                    ub = np.zeros((26, 26)); srv.download(sb, ub, None, None)
                    srv.remove(sb)
                    done.add("b")
            if "a7" in done and "b" in done:
                break
        assert "a7" in done and "b" in done, (prec, chunks)
        ua = np.zeros((42, 42)); srv.download(sa, ua, None, None)
        assert np.array_equal(ua, ua_ref), prec
        assert np.array_equal(ub, ub_ref), prec
        assert retired_b.rp_iterations == db_ref.rp_iterations and retired_b.continuity == db_ref.continuity
        for b in (ref_a, ref_b, srv):
            b.close()


def test_uneven_batch_equals_alone_mixed():
    """Regression: eight cavities of uneven sizes (partial tile columns, one to
    49 regions each) advanced together in mixed precision give the same bits
    as each alone. Streaming stores in the PCG round once broke this."""
    from pyflow._device import DeviceBatch, cavity_params
    sizes = [448, 352, 288, 160, 160, 32, 64, 32]
    res = [1200.0, 300.0, 2500.0, 150.0, 900.0, 4000.0, 100.0, 700.0]
    mk = lambda n, re: cavity_params(n, n, dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0, Re=re,  # noqa: E731
                                     lin_tol=1e-6, lin_maxiter=10000)
    ps = [mk(n, re) for n, re in zip(sizes, res)]
    mix = DeviceBatch(ps, "mixed")
    dm = mix.run(5)
    for k, (n, p) in enumerate(zip(sizes, ps)):
        alone = DeviceBatch([p], "mixed")
        da = alone.run(5)[0]
        ua = np.zeros((n + 2, n + 2)); alone.download(0, ua, None, None)
        um = np.zeros((n + 2, n + 2)); mix.download(k, um, None, None)
        alone.close()
        assert np.array_equal(ua, um), (k, n)
        assert da.rp_iterations_run == dm[k].rp_iterations_run, (k, n)
    mix.close()


def test_batch_more_workgroups_than_cus_mixed():
    """Regression: a mixed-precision batch whose rounds hold more workgroups
    than the GPU has CUs (sixteen cavities of 32..512 cells a side, 400+
    regions), so some workgroups of a round run after others have finished,
    gives each cavity the same bits as alone."""
    from pyflow._device import DeviceBatch, cavity_params
    sizes = [448, 352, 288, 160, 160, 32, 64, 32, 96, 448, 352, 480, 288, 320, 512, 384]
    mk = lambda n, k: cavity_params(n, n, dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0,  # noqa: E731
                                    Re=150.0 + 97.0 * k, lin_tol=1e-6, lin_maxiter=10000)
    ps = [mk(n, k) for k, n in enumerate(sizes)]
    mix = DeviceBatch(ps, "mixed")
    dm = mix.run(2)
    assert not any(d.nan_detected for d in dm.values())
    for k in (0, 3, 8, 11, 14):
        n = sizes[k]
        alone = DeviceBatch([ps[k]], "mixed")
        da = alone.run(2)[0]
        ua = np.zeros((n + 2, n + 2)); alone.download(0, ua, None, None)
        um = np.zeros((n + 2, n + 2)); mix.download(k, um, None, None)
        alone.close()
        assert np.array_equal(ua, um), (k, n)
        assert da.rp_iterations_run == dm[k].rp_iterations_run, (k, n)
    mix.close()


@pytest.mark.parametrize("precision,cap", [("fp64", 1e-8), ("mixed", 1e-5)])
def test_several_regions_per_workgroup(precision, cap, monkeypatch):
    """Workgroups that own several regions (a cavity capped at two workgroups:
    130x97 cells = 6 regions, 4 + 2 per workgroup) follow the FP64 oracle."""
    from pyflow._device import DeviceBatch, cavity_params
    monkeypatch.setenv("PYFLOW_MAXWG", "2")
    nx, ny = 130, 97
    p = cavity_params(nx, ny, dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0, Re=400.0, lin_tol=1e-10,
                      lin_maxiter=5000)
    b = DeviceBatch([p], precision)
    d = b.run(6)[0]
    u = np.zeros((ny + 2, nx + 2)); v = np.zeros_like(u)
    b.download(0, u, v, None)
    b.close()
    cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05,
                          Re=400.0, lid_velocity=1.0, lin_tol=1e-10, lin_maxiter=5000)
    ref = O.allocate_state(nx, ny)
    for it in range(6):
        O.step(cfg, ref, it)
    assert d.steps_done == 6 and not d.nan_detected
    assert float(np.max(np.abs(u - ref.fields["u"]))) <= cap
    assert float(np.max(np.abs(v - ref.fields["v"]))) <= cap


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_equal_direct_launches(precision):
    """Rounds queued as hipGraph replays (16- and 6-round chains, the rest of
    each chunk launched directly) give the same bits as one launch per round,
    over several runs of the same batch with host work in between."""
    from pyflow._device import DeviceBatch, cavity_params
    sizes, res = [200, 96, 40], [1000.0, 300.0, 2000.0]
    ps = [cavity_params(n, n, dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0, Re=re, lin_tol=1e-6,
                        lin_maxiter=5000) for n, re in zip(sizes, res)]
    outs = []
    for glen in (0, 16, 6):
        b = DeviceBatch(ps, precision)
        b.set_graph(glen)
        runs = []
        for nsteps in (2, 1, 3):  # three runs: each replays the graphs of the same exec
            d = b.run(nsteps)
            fields = []
            for k, n in enumerate(sizes):
                u = np.zeros((n + 2, n + 2)); v = np.zeros_like(u); p = np.zeros_like(u)
                b.download(k, u, v, p)
                fields.append((u, v, p))
            runs.append((fields, [(d[k].rp_iterations_run, d[k].continuity, d[k].dt) for k in range(len(sizes))]))
        b.close()
        outs.append(runs)
    for glen_runs in outs[1:]:
        for (f0, s0), (f1, s1) in zip(outs[0], glen_runs):
            assert s0 == s1
            for a, c in zip(f0, f1):
                for x, y in zip(a, c):
                    assert np.array_equal(x, y)
