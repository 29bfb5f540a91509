"""GPU checks at the full workload sizes, through properties that do not need
the FP64 oracle to run at those sizes (it would take minutes to hours on the
host): the true FP64 residual of the 4096x4096 pinned Poisson solve, computed
on the host from the downloaded solution, and the 1024x1024 lid-driven cavity
in mixed precision against the FP64 device path (itself held to the oracle
and the reference's golden trajectories on smaller grids)."""
import numpy as np
import pytest

from oracle import pyflow_oracle as O

pytestmark = pytest.mark.gpu


def test_poisson_4096_true_residual():
    """The 4096x4096 pinned-Neumann solve (right-hand side N(0,1), seed 0, mean
    removed, cell 0 pinned) reaches ||b - A x|| <= 1e-10 ||b|| on the host's
    FP64 operator (pressure_solver.py:97-106), not only on the device's own
    residual."""
    from pyflow.linear_solvers.interface import PinnedNegLaplacian, device_cg
    n, tol = 4096, 1e-10
    b = np.random.default_rng(0).standard_normal((n, n))
    b -= b.mean()
    b[0, 0] = 0.0
    b = b.reshape(-1)
    x, info, its, rnorm, conv = device_cg(PinnedNegLaplacian(n, n, 1.0, 1.0), b, tol, 400_000, jacobi=True,
                                          precision="mixed")
    bn = float(np.linalg.norm(b))
    true = float(np.linalg.norm(b - O.neg_laplacian_pinned(x, n, n, 1.0, 1.0)))
    print(f"[poisson 4096] iterations={its} device ||r||/||b||={rnorm / bn:.3e} host {true / bn:.3e}", flush=True)
    assert info == 0 and conv and np.all(np.isfinite(x))
    assert rnorm <= tol * bn
    assert true <= tol * bn * 1.0001


def _pressure_envelope(n, step):
    """The pressure freedom the reference's own stopping rule leaves at this
    grid and step, from tests/golden/pressure_envelope.json (written by the
    FP64 oracle): the largest relative L2 change of p, and of dt * grad p, of
    a trajectory the reference accepts (every solve without the Jacobi
    preconditioner, or every solve to 1e-13) against the reference's
    trajectory, and of p solved exactly from the reference's own velocity."""
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "pressure_envelope.json")
    with open(path) as f:
        g = json.load(f)["grids"][str(n)]
    traj, tol = g["trajectory"][step], g["tolerance"][step]
    alts = [v for k, v in traj.items() if isinstance(v, dict) and k in ("exact", "no_jacobi")]
    p_env = max([a["p_rel_l2"] for a in alts] + [tol["p_rel_l2_ref_vs_exact"]])
    g_env = max([a["corr_rel_l2"] for a in alts] + [tol["corr_rel_l2_ref_vs_exact"]])
    return p_env, g_env


def test_cavity_1024_mixed_matches_fp64():
    """Three projection steps of the 1024x1024 cavity (Re 1000, U 1, QUICK,
    tol 1e-6) in mixed precision against the FP64 device path: the same step
    sizes, every solve converged, continuity to 1e-3, u and v within 1e-5,
    and p and grad p within 4x the pressure envelope of the reference itself.

    A solve is held only to ||b - A p|| <= 1e-6 ||b||, and along a trajectory
    the next step's pressure absorbs the gradient part of the velocity that
    this leaves free; the oracle measures how far p and dt * grad p of other
    trajectories the reference accepts move (1024^2, step 3: 6.4e-4 and
    1.9e-5). Two accepted solutions may each sit that far from a third, so
    the bound is 4x. Without the deflation step, FP32 rounding left a
    near-constant error along the pinned operator's slowest mode in the first
    step's pressure (2.7e-4 relative), and step 3's p, a hundred times
    smaller, inherited it as a 5.4% difference."""
    from pyflow._device import DeviceBatch, cavity_params
    n, steps = 1024, 3
    out = {}
    for prec in ("fp64", "mixed"):
        p = cavity_params(n, n, dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0, Re=1000.0, cfl_target=0.5,
                          cfl_growth=1.05, lin_tol=1e-6, lin_maxiter=10000, jacobi=True)
        b = DeviceBatch([p], prec)
        ds = [b.run(1)[0] for _ in range(steps)]
        u = np.zeros((n + 2, n + 2)); v = np.zeros_like(u); pr = np.zeros_like(u)
        b.download(0, u, v, pr)
        b.close()
        for d in ds:
            assert d.steps_done == 1 and not d.nan_detected and d.rp_converged and d.rp_info == 0
            assert d.rp_residual <= 1e-6 * d.bnorm
        out[prec] = (ds, u, v, pr)
    d64, u64, v64, p64 = out["fp64"]
    dmx, umx, vmx, pmx = out["mixed"]
    du = float(np.max(np.abs(umx - u64)))
    dv = float(np.max(np.abs(vmx - v64)))
    pi64, pimx = p64[1:-1, 1:-1], pmx[1:-1, 1:-1]
    dp = float(np.linalg.norm(pimx - pi64) / np.linalg.norm(pi64))
    g64 = np.concatenate([g.ravel() for g in O.gradient_coherent(pi64, 1.0, 1.0)])
    gmx = np.concatenate([g.ravel() for g in O.gradient_coherent(pimx, 1.0, 1.0)])
    dg = float(np.linalg.norm(gmx - g64) / np.linalg.norm(g64))
    p_env, g_env = _pressure_envelope(n, steps - 1)
    print(f"[cavity 1024] max|du|={du:.3e} max|dv|={dv:.3e} rel|dp|={dp:.3e} (envelope {p_env:.3e}, "
          f"ratio {dp / p_env:.2f}) rel|d grad p|={dg:.3e} (envelope {g_env:.3e}, ratio {dg / g_env:.2f}) "
          f"continuity {dmx[-1].continuity:.6e} vs {d64[-1].continuity:.6e} "
          f"iterations {[d.rp_iterations for d in dmx]} vs {[d.rp_iterations for d in d64]}", flush=True)
    for a, b in zip(dmx, d64):
        assert abs(a.dt - b.dt) <= 1e-6 * b.dt
    assert du <= 1e-5 and dv <= 1e-5
    assert abs(dmx[-1].continuity - d64[-1].continuity) <= 1e-3 * abs(d64[-1].continuity) + 1e-12
    assert dp <= 4.0 * p_env
    assert dg <= 4.0 * g_env
