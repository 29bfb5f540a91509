"""FP64 cavities of one region run the CG iterations of a solve inside one
round (PfCav::local, kernel function local_cg). That path must give the bits
of one round per P or Q phase (PYFLOW_LOCAL=0): fields, diagnostics and
iteration counts, on steps and on Poisson solves of both operators."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _with_local(flag, make):
    old = os.environ.get("PYFLOW_LOCAL")
    os.environ["PYFLOW_LOCAL"] = "1" if flag else "0"  # read when a cavity is added
    try:
        return make()
    finally:
        if old is None:
            del os.environ["PYFLOW_LOCAL"]
        else:
            os.environ["PYFLOW_LOCAL"] = old


def _steps(nx, ny, rows, local, nsteps, scheme="quick", re=100.0, tol=1e-8, maxiter=300, u0=None):
    from pyflow._device import DeviceBatch, cavity_params

    def make():
        p = cavity_params(nx, ny, dx=1.0 / (nx - 1), dy=1.0 / (ny - 1), scheme=scheme, lid_velocity=1.0, Re=re,
                          lin_tol=tol, lin_maxiter=maxiter)
        p.tile_rows = rows
        return DeviceBatch([p], "fp64")

    b = _with_local(local, make)
    try:
        if u0 is not None:
            b.upload(0, u0, np.zeros_like(u0), np.zeros_like(u0))
        diags = [b.run(1)[0] for _ in range(nsteps)]
        u = np.zeros((ny + 2, nx + 2)); v = np.zeros_like(u); p = np.zeros_like(u)
        b.download(0, u, v, p)
        rounds = b.rounds()
    finally:
        b.close()
    keys = ("dt", "rp_iterations", "rp_info", "rp_residual", "div_norm", "ru", "rv", "continuity")
    return (u, v, p), [tuple(getattr(d, k) for k in keys) for d in diags], rounds


@pytest.mark.parametrize("nx,ny,rows", [(64, 64, 4), (40, 24, 4), (64, 96, 8), (17, 9, 1)])
def test_local_cg_steps_bitwise(nx, ny, rows):
    """Ten QUICK steps (C1's lid-driven cavity at Re 100, lin_tol 1e-8):
    the in-round CG gives the bits of the round-per-phase CG in u, v, p and
    in every step's diagnostics, in far fewer rounds. 64x96 with 8-tile
    regions is the largest cavity that takes the path (6144 cells)."""
    print(f"[local] steps {nx}x{ny} rows={rows}", flush=True)
    f1, d1, r1 = _steps(nx, ny, rows, True, 10)
    f0, d0, r0 = _steps(nx, ny, rows, False, 10)
    print(f"[local] {nx}x{ny}: last-step rounds {r1} in-round CG, {r0} one per phase; "
          f"iterations {[d[1] for d in d1]}", flush=True)
    for a, b in zip(f1, f0):
        assert np.array_equal(a, b)
    assert d1 == d0
    assert r1 < r0


def test_local_cg_iteration_cap_bitwise():
    """A cap the solves hit (lin_maxiter 7): same bits, same info and
    iteration counts."""
    print("[local] cap", flush=True)
    f1, d1, _ = _steps(48, 48, 4, True, 4, maxiter=7)
    f0, d0, _ = _steps(48, 48, 4, False, 4, maxiter=7)
    for a, b in zip(f1, f0):
        assert np.array_equal(a, b)
    assert d1 == d0
    # step 1 from rest has a zero right-hand side (info 0); the others stop at the cap (info = maxiter)
    assert d1[0][2] == 0 and all(d[2] == 7 and d[1] == 7 for d in d1[1:])


@pytest.mark.parametrize("op", ["pinned", "dirichlet"])
def test_local_cg_solve_bitwise(op):
    """Poisson solves of both operators on a 64x56 grid to 1e-10 within 2000
    iterations (the pinned one stops at the cap): same solution bits, info
    and iteration counts."""
    from pyflow import _hip as H
    from pyflow._device import DeviceBatch, cavity_params
    nx, ny = 64, 56
    rng = np.random.default_rng(11)
    rhs = rng.standard_normal((ny, nx))
    opc = H.OP_NEG_LAP_PINNED if op == "pinned" else H.OP_POS_LAP_DIRICHLET
    if op == "pinned":
        rhs -= rhs.mean()
        rhs[0, 0] = 0.0
    else:
        rhs[0, :] = rhs[-1, :] = rhs[:, 0] = rhs[:, -1] = 0.0
    out = {}
    for local in (True, False):
        print(f"[local] solve {op} local={local}", flush=True)

        def make():
            p = cavity_params(nx, ny, dx=1.0 / (nx - 1), dy=1.0 / (ny - 1), lin_tol=1e-10, lin_maxiter=2000,
                              op=opc, mode=H.MODE_SOLVE)
            return DeviceBatch([p], "fp64")

        b = _with_local(local, make)
        try:
            b.set_rhs(0, rhs)
            d = b.run(1)[0]
            out[local] = (b.solution(0), d.rp_iterations, d.rp_info, d.rp_residual)
        finally:
            b.close()
    assert np.array_equal(out[True][0], out[False][0])
    assert out[True][1:] == out[False][1:]


def test_local_cg_nan_runs_round_per_phase():
    """A NaN in u spreads into the pressure solve: that solve leaves the
    in-round loop and runs one round per phase, with the same bits (NaN where
    NaN) and counts as the round-per-phase path, and lin_maxiter iterations."""
    nx, ny = 48, 40
    u0 = np.zeros((ny + 2, nx + 2))
    u0[ny // 2, nx // 3] = np.nan
    print("[local] nan", flush=True)
    f1, d1, r1 = _steps(nx, ny, 4, True, 2, maxiter=50, u0=u0)
    f0, d0, r0 = _steps(nx, ny, 4, False, 2, maxiter=50, u0=u0)
    for a, b in zip(f1, f0):
        assert np.array_equal(a, b, equal_nan=True)
    assert np.array_equal(np.array(d1, dtype=float), np.array(d0, dtype=float), equal_nan=True)
    assert d1[-1][1] == 50 and r1 == r0
