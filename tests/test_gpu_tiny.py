"""Lid-driven cavities at the smallest sizes the reference accepts (2x2 up to
7x4), and ragged sizes one cell either side of the 64-cell tile width and the
16-row tile height: three QUICK steps in FP64 and in mixed precision against
the FP64 oracle. In the small ones every cell is on or next to the boundary,
so the operators' edge cases carry the whole step; the ragged ones have
partial tiles and, in FP64, run one round per CG phase (several regions) or
the CG inside one round (one region)."""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import pyflow_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("nx,ny", [(2, 2), (3, 2), (2, 5), (5, 3), (7, 4), (63, 65), (65, 17), (129, 127)])
def test_tiny_cavity_steps_follow_oracle(nx, ny, prec):
    from pyflow._device import DeviceBatch, cavity_params
    re, tol, maxiter, nsteps = 100.0, 1e-10, 20000, 3  # every solve converges (unconverged ones differ by path)
    print(f"[tiny] {prec} {nx}x{ny}", flush=True)
    p = cavity_params(nx, ny, dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0, Re=re, lin_tol=tol,
                      lin_maxiter=maxiter)
    b = DeviceBatch([p], prec)
    try:
        d = b.run(nsteps)[0]
        u = np.zeros((ny + 2, nx + 2)); v = np.zeros_like(u); pp = np.zeros_like(u)
        b.download(0, u, v, pp)
    finally:
        b.close()
    cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05,
                          Re=re, lid_velocity=1.0, lin_tol=tol, lin_maxiter=maxiter)
    ref = O.allocate_state(nx, ny)
    for it in range(nsteps):
        O.step(cfg, ref, it)
    assert d.steps_done == nsteps and not d.nan_detected
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(v)) and np.all(np.isfinite(pp))
    bound = 1e-9 if prec == "fp64" else 1e-5
    du = float(np.max(np.abs(u - ref.fields["u"])))
    dv = float(np.max(np.abs(v - ref.fields["v"])))
    print(f"[tiny] {prec} {nx}x{ny}: max|du| {du:.2e} max|dv| {dv:.2e}", flush=True)
    assert du <= bound and dv <= bound
