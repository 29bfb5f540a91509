"""GPU checks of the tile/schedule fields and the autotuner: every tiling the
tuner may pick follows the FP64 oracle, a tuned cavity gives the same bits
alone and in a batch with the same tiles, and the tuner returns one of its
candidates with a measured cost."""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import pyflow_oracle as O

pytestmark = pytest.mark.gpu


def _cavity(nx, ny, re, tol, maxiter):
    from pyflow._device import cavity_params
    return cavity_params(nx, ny, dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0, Re=re, lin_tol=tol,
                         lin_maxiter=maxiter)


@pytest.mark.parametrize("precision,rows,cap,s,bound", [
    ("mixed", 1, 3, 2, 1e-5), ("mixed", 2, 1024, 1, 1e-5), ("mixed", 8, 2, 1, 1e-5), ("mixed", 4, 1, 2, 1e-5),
    ("fp64", 1, 5, 1, 1e-8), ("fp64", 8, 1024, 1, 1e-8),
])
def test_tilings_follow_oracle(precision, rows, cap, s, bound):
    """Six steps of a 130x97 cavity with a given region height, workgroup cap
    and iterations per round stay within the precision's bound of the FP64
    oracle (u, v max-abs, lid U = 1)."""
    from pyflow import autotune as A
    from pyflow._device import DeviceBatch
    nx, ny, re, tol, maxiter = 130, 97, 400.0, 1e-10, 5000
    p = A.apply(_cavity(nx, ny, re, tol, maxiter), A.Tile(rows, cap, s))
    b = DeviceBatch([p], precision)
    d = b.run(6)[0]
    u = np.zeros((ny + 2, nx + 2)); v = np.zeros_like(u)
    b.download(0, u, v, None)
    b.close()
    cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05,
                          Re=re, lid_velocity=1.0, lin_tol=tol, lin_maxiter=maxiter)
    ref = O.allocate_state(nx, ny)
    for it in range(6):
        O.step(cfg, ref, it)
    assert d.steps_done == 6 and not d.nan_detected and d.rp_converged
    assert float(np.max(np.abs(u - ref.fields["u"]))) <= bound
    assert float(np.max(np.abs(v - ref.fields["v"]))) <= bound


def test_tuned_cavity_alone_equals_in_batch():
    from pyflow import autotune as A
    from pyflow._device import DeviceBatch
    sizes = [(200, 136, 700.0), (96, 96, 150.0), (64, 40, 3000.0)]
    tiles = [A.Tile(2, 8, 2), A.Tile(8, 3, 1), A.Tile(1, 1024, 2)]
    ps = [A.apply(_cavity(nx, ny, re, 1e-6, 5000), t) for (nx, ny, re), t in zip(sizes, tiles)]
    mix = DeviceBatch(ps, "mixed")
    dm = mix.run(4)
    for k, ((nx, ny, _), p) in enumerate(zip(sizes, ps)):
        alone = DeviceBatch([p], "mixed")
        da = alone.run(4)[0]
        ua = np.zeros((ny + 2, nx + 2)); alone.download(0, ua, None, None)
        um = np.zeros((ny + 2, nx + 2)); mix.download(k, um, None, None)
        alone.close()
        assert np.array_equal(ua, um), k
        assert da.rp_iterations_run == dm[k].rp_iterations_run, k
    mix.close()


def test_autotune_returns_a_measured_candidate(tmp_path, monkeypatch):
    from pyflow import autotune as A
    monkeypatch.setenv("PYFLOW_TUNE_CACHE", str(tmp_path / "tiles.json"))
    nx, ny = 200, 136
    t = A.tune(nx, ny, "mixed", warmup=8, rounds=16, force=True)
    assert np.isfinite(t.us_per_iteration) and t.us_per_iteration > 0
    keys = {(c.tile_rows, c.max_workgroups, c.pcg_per_round) for c in A.candidates(nx, ny, "mixed")}
    assert (t.tile_rows, t.max_workgroups, t.pcg_per_round) in keys
    assert A.cached(nx, ny, "mixed") == t
    A._cache.clear()
    assert A.cached(nx, ny, "mixed") == t  # from the JSON file
    best, res = A.tune_batch([_cavity(96, 96, 100.0, 1e-6, 100), _cavity(40, 72, 100.0, 1e-6, 100)], "mixed",
                             tilings=((4, 2), (4, 16), (2, 1024)), warmup=4, rounds=8)
    assert best in ((4, 2), (4, 16), (2, 1024)) and all(np.isfinite(v) and v > 0 for v in res.values())


def test_fp64_iterate_kinds_mixed_in_one_batch(monkeypatch):
    """FP64-iterate cavities with one and with two PCG iterations per round in
    one batch (the kernel instance that compiles both bodies) give the same
    bits as each alone (instances with one body), and follow the FP64 oracle."""
    from pyflow import autotune as A
    from pyflow._device import DeviceBatch
    monkeypatch.setenv("PYFLOW_XD", "1")
    nx, ny, re, tol, maxiter = 130, 97, 400.0, 1e-10, 5000
    specs = [(A.Tile(4, 1024, 2), 130, 97), (A.Tile(2, 3, 1), 96, 72), (A.Tile(1, 4, 2), 70, 40)]
    ps = [A.apply(_cavity(x, y, re, tol, maxiter), t) for t, x, y in specs]
    mix = DeviceBatch(ps, "mixed")
    dm = mix.run(3)
    for k, ((t, x, y), p) in enumerate(zip(specs, ps)):
        alone = DeviceBatch([p], "mixed")
        da = alone.run(3)[0]
        ua = np.zeros((y + 2, x + 2)); alone.download(0, ua, None, None)
        um = np.zeros((y + 2, x + 2)); mix.download(k, um, None, None)
        alone.close()
        assert np.array_equal(ua, um), k
        assert da.rp_iterations_run == dm[k].rp_iterations_run, k
        assert dm[k].rp_converged and not dm[k].nan_detected
        if k == 0:
            cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5,
                                  cfl_growth=1.05, Re=re, lid_velocity=1.0, lin_tol=tol, lin_maxiter=maxiter)
            ref = O.allocate_state(x, y)
            for it in range(3):
                O.step(cfg, ref, it)
            assert float(np.max(np.abs(um - ref.fields["u"]))) <= 1e-5
    mix.close()


def test_two_workgroups_per_cu_kernel_equals_default(monkeypatch):
    """A batch whose cavities all use regions of at most two tiles, one per
    workgroup, runs the kernel instance with two workgroups per CU; it gives
    the same bits as the one-workgroup-per-CU instance (PYFLOW_OCC2=0) and
    follows the FP64 oracle."""
    from pyflow import autotune as A
    from pyflow._device import DeviceBatch
    re, tol, maxiter = 400.0, 1e-10, 5000
    specs = [(130, 97, A.Tile(2, 1024, 2)), (70, 40, A.Tile(1, 1024, 2)), (32, 32, A.Tile(0, 0, 0))]
    ps = [A.apply(_cavity(x, y, re, tol, maxiter), t) for x, y, t in specs]
    out = {}
    for occ in ("1", "0"):
        monkeypatch.setenv("PYFLOW_OCC2", occ)
        b = DeviceBatch(ps, "mixed")
        d = b.run(4)
        fields = []
        for k, (x, y, _) in enumerate(specs):
            u = np.zeros((y + 2, x + 2)); v = np.zeros_like(u)
            b.download(k, u, v, None)
            fields.append((u, v, d[k].rp_iterations_run))
        b.close()
        out[occ] = fields
    for (u1, v1, i1), (u0, v0, i0) in zip(out["1"], out["0"]):
        assert np.array_equal(u1, u0) and np.array_equal(v1, v0) and i1 == i0
    x, y, _ = specs[0]
    cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05,
                          Re=re, lid_velocity=1.0, lin_tol=tol, lin_maxiter=maxiter)
    ref = O.allocate_state(x, y)
    for it in range(4):
        O.step(cfg, ref, it)
    assert float(np.max(np.abs(out["1"][0][0] - ref.fields["u"]))) <= 1e-5
