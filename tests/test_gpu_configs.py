"""GPU checks of the batched and served configurations at their real sizes.

* Batched: the seeded 256-cavity batch (n_i = 32*U{1..16}, Re_i log-uniform
  [100, 5000], seed 0), with the tiles the autotuner picks for it, advances a
  few steps; sampled cavities run alone give the same bits as in the mix, and
  the 32^2 and 64^2 cavities follow the FP64 NumPy oracle.
* Served: the seeded request stream (n ~ round(exp(U(ln 64, ln 2048))), Re
  log-uniform [100, 1e4], seed 0; steps U{100..2000} scaled down so the test
  stays short) through the continuous-batching server under an HBM budget
  that makes requests wait; every result equals the request run alone.
"""
import numpy as np
import pytest

from oracle import pyflow_oracle as O

pytestmark = pytest.mark.gpu


def _fields(batch, slot, n):
    u = np.zeros((n + 2, n + 2)); v = np.zeros_like(u); p = np.zeros_like(u)
    batch.download(slot, u, v, p)
    return u, v, p


def test_batch_of_256_seeded_cavities_alone_equals_mix_and_oracle():
    from types import SimpleNamespace
    from pyflow import autotune
    from pyflow._device import DeviceBatch
    from pyflow.workloads import batch_params
    steps = 3
    params, ns = batch_params(256, 0)
    assert len(params) == 256 and ns.min() == 32 and ns.max() == 512
    (rows, cap), _ = autotune.tune_batch(params, "mixed")
    print(f"[c4] tuned tile_rows={rows} max_workgroups={cap}", flush=True)
    for p in params:
        p.tile_rows, p.max_workgroups = rows, cap
    batch = DeviceBatch(params, "mixed")
    diags = batch.run(steps)
    assert all(diags[k].steps_done == steps and not diags[k].nan_detected for k in range(256))
    assert all(diags[k].rp_converged for k in range(256))
    order = np.argsort(ns, kind="stable")
    sample = sorted({int(order[0]), int(order[-1]), int(order[128]), int(order[64]), int(order[200]), 7})
    mix = {k: _fields(batch, k, int(ns[k])) for k in sample}
    mix_diag = {k: (diags[k].rp_iterations_run, diags[k].continuity, diags[k].dt) for k in sample}
    batch.close()
    for k in sample:
        alone = DeviceBatch([params[k]], "mixed")
        d = alone.run(steps)[0]
        got = _fields(alone, 0, int(ns[k]))
        alone.close()
        for f, a, b in zip("uvp", got, mix[k]):
            assert np.array_equal(a, b), (k, int(ns[k]), f)
        assert (d.rp_iterations_run, d.continuity, d.dt) == mix_diag[k]
    print(f"[c4] alone == in the mix, bitwise, for cavities {sample} (n = {[int(ns[k]) for k in sample]})",
          flush=True)
    small = [k for k in range(256) if ns[k] <= 64][:4]
    assert small
    for k in small:
        n = int(ns[k])
        cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05,
                              Re=1.0 / params[k].nu, lid_velocity=1.0, lin_tol=1e-6, lin_maxiter=10000)
        ref = O.allocate_state(n, n)
        for it in range(steps):
            res, _ = O.step(cfg, ref, it)
        u, v, p = mix.get(k) or (None, None, None)
        if u is None:
            b1 = DeviceBatch([params[k]], "mixed")
            b1.run(steps)
            u, v, p = _fields(b1, 0, n)
            b1.close()
        du = float(np.max(np.abs(u - ref.fields["u"])))
        dv = float(np.max(np.abs(v - ref.fields["v"])))
        print(f"[c4] cavity {k} ({n}^2) vs FP64 oracle: max|du|={du:.2e} max|dv|={dv:.2e}", flush=True)
        assert du <= 1e-5 and dv <= 1e-5


def test_served_stream_at_config_sizes_alone_equals_mix():
    from pyflow._device import DeviceBatch
    from pyflow.serving import CavityRequest, CavityServer, footprint
    from pyflow.workloads import serve_requests
    count = 12
    ns, res, steps = serve_requests(count, seed=0, step_scale=0.002)
    reqs = [CavityRequest(int(n), int(n), int(s), Re=float(re)) for n, re, s in zip(ns, res, steps)]
    print(f"[c5] sizes {sorted(int(n) for n in ns)}, steps {list(map(int, steps))}", flush=True)
    assert ns.min() >= 64 and ns.max() <= 2048
    fp = sorted(footprint(r, "mixed") for r in reqs)
    budget = fp[-1] + fp[-2] + fp[0]  # the two largest and the smallest fit together
    srv = CavityServer("mixed", hbm_budget=budget, max_cavities=16, chunk_rounds=64)
    results, tickets = {}, []
    waited = False
    for k, r in enumerate(reqs):  # arrivals: two requests per scheduling step
        tickets.append(srv.submit(r))
        if k % 2 == 1:
            waited = waited or srv.waiting > 0
            for out in srv.pump():
                results[out.ticket] = out
    while not srv.idle():
        waited = waited or srv.waiting > 0
        for out in srv.pump():
            results[out.ticket] = out
    assert waited and srv.peak_bytes <= budget and srv.peak_running >= 2
    srv.close()
    assert sorted(results) == tickets
    for t, r in zip(tickets, reqs):
        got = results[t]
        assert got.status == "done" and got.diag.steps_done == r.steps
        b = DeviceBatch([r.params()], "mixed")
        d = b.run(r.steps)[0]
        u, v, p = _fields(b, 0, r.nx)
        b.close()
        assert np.array_equal(got.u, u) and np.array_equal(got.v, v) and np.array_equal(got.p, p), (t, r.nx)
        assert got.diag.rp_iterations_run == d.rp_iterations_run
    print(f"[c5] {count} requests: every result equals the request run alone (peak {srv.peak_running} running, "
          f"{srv.peak_bytes} of {budget} bytes)", flush=True)
