"""GPU checks of the continuous-batching server: requests admitted by HBM
footprint into a running batch, overlapping on the device, give the same bits
as each request run alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _alone(req, precision):
    from pyflow._device import DeviceBatch
    b = DeviceBatch([req.params()], precision)
    d = b.run(req.steps)[0]
    u = np.zeros((req.ny + 2, req.nx + 2)); v = np.zeros_like(u); p = np.zeros_like(u)
    b.download(0, u, v, p)
    b.close()
    return u, v, p, d


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_server_admits_by_footprint_and_matches_alone(precision):
    from pyflow.serving import CavityRequest, CavityServer, footprint
    reqs = [CavityRequest(130, 97, 4, Re=400.0), CavityRequest(64, 64, 6, Re=2500.0),
            CavityRequest(96, 40, 3, Re=150.0), CavityRequest(150, 150, 2, Re=1000.0),
            CavityRequest(33, 70, 5, Re=700.0)]
    fp = [footprint(r, precision) for r in reqs]
    # Room for the two largest at once: the rest must wait for memory.
    budget = sorted(fp)[-1] + sorted(fp)[-2]
    srv = CavityServer(precision, hbm_budget=budget, max_cavities=8, chunk_rounds=24)
    results = {}
    tickets = [srv.submit(r) for r in reqs[:3]]
    saw_waiting = srv.waiting > 0
    for r in srv.pump():
        results[r.ticket] = r
    tickets += [srv.submit(r) for r in reqs[3:]]  # arrive while others run
    while not srv.idle():
        saw_waiting = saw_waiting or srv.waiting > 0
        assert srv.in_use_bytes <= budget
        for r in srv.pump():
            results[r.ticket] = r
    assert saw_waiting and srv.peak_bytes <= budget and srv.peak_running >= 2
    assert sorted(results) == tickets
    for t, req in zip(tickets, reqs):
        r = results[t]
        u, v, p, d = _alone(req, precision)
        assert np.array_equal(r.u, u) and np.array_equal(r.v, v) and np.array_equal(r.p, p), t
        assert r.diag.steps_done == req.steps and r.diag.rp_iterations_run == d.rp_iterations_run
        assert r.latency >= 0 and r.t_admit >= r.t_submit
    srv.close()


def test_server_refuses_request_larger_than_budget():
    from pyflow.serving import CavityRequest, CavityServer, footprint
    small = CavityRequest(32, 32, 1)
    srv = CavityServer("mixed", hbm_budget=footprint(small, "mixed"))
    with pytest.raises(ValueError):
        srv.submit(CavityRequest(256, 256, 1))
    srv.submit(small)
    assert len(srv.run_until_idle()) == 1
    srv.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_nan_request_is_evicted_without_disturbing_neighbours(precision):
    """A request whose fields hold a NaN joins a running batch: the server
    evicts it at the first chunk boundary after its first step started
    (status nan_detected, no step completed, its arena block released) -- in
    the reference's semantics its every PCG solve would run to lin_maxiter --
    and the co-resident requests still give the same bits as alone."""
    from pyflow.serving import CavityRequest, CavityServer, footprint
    nb = CavityRequest(64, 64, 40, Re=1000.0)
    u = np.zeros((nb.ny + 2, nb.nx + 2))
    u[nb.ny // 2, nb.nx // 3] = np.nan
    nb.fields = (u, np.zeros_like(u), np.zeros_like(u))
    reqs = [CavityRequest(96, 80, 6, Re=400.0), nb, CavityRequest(130, 70, 5, Re=150.0)]
    budget = sum(footprint(r, precision) for r in reqs)
    srv = CavityServer(precision, hbm_budget=budget, max_cavities=4, chunk_rounds=16)
    tickets = [srv.submit(reqs[0])]
    results = {r.ticket: r for r in srv.pump()}
    tickets += [srv.submit(r) for r in reqs[1:]]  # the NaN request arrives while another runs
    print(f"[serve nan] {precision}: request {tickets[1]} submitted with a NaN in u", flush=True)
    while not srv.idle():
        for r in srv.pump():
            results[r.ticket] = r
    assert srv.evicted == 1 and srv.in_use_bytes == 0 and srv.arena.free_bytes == srv.arena.size
    bad = results[tickets[1]]
    assert bad.status == "nan_detected" and bad.diag.steps_done == 0 and bad.diag.rp_iterations_run < 2 * 16
    assert np.isnan(bad.diag.umax) and np.isnan(bad.u).any()
    for t, req in ((tickets[0], reqs[0]), (tickets[2], reqs[2])):
        r = results[t]
        assert r.status == "done"
        u0, v0, p0, d = _alone(req, precision)
        assert np.array_equal(r.u, u0) and np.array_equal(r.v, v0) and np.array_equal(r.p, p0), t
    srv.close()
