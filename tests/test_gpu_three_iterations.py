"""The three-iteration mixed PCG round (s-step PCG with s = 3, pcg_per_round
3, kernel k_round_ca3): true-residual convergence on both operators against
the two-iteration round, exact iteration caps, cavity steps against the FP64
oracle for every region height and grouping it runs with, bitwise equality
alone and in a batch, and the rule that such cavities do not share a batch
with others."""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import pyflow_oracle as O

pytestmark = pytest.mark.gpu


def _rhs(n, seed):
    b = np.random.default_rng(seed).standard_normal(n * n)
    b -= b.mean()
    b[0] = 0.0
    return b


def _solve(n, op, b, tol, maxiter, jacobi, s):
    from pyflow import _hip as H
    from pyflow._device import DeviceBatch, cavity_params
    dx = 1.0 if op == H.OP_NEG_LAP_PINNED else 1.0 / (n - 1)
    p = cavity_params(n, n, dx=dx, dy=dx, lin_tol=tol, lin_maxiter=maxiter, jacobi=jacobi, op=op,
                      mode=H.MODE_SOLVE)
    p.pcg_per_round = s
    bt = DeviceBatch([p], "mixed")
    bt.set_rhs(0, b.reshape(n, n))
    d = bt.run(1)[0]
    x = bt.solution(0).reshape(-1).copy()
    got = bt.resolved_params(0).pcg_per_round
    bt.close()
    return x, d, got


@pytest.mark.parametrize("n", [64, 200])
def test_three_iterations_reach_fp64_tolerance(n):
    """Pinned operator with and without Jacobi and the Dirichlet operator, to
    1e-6 and 1e-10: the FP64 true residual (computed on the host with the
    oracle's operators) meets the tolerance, in at most 20% (+6) more
    iterations than the two-iteration round."""
    from pyflow import _hip as H
    b = _rhs(n, 5)
    bd = b.reshape(n, n).copy()
    bd[0, :] = bd[-1, :] = bd[:, 0] = bd[:, -1] = 0.0
    h = 1.0 / (n - 1)
    cases = [("pinned", H.OP_NEG_LAP_PINNED, lambda z: O.neg_laplacian_pinned(z, n, n, 1.0, 1.0), True, b),
             ("pinned-noM", H.OP_NEG_LAP_PINNED, lambda z: O.neg_laplacian_pinned(z, n, n, 1.0, 1.0), False, b),
             ("dirichlet", H.OP_POS_LAP_DIRICHLET, lambda z: O.pos_laplacian_dirichlet(z, n, n, h, h), True,
              bd.reshape(-1))]
    bad = []
    for name, op, mv, jac, rhs in cases:
        for tol in (1e-6, 1e-10):
            its = {}
            for s in (3, 2):
                print(f"[s3] n={n} {name} tol={tol:g} s={s}", flush=True)
                x, d, got = _solve(n, op, rhs, tol, 40000, jac, s)
                assert got == s
                rel = float(np.linalg.norm(rhs - mv(x)) / np.linalg.norm(rhs))
                print(f"[s3] n={n} {name} tol={tol:g} s={s}: info={d.rp_info} its={d.rp_iterations} "
                      f"passes={d.outer_passes} rel={rel:.3e}", flush=True)
                its[s] = d.rp_iterations
                if not (d.rp_info == 0 and d.rp_converged and rel <= tol * 1.0001):
                    bad.append((name, tol, s, d.rp_info, d.rp_iterations, rel))
            assert its[3] <= 1.2 * its[2] + 6, (name, tol, its)
    assert not bad, bad


def test_three_iterations_honour_iteration_caps():
    """Caps that are not multiples of three end the solve after exactly that
    many iterations (the last round runs one or two)."""
    from pyflow import _hip as H
    n = 96
    b = _rhs(n, 7)
    for cap in (1, 2, 7, 8, 13):
        print(f"[s3] cap {cap}", flush=True)
        x, d, got = _solve(n, H.OP_NEG_LAP_PINNED, b, 1e-14, cap, True, 3)
        assert got == 3
        assert d.rp_info == cap and d.rp_iterations == cap and not d.rp_converged, (cap, d.rp_iterations)
        assert np.all(np.isfinite(x))


def _cavity(nx, ny, re, tol, maxiter, rows, cap, s=3):
    from pyflow._device import cavity_params
    p = cavity_params(nx, ny, dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0, Re=re, lin_tol=tol,
                      lin_maxiter=maxiter)
    p.tile_rows, p.max_workgroups, p.pcg_per_round = rows, cap, s
    return p


@pytest.mark.parametrize("rows,cap", [(4, 1024), (2, 1024), (4, 3), (2, 2)])
def test_three_iteration_steps_follow_oracle(rows, cap):
    """Six steps of a 130x97 cavity (Re 400, tol 1e-10) with 4- or 2-tile
    regions, one or several regions per workgroup, stay within 1e-5 of the
    FP64 oracle in u and v."""
    from pyflow._device import DeviceBatch
    nx, ny, re, tol, maxiter = 130, 97, 400.0, 1e-10, 5000
    print(f"[s3] steps rows={rows} cap={cap}", flush=True)
    b = DeviceBatch([_cavity(nx, ny, re, tol, maxiter, rows, cap)], "mixed")
    assert b.resolved_params(0).pcg_per_round == 3
    d = b.run(6)[0]
    u = np.zeros((ny + 2, nx + 2)); v = np.zeros_like(u)
    b.download(0, u, v, None)
    b.close()
    cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05,
                          Re=re, lid_velocity=1.0, lin_tol=tol, lin_maxiter=maxiter)
    ref = O.allocate_state(nx, ny)
    for it in range(6):
        O.step(cfg, ref, it)
    du = float(np.max(np.abs(u - ref.fields["u"])))
    dv = float(np.max(np.abs(v - ref.fields["v"])))
    print(f"[s3] rows={rows} cap={cap}: max|du|={du:.3e} max|dv|={dv:.3e}", flush=True)
    assert d.steps_done == 6 and not d.nan_detected and d.rp_converged
    assert du <= 1e-5 and dv <= 1e-5


def test_three_iteration_batch_equals_alone():
    """Uneven cavities with three iterations per round give the same bits
    alone and together, and a batch that would mix them with cavities of
    another round kind is refused."""
    from pyflow import _hip as H
    from pyflow._device import DeviceBatch
    shapes = [(150, 120, 4, 1024), (64, 64, 2, 1024), (200, 90, 4, 2), (96, 40, 2, 1024)]
    params = [_cavity(nx, ny, 300.0 + 100 * k, 1e-8, 3000, rows, cap) for k, (nx, ny, rows, cap) in
              enumerate(shapes)]

    def fields(batch, slot, nx, ny):
        u = np.zeros((ny + 2, nx + 2)); v = np.zeros_like(u); p = np.zeros_like(u)
        batch.download(slot, u, v, p)
        return u, v, p

    print("[s3] batch of 4", flush=True)
    b = DeviceBatch(params, "mixed")
    dall = b.run(4)
    together = [fields(b, k, nx, ny) for k, (nx, ny, _, _) in enumerate(shapes)]
    b.close()
    for k, (nx, ny, _, _) in enumerate(shapes):
        print(f"[s3] alone {k}", flush=True)
        a = DeviceBatch([params[k]], "mixed")
        d = a.run(4)[0]
        alone = fields(a, 0, nx, ny)
        a.close()
        assert d.rp_iterations_run == dall[k].rp_iterations_run
        for f0, f1 in zip(alone, together[k]):
            assert np.array_equal(f0, f1), k
    mixed = [params[0], _cavity(64, 64, 400.0, 1e-8, 3000, 4, 1024, s=2)]
    m = DeviceBatch(mixed, "mixed")
    try:
        with pytest.raises(RuntimeError, match="three PCG iterations"):
            m.run(1)
    finally:
        m.close()
