"""Closed-form, invariant and brute-force pins of the oracle, chosen so a
plausible mistake (dropped term, wrong sign or index, transposed operand)
fails at least one of them independently of the golden vectors."""
from types import SimpleNamespace

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import pyflow_oracle as O


@pytest.mark.parametrize("nx,ny", [(8, 8), (16, 12), (5, 9)])
def test_lap_equals_div_grad(nx, ny):
    rng = np.random.default_rng(42)
    p = rng.standard_normal((ny, nx))
    dx, dy = 1.0 / (nx - 1), 1.0 / (ny - 1)
    gx, gy = O.gradient_coherent(p, dx, dy)
    assert np.array_equal(O.divergence_coherent(gx, gy, dx, dy), O.laplacian_coherent(p, dx, dy))


@pytest.mark.parametrize("nx,ny", [(8, 8), (11, 7)])
def test_constant_null_space_and_zero_mean(nx, ny):
    lap = O.laplacian_coherent(np.full((ny, nx), 3.14), 0.3, 0.7)
    assert np.allclose(lap, 0.0, atol=1e-14)
    p = np.random.default_rng(7).standard_normal((ny, nx))
    assert abs(np.mean(O.laplacian_coherent(p, 0.3, 0.7))) < 5e-13


def _dense_neumann_5pt(nx, ny, dx, dy):
    """Brute-force reflective-Neumann 5-point matrix: -Lap with missing neighbours dropped."""
    N = nx * ny
    A = np.zeros((N, N))
    for j in range(ny):
        for i in range(nx):
            k = j * nx + i
            for (jj, ii, c) in ((j, i - 1, dx), (j, i + 1, dx), (j - 1, i, dy), (j + 1, i, dy)):
                if 0 <= jj < ny and 0 <= ii < nx:
                    A[k, k] += 1.0 / c ** 2
                    A[k, jj * nx + ii] -= 1.0 / c ** 2
    return A


@pytest.mark.parametrize("nx,ny,dx,dy", [(6, 5, 1.0, 1.0), (7, 4, 0.3, 0.9), (2, 3, 1.0, 2.0)])
def test_pinned_operator_matches_dense_assembly(nx, ny, dx, dy):
    A = _dense_neumann_5pt(nx, ny, dx, dy)
    A[0, :] = 0.0
    A[0, 0] = 1.0
    rng = np.random.default_rng(3)
    for _ in range(3):
        x = rng.standard_normal(nx * ny)
        np.testing.assert_allclose(O.neg_laplacian_pinned(x, nx, ny, dx, dy), A @ x, rtol=1e-12, atol=1e-12)


def test_dirichlet_operator_matches_dense_assembly():
    nx, ny, dx, dy = 6, 5, 0.5, 0.25
    A = -_dense_neumann_5pt(nx, ny, dx, dy)
    for j in range(ny):
        for i in range(nx):
            if i in (0, nx - 1) or j in (0, ny - 1):
                k = j * nx + i
                A[k, :] = 0.0
                A[k, k] = 1.0
    x = np.random.default_rng(5).standard_normal(nx * ny)
    np.testing.assert_allclose(O.pos_laplacian_dirichlet(x, nx, ny, dx, dy), A @ x, rtol=1e-12, atol=1e-12)


def test_interior_stencil_known_answer():
    """A (x^2+y^2)/4 = -1 inside for dx=dy=1 (tests/test_projection_method_contract.py:1-22)."""
    nx = ny = 8
    X, Y = np.meshgrid(np.arange(1, nx + 1.0), np.arange(1, ny + 1.0))
    p = (X ** 2 + Y ** 2) / 4.0
    out = O.neg_laplacian_pinned(p.reshape(-1), nx, ny, 1.0, 1.0).reshape(ny, nx)
    assert np.allclose(out[1:-1, 1:-1], -1.0, atol=1e-12)


def test_jacobi_diagonal_closed_form():
    inv = O.jacobi_inv_diag(5, 4, 0.5, 0.25).reshape(4, 5)
    assert np.all(inv[0, :] == 1.0) and np.all(inv[-1, :] == 1.0)
    assert np.all(inv[:, 0] == 1.0) and np.all(inv[:, -1] == 1.0)
    assert np.all(inv[1:-1, 1:-1] == 1.0 / (2.0 / 0.25 + 2.0 / 0.0625))


def test_central_divergence_linear_fields_and_right_column():
    nx, ny = 9, 7
    X, Y = np.meshgrid(np.arange(nx, dtype=float), np.arange(ny, dtype=float))
    d = O.divergence_central(X, Y, 1.0, 1.0)
    # linear fields are exact in the interior, on the left column, rows 0 / ny-1 and corners
    assert np.allclose(d[:, :-1], 2.0)
    assert np.allclose(d[[0, -1], -1], 2.0)
    # the reference never assigns the right column between the corners
    assert np.all(d[1:-1, -1] == 0.0)
    d2 = O.divergence_central(3.0 * X, -Y, 1.0, 1.0)
    assert np.allclose(d2[:, :-1], 2.0)


def test_advection_orders_on_quadratic():
    """x^2+y^2 with u=v=1: QUICK beats upwind, both small (contract test :270-296)."""
    nx = ny = 16
    X, Y = np.meshgrid(np.arange(nx, dtype=float), np.arange(ny, dtype=float))
    phi = X ** 2 + Y ** 2
    one = np.ones_like(phi)
    exp = 2 * X + 2 * Y
    m = (slice(2, -2), slice(2, -2))
    eu = np.mean(np.abs(O.advect_upwind(one, one, phi, 1.0, 1.0)[m] - exp[m]))
    eq = np.mean(np.abs(O.advect_quick(one, one, phi, 1.0, 1.0)[m] - exp[m]))
    assert eq < eu < 5.0
    # interior QUICK with u>0 is exact on a quadratic: e - w = 2x per unit spacing
    q = O.advect_quick(one, one, phi, 1.0, 1.0)
    assert np.allclose(q[m], exp[m], rtol=0, atol=1e-12)


def test_quick_negative_flow_face_is_reference_formula():
    """For u<0 the east face is 3/8 W + 6/8 P - 1/8 E (advection.py:65), not the textbook form."""
    f = np.array([[0.0, 1.0, 4.0, 9.0, 16.0]])
    u = -np.ones_like(f)
    v = np.zeros_like(f)
    c = O.advect_quick(u, v, f, 1.0, 1.0)
    i = 2
    W, P, E, WW = f[0, i - 1], f[0, i], f[0, i + 1], f[0, i - 2]
    pe = (3 / 8 * W + 6 / 8 * P) - 1 / 8 * E
    pw = (3 / 8 * WW + 6 / 8 * W) - 1 / 8 * P
    assert c[0, i] == (-1.0 * pe - -1.0 * pw)


@pytest.mark.parametrize("pc", [True, False])
def test_cg_bitwise_equals_scipy(pc):
    nx, ny, dx, dy = 20, 17, 1.0, 1.0
    rng = np.random.default_rng(11)
    b = rng.standard_normal(nx * ny)
    b -= b.mean()
    b[0] = 0.0
    mv = lambda x: O.neg_laplacian_pinned(x, nx, ny, dx, dy)  # noqa: E731
    invD = O.jacobi_inv_diag(nx, ny, dx, dy)
    ps = (lambda r: invD * r) if pc else None
    A = spla.LinearOperator((nx * ny, nx * ny), matvec=mv, dtype=float)
    M = spla.LinearOperator((nx * ny, nx * ny), matvec=lambda r: invD * r, dtype=float) if pc else None
    for maxiter in (5, 37, 400):
        x_ref, info_ref = spla.cg(A, b, x0=np.zeros_like(b), rtol=1e-9, maxiter=maxiter, M=M)
        x, info, _ = O.cg(mv, b, 1e-9, maxiter, ps)
        assert info == info_ref
        assert np.array_equal(x, x_ref)


def test_cg_solves_dense_system():
    nx, ny = 9, 8
    mv = lambda x: O.neg_laplacian_pinned(x, nx, ny, 1.0, 1.0)  # noqa: E731
    A = np.stack([mv(e) for e in np.eye(nx * ny)], axis=1)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(nx * ny)
    b -= b.mean()
    b[0] = 0.0
    x, info, its = O.cg(mv, b, 1e-13, 1000, None)
    assert info == 0 and its > 0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-10)


def test_cg_edge_cases():
    mv = lambda x: 2.0 * x  # noqa: E731
    x, info, its = O.cg(mv, np.zeros(4), 1e-8, 10, None)
    assert info == 0 and its == 0 and not x.any()
    x, info, its = O.cg(mv, np.array([np.nan, 1.0]), 1e-8, 7, None)
    assert info == 7 and its == 7 and np.isnan(x).all()


def test_safeguard_halves_dt():
    """Rising central divergence halves dt (contract test :482-516)."""
    cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.9, cfl_growth=1.2,
                          Re=100.0, lid_velocity=1.0, lin_tol=1e-10, lin_maxiter=50)
    st = O.allocate_state(8, 8)
    u = st.fields["u"][1:-1, 1:-1]
    for j in range(8):
        for i in range(8):
            u[j, i] = ((-1) ** (i + j)) * (i + j) * 0.05
    dts = []
    for k in range(6):
        O.step(cfg, st, k)
        dts.append(st.meta["dt_prev"])
    assert any(dts[i + 1] < dts[i] for i in range(5))


def test_planted_mistakes_fail_the_pins(golden_ops, monkeypatch):
    """A wrong QUICK weight or a wrong Laplacian sign in the oracle is caught."""
    g = golden_ops
    u, v, f = g["ops2_u"], g["ops2_v"], g["ops2_f"]
    dx, dy = g["ops2_dxdy"]
    monkeypatch.setattr(O, "_C6", 5 / 8)
    assert not np.array_equal(O.advect_quick(u, v, f, dx, dy), g["ops2_quick"])
    monkeypatch.undo()
    orig = O.laplacian_coherent
    monkeypatch.setattr(O, "laplacian_coherent", lambda p, a, b: -orig(p, a, b))
    X, Y = np.meshgrid(np.arange(1, 9.0), np.arange(1, 9.0))
    out = O.neg_laplacian_pinned(((X ** 2 + Y ** 2) / 4.0).reshape(-1), 8, 8, 1.0, 1.0).reshape(8, 8)
    assert not np.allclose(out[1:-1, 1:-1], -1.0)


def test_planted_structural_mistakes_fail_the_pins(golden_ops, monkeypatch):
    """A dropped term, a transposed operand and an off-by-one pin, each planted
    in the oracle, are caught by the reference-written golden vectors."""
    g = golden_ops
    u, v, f = g["ops2_u"], g["ops2_v"], g["ops2_f"]
    dx, dy = g["ops2_dxdy"]
    # dropped term: the coherent divergence without its y part
    orig_div = O.divergence_coherent

    def div_no_y(uu, vv, a, b):
        return orig_div(uu, np.zeros_like(vv), a, b)

    monkeypatch.setattr(O, "divergence_coherent", div_no_y)
    assert not np.array_equal(O.divergence_coherent(u, v, dx, dy), g["ops2_hdiv"])
    monkeypatch.undo()
    # transposed operand: the x derivative differenced along rows (axis 0)
    # and the y derivative along columns
    orig_grad = O.gradient_coherent

    def grad_t(p, a, b):
        gx_true, gy_true = orig_grad(p, a, b)
        return gy_true, gx_true

    monkeypatch.setattr(O, "gradient_coherent", grad_t)
    gx, gy = O.gradient_coherent(f, dx, dy)
    assert not np.array_equal(gx, g["ops2_gx"]) and not np.array_equal(gy, g["ops2_gy"])
    monkeypatch.undo()
    # off-by-one pin: the identity row at cell 1 instead of cell 0
    orig_lap = O.laplacian_coherent

    def pinned_at_1(x, nx, ny, a, b):
        out = (-orig_lap(x.reshape(ny, nx), a, b)).reshape(-1)
        out[1] = x[1]
        return out

    monkeypatch.setattr(O, "neg_laplacian_pinned", pinned_at_1)
    nx, ny = 24, 20
    uu = np.zeros((ny, nx)); vv = np.zeros((ny, nx)); pp = np.zeros((ny, nx))
    uu[ny // 3, nx // 3] = 1.0
    cfg = SimpleNamespace(lin_tol=1e-10, lin_maxiter=300, enable_jacobi_pc=True)
    psol, _ = O.solve_pressure_poisson(uu, vv, pp, 0.01, 1.0, 1.0, cfg)
    assert not np.array_equal(psol, g["ps_bump_pc_p"])


def test_loop_faithful_forms_equal_vectorised():
    """The per-cell loop forms (the reference's cost shape, used for the
    loop-faithful CPU baseline) give the same bits as the vectorised oracle:
    advection on random fields of both signs, the per-step LIL Jacobi
    assembly, and a whole step."""
    from types import SimpleNamespace
    rng = np.random.default_rng(11)
    for ny, nx, dx, dy in ((7, 9, 1.0, 1.0), (12, 5, 0.1, 0.2), (2, 2, 1.0, 1.0)):
        u, v, f = (rng.standard_normal((ny, nx)) for _ in range(3))
        u[0, 0] = 0.0  # the >= 0 / > 0 branch boundary
        assert np.array_equal(O.advect_quick_loops(u, v, f, dx, dy), O.advect_quick(u, v, f, dx, dy))
        assert np.array_equal(O.advect_upwind_loops(u, v, f, dx, dy), O.advect_upwind(u, v, f, dx, dy))
        assert np.array_equal(O.jacobi_inv_diag_assembled(nx, ny, dx, dy), O.jacobi_inv_diag(nx, ny, dx, dy))
    cfg = SimpleNamespace(disable_advection=False, advection_scheme="quick", cfl_target=0.5, cfl_growth=1.05,
                          Re=100.0, lid_velocity=1.0, lin_tol=1e-8, lin_maxiter=300)
    a, b = O.allocate_state(16, 12), O.allocate_state(16, 12)
    for it in range(3):
        ra, _ = O.step(cfg, a, it)
        rb, _ = O.step(cfg, b, it, loops=True)
        assert ra == rb
    for k in ("u", "v", "p"):
        assert np.array_equal(a.fields[k], b.fields[k])
