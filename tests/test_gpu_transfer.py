"""Field transfers between host and device (pf_batch_upload / download,
pf_batch_set_rhs / get_solution). They run in chunks of rows through two
pinned halves; these cases span one chunk, several chunks with a partial
last one, and short wide cavities."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(64, 48), (2050, 1300), (3000, 40)]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("nx,ny", SHAPES)
def test_field_round_trip(prec, nx, ny):
    """Upload u, v, p with their ghost ring, download into arrays whose ring
    holds a marker: the interiors come back exactly (rounded to FP32 in
    mixed precision) and the rings keep the marker."""
    from pyflow._device import DeviceBatch, cavity_params
    print(f"[transfer] {prec} {nx}x{ny}", flush=True)
    b = DeviceBatch([cavity_params(nx, ny, dx=1.0, dy=1.0)], prec)
    try:
        rng = np.random.default_rng(nx + ny)
        fields = [rng.standard_normal((ny + 2, nx + 2)) for _ in range(3)]
        b.upload(0, *fields)
        out = [np.full((ny + 2, nx + 2), 7.0) for _ in range(3)]
        b.download(0, *out)
    finally:
        b.close()
    for a, o in zip(fields, out):
        inner = a[1:-1, 1:-1]
        want = inner if prec == "fp64" else inner.astype(np.float32).astype(np.float64)
        assert np.array_equal(o[1:-1, 1:-1], want)
        ring = np.ones_like(o, dtype=bool)
        ring[1:-1, 1:-1] = False
        assert np.all(o[ring] == 7.0)


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_rhs_and_solution_round_trip(nx, ny):
    """The right-hand side goes up with set_rhs and the solution comes back
    with get_solution: the solution meets the solve's tolerance on the
    host's FP64 operator, so both transfers carried every row."""
    from oracle import pyflow_oracle as O
    from pyflow import _hip as H
    from pyflow._device import DeviceBatch, cavity_params
    print(f"[transfer] rhs {nx}x{ny}", flush=True)
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal((ny, nx))
    rhs -= rhs.mean()
    rhs[0, 0] = 0.0
    p = cavity_params(nx, ny, dx=1.0, dy=1.0, lin_tol=1e-8, lin_maxiter=200000, op=H.OP_NEG_LAP_PINNED,
                      mode=H.MODE_SOLVE)
    b = DeviceBatch([p], "fp64" if nx * ny < 100000 else "mixed")
    try:
        b.set_rhs(0, rhs)
        d = b.run(1)[0]
        x = b.solution(0)
    finally:
        b.close()
    assert d.rp_info == 0
    res = rhs.reshape(-1) - O.neg_laplacian_pinned(x.reshape(-1), nx, ny, 1.0, 1.0)
    assert np.linalg.norm(res) <= 1e-8 * np.linalg.norm(rhs) * 1.0001
