"""The allocation-free solve (pf_solve_noalloc) against the same solve through
the batch entry points (pf_batch_create + set_rhs + run + get_solution):
same bits, two calls in a row with the second on a smaller grid, both
precisions, the Dirichlet operator, the deflation step (its vector y in
the scratch array too) and the three-iteration round; called through the C
ABI with device pointers."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rhs(nx, ny, seed, pinned=True):
    r = np.random.default_rng(seed).standard_normal((ny, nx))
    if pinned:
        r -= r.mean()
        r[0, 0] = 0.0
    return r


def _params(nx, ny, op, tol, deflate=0, s=0):
    from pyflow import _hip as H
    from pyflow._device import cavity_params
    p = cavity_params(nx, ny, dx=1.0, dy=1.0, lin_tol=tol, lin_maxiter=20000, jacobi=(op == H.OP_NEG_LAP_PINNED),
                      op=op, mode=H.MODE_SOLVE, deflate=deflate)
    p.pcg_per_round = s
    return p


def _noalloc(prec, p, rhs):
    import torch
    from pyflow import _hip as H
    dev = torch.device("cuda", torch.cuda.current_device())
    b = torch.from_numpy(np.ascontiguousarray(rhs)).to(dev)
    x = torch.full_like(b, float("nan"))  # every element must be written
    d = H.StepDiag()
    stream = torch.cuda.current_stream().cuda_stream
    H.check(H.lib().pf_solve_noalloc(prec, C.byref(p), C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
                                     C.c_void_p(stream), C.cast(C.byref(d), C.c_void_p)), "pf_solve_noalloc")
    return x.cpu().numpy(), d


def _batch(prec, p, rhs):
    from pyflow._device import DeviceBatch
    b = DeviceBatch([p], prec)
    b.set_rhs(0, rhs)
    d = b.run(1)[0]
    x = b.solution(0)
    b.close()
    return x, d


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_solve_equals_batch_solve(precision):
    from pyflow import _hip as H
    from pyflow._device import precision_code
    prec = precision_code(precision)
    cases = [(320, 240, H.OP_NEG_LAP_PINNED, 1e-6, 5, 2, 0), (300, 200, H.OP_NEG_LAP_PINNED, 1e-9, 1, 0, 0),
             (256, 192, H.OP_NEG_LAP_PINNED, 1e-9, 7, 0, 3), (96, 70, H.OP_NEG_LAP_PINNED, 1e-9, 2, 0, 0),
             (64, 40, H.OP_POS_LAP_DIRICHLET, 1e-10, 3, 0, 0), (40, 36, H.OP_NEG_LAP_PINNED, 1e-6, 6, 2, 0),
             (33, 17, H.OP_NEG_LAP_PINNED, 1e-8, 4, 0, 0)]
    for nx, ny, op, tol, seed, dfl, s in cases:  # consecutive calls, later ones smaller
        rhs = _rhs(nx, ny, seed, pinned=(op == H.OP_NEG_LAP_PINNED))
        print(f"[noalloc] {precision} {nx}x{ny} op={op} deflate={dfl} pcg_per_round={s}", flush=True)
        xa, da = _noalloc(prec, _params(nx, ny, op, tol, dfl, s), rhs)
        xb, db = _batch(prec, _params(nx, ny, op, tol, dfl, s), rhs)
        assert np.array_equal(xa, xb), (precision, nx, ny)
        assert (da.rp_iterations, da.rp_info, da.rp_converged, da.outer_passes) == \
               (db.rp_iterations, db.rp_info, db.rp_converged, db.outer_passes)
        assert da.rp_residual == db.rp_residual and da.bnorm == db.bnorm
        assert da.rp_converged and da.rp_residual <= tol * da.bnorm
    # a zero right-hand side: x = b = 0, no iteration (SciPy's early return)
    xz, dz = _noalloc(prec, _params(40, 30, H.OP_NEG_LAP_PINNED, 1e-9), np.zeros((30, 40)))
    assert np.array_equal(xz, np.zeros((30, 40))) and dz.rp_iterations == 0 and dz.rp_converged
