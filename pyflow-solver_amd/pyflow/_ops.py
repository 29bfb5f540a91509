"""Standalone float64 operators on the GPU (pf_op_apply), NumPy in / NumPy out."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _hip as H

__all__ = ["apply_op"]


def apply_op(name: str, shape, dx: float, dy: float, *inputs: np.ndarray, nout: int = 1):
    torch = H.require_gpu()
    ny, nx = int(shape[0]), int(shape[1])
    dev = torch.device("cuda", torch.cuda.current_device())
    ins = [torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(ny, nx))).to(dev)
           for a in inputs]
    outs = [torch.empty((ny, nx), dtype=torch.float64, device=dev) for _ in range(nout)]
    ptr = [C.c_void_p(t.data_ptr()) for t in ins] + [None] * (3 - len(ins))
    optr = [C.c_void_p(t.data_ptr()) for t in outs] + [None] * (2 - len(outs))
    stream = torch.cuda.current_stream().cuda_stream
    H.check(H.lib().pf_op_apply(H.OPK[name], nx, ny, float(dx), float(dy), ptr[0], ptr[1], ptr[2],
                                optr[0], optr[1], C.c_void_p(stream)), f"pf_op_apply({name})")
    res = [o.cpu().numpy() for o in outs]
    return res[0] if nout == 1 else tuple(res)
