"""ctypes binding of ``libpyflow_hip.so`` (C ABI declared in include/pyflow_hip.h).

The library is the only compute path of this package: there is no CPU
fallback. Importing this module never touches the GPU; ``lib()`` loads the
shared object and raises ``HipUnavailableError`` when it is missing, and
``require_gpu()`` raises when no ROCm device is visible.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

__all__ = [
    "HipUnavailableError", "PfError", "lib", "require_gpu", "CavityParams", "StepDiag",
    "PREC_FP64", "PREC_MIXED", "SCHEME_UPWIND", "SCHEME_QUICK", "SCHEME_NONE",
    "OP_NEG_LAP_PINNED", "OP_POS_LAP_DIRICHLET", "MODE_STEP", "MODE_PROJECT", "MODE_SOLVE",
    "OPK", "check", "library_path", "EXPORTED_SYMBOLS",
]

PREC_FP64, PREC_MIXED = 0, 1
SCHEME_UPWIND, SCHEME_QUICK, SCHEME_NONE = 0, 1, 2
OP_NEG_LAP_PINNED, OP_POS_LAP_DIRICHLET = 0, 1
MODE_STEP, MODE_PROJECT, MODE_SOLVE = 0, 1, 2
OPK = {
    "grad_coherent": 0, "div_coherent": 1, "lap_coherent": 2, "div_central": 3,
    "adv_quick": 4, "adv_upwind": 5, "neg_lap_pinned": 6, "pos_lap_dirichlet": 7,
}

# Every symbol include/pyflow_hip.h declares.
EXPORTED_SYMBOLS = (
    "pf_abi_version", "pf_last_error", "pf_sizeof_params", "pf_sizeof_diag", "pf_device_init", "pf_batch_workspace_bytes",
    "pf_batch_create", "pf_batch_destroy", "pf_batch_create_empty", "pf_batch_add",
    "pf_batch_remove", "pf_batch_set_params", "pf_batch_upload", "pf_batch_download",
    "pf_batch_zero_fields", "pf_batch_set_meta", "pf_batch_set_rhs", "pf_batch_get_solution",
    "pf_batch_set_dt", "pf_batch_run", "pf_batch_launch", "pf_batch_poll", "pf_batch_diag",
    "pf_batch_centerline", "pf_batch_rounds", "pf_batch_set_timing", "pf_batch_timing", "pf_op_apply",
    "pf_batch_start", "pf_batch_advance", "pf_batch_finished", "pf_batch_set_graph",
    "pf_batch_resolved_params", "pf_solve_noalloc",
)


class HipUnavailableError(RuntimeError):
    """The HIP library or a ROCm GPU is not available."""


class PfError(RuntimeError):
    """A libpyflow_hip call returned an error status."""


class CavityParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("scheme", C.c_int32), ("lid_enabled", C.c_int32),
        ("jacobi", C.c_int32), ("lin_maxiter", C.c_int32), ("op", C.c_int32), ("mode", C.c_int32),
        ("lid_velocity", C.c_double), ("dx", C.c_double), ("dy", C.c_double), ("nu", C.c_double),
        ("cfl_target", C.c_double), ("cfl_growth", C.c_double), ("lin_tol", C.c_double),
        ("inv_diag", C.c_double), ("inner_rtol", C.c_double),
        ("tile_rows", C.c_int32), ("max_workgroups", C.c_int32), ("pcg_per_round", C.c_int32),
        ("iterate64", C.c_int32),
        ("deflate", C.c_int32),
        ("reserved", C.c_int32),
    ]


class StepDiag(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("cfl", C.c_double), ("umax", C.c_double), ("vmax", C.c_double),
        ("dt_prev", C.c_double), ("hist", C.c_double * 5), ("hist_len", C.c_int32),
        ("safeguard", C.c_int32), ("div_before", C.c_double), ("div_initial", C.c_double),
        ("div_norm", C.c_double), ("rhs_mean", C.c_double), ("bnorm", C.c_double),
        ("rp_residual", C.c_double), ("ru", C.c_double), ("rv", C.c_double),
        ("continuity", C.c_double), ("rp_info", C.c_int32), ("rp_iterations", C.c_int32),
        ("rp_converged", C.c_int32), ("nan_detected", C.c_int32), ("outer_passes", C.c_int32),
        ("steps_done", C.c_int32), ("rp_iterations_run", C.c_int32), ("outer_passes_run", C.c_int32),
    ]

    def history(self):
        return [float(self.hist[k]) for k in range(self.hist_len)]


_lock = threading.Lock()
_lib = None


def library_path() -> str:
    override = os.environ.get("PYFLOW_HIP_LIB")
    if override:
        return override
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libpyflow_hip.so")


def _declare(L):
    P, I, I64, D, Z = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_size_t
    PP = C.POINTER(C.c_void_p)
    sig = {
        "pf_abi_version": ([], I),
        "pf_last_error": ([], C.c_char_p),
        "pf_sizeof_params": ([], Z),
        "pf_sizeof_diag": ([], Z),
        "pf_device_init": ([I], I),
        "pf_batch_workspace_bytes": ([I, I, C.POINTER(CavityParams), C.POINTER(Z)], I),
        "pf_batch_create": ([I, I, C.POINTER(CavityParams), P, Z, P, PP], I),
        "pf_batch_destroy": ([P], I),
        "pf_batch_create_empty": ([I, I, I64, P, PP], I),
        "pf_batch_add": ([P, C.POINTER(CavityParams), P, Z, C.POINTER(I)], I),
        "pf_batch_remove": ([P, I], I),
        "pf_batch_set_params": ([P, I, C.POINTER(CavityParams)], I),
        "pf_batch_resolved_params": ([P, I, C.POINTER(CavityParams)], I),
        "pf_batch_upload": ([P, I, P, P, P, I64, I], I),
        "pf_batch_download": ([P, I, P, P, P, I64, I], I),
        "pf_batch_zero_fields": ([P], I),
        "pf_batch_set_meta": ([P, I, D, P, I], I),
        "pf_batch_set_rhs": ([P, I, P, I64], I),
        "pf_batch_get_solution": ([P, I, P, I64], I),
        "pf_batch_set_dt": ([P, I, D], I),
        "pf_batch_run": ([P, I, P, P], I),
        "pf_batch_launch": ([P, I, P], I),
        "pf_batch_poll": ([P, C.POINTER(I)], I),
        "pf_batch_diag": ([P, P], I),
        "pf_batch_centerline": ([P, I, P], I),
        "pf_batch_rounds": ([P, C.POINTER(I64)], I),
        "pf_batch_set_timing": ([P, I], I),
        "pf_batch_set_graph": ([P, I], I),
        "pf_batch_timing": ([P, C.POINTER(D), C.POINTER(I64)], I),
        "pf_op_apply": ([I, I, I, D, D, P, P, P, P, P, P], I),
        "pf_batch_start": ([P, I, I], I),
        "pf_batch_advance": ([P, I64], I),
        "pf_batch_finished": ([P, P], I),
        "pf_solve_noalloc": ([I, C.POINTER(CavityParams), P, P, P, P], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


def _torch_runtime_first() -> None:
    # On a machine with a GPU, torch initialises its HIP runtime before this
    # library is loaded (the library registers its code objects when it is
    # loaded). Loaded the other way round, the library's hipSetDevice found no
    # device in that process.
    try:
        import torch
    except ImportError:  # pragma: no cover - torch is a dependency
        return
    torch.cuda.is_available()


def lib():
    """Load (once) and return the ctypes handle of libpyflow_hip.so."""
    global _lib
    if _lib is not None:
        return _lib
    _torch_runtime_first()
    with _lock:
        if _lib is None:
            path = library_path()
            if not os.path.exists(path):
                raise HipUnavailableError(
                    f"libpyflow_hip.so not found at {path}; build it with "
                    "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
            try:
                L = C.CDLL(path)
            except OSError as e:  # pragma: no cover - environment specific
                raise HipUnavailableError(f"cannot load {path}: {e}") from e
            _declare(L)
            if L.pf_abi_version() != 2:
                raise HipUnavailableError("libpyflow_hip ABI version mismatch")
            if L.pf_sizeof_params() != C.sizeof(CavityParams) or L.pf_sizeof_diag() != C.sizeof(StepDiag):
                raise HipUnavailableError("libpyflow_hip struct layout mismatch")
            _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().pf_last_error().decode(errors="replace")
        raise PfError(f"{what or 'libpyflow_hip'} failed ({rc}): {msg}")


_gpu_checked = False


def require_gpu():
    """Raise unless a ROCm GPU is visible; returns the torch module."""
    global _gpu_checked
    import torch
    if not _gpu_checked:
        if not torch.cuda.is_available():
            raise HipUnavailableError(
                "PyFlow-HIP needs a ROCm GPU (torch.cuda.is_available() is False); "
                "there is no CPU fallback")
        lib()
        check(lib().pf_device_init(torch.cuda.current_device()), "pf_device_init")
        _gpu_checked = True
    return torch
