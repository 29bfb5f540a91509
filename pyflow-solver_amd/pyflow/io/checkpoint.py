"""Checkpoint I/O (reference: io/checkpoint.py:34-127).

Same ``.npz`` format: key ``meta`` (JSON string) and ``field_<name>`` arrays
with the ghost ring, atomic write (temp file, fsync, rename), SHA1 field
hashes (``sha1(arr.tobytes())[:16]``), ``config_hash`` and
``structured_config_hash``. Added meta keys (ignored by the reference's
loader), all taken from and restored into ``state.meta``:

* ``dt_prev`` and ``divergence_history``: the dt controller's memory, which
  the reference does not persist;
* ``kernel_config``: the precision and the tiling the device cavity ran with
  (region height, workgroups, PCG iterations per round, FP64 iterate, inner
  tolerance). The tiling fixes the order of the FP64 partial sums; the next
  step of a loaded state runs with it again, whatever the environment.

With them a restart continues the uninterrupted trajectory bit for bit.
"""
from __future__ import annotations

import hashlib
import io
import json
import os
import tempfile
from typing import Any, cast

import numpy as np

from ..core.ghost_fields import State

try:
    from ..config.model import SimulationConfig
except Exception:  # pragma: no cover
    SimulationConfig = None  # type: ignore

__all__ = ["save_checkpoint", "load_checkpoint", "save_checkpoint_bytes", "load_checkpoint_bytes"]

_RESTART_META = ("dt_prev", "divergence_history", "kernel_config")


def _hash_config(cfg: Any) -> str:
    try:
        if hasattr(cfg, "__dict__"):
            data = {k: v for k, v in vars(cfg).items() if not k.startswith("_")}
        else:
            data = {k: getattr(cfg, k) for k in dir(cfg) if not k.startswith("_")}
        blob = json.dumps(data, sort_keys=True, default=str).encode()
        return hashlib.sha1(blob).hexdigest()[:10]
    except Exception:
        return "unknown"


def _hash_array(arr: np.ndarray) -> str:
    return hashlib.sha1(arr.tobytes()).hexdigest()[:16]


def _fields(state: State) -> dict[str, np.ndarray]:
    getter = getattr(state, "_host_readonly", None)
    return getter() if getter is not None else dict(state.fields.items())


def _restart_meta(state: State, cfg: Any) -> dict[str, Any]:
    out: dict[str, Any] = {}
    if "dt_prev" in state.meta:
        out["dt_prev"] = float(state.meta["dt_prev"])
    if "divergence_history" in state.meta:
        out["divergence_history"] = [float(x) for x in state.meta["divergence_history"]]
    if state.meta.get("kernel_config"):
        out["kernel_config"] = dict(state.meta["kernel_config"])
    return out


def save_checkpoint(path: str, state: State, iteration: int, sim_time: float, cfg: Any) -> str:
    """Atomically persist simulation state (temp file -> fsync -> os.replace)."""
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    fields = _fields(state)
    field_hashes = {k: _hash_array(v) for k, v in fields.items()}
    cfg_hash = None
    if SimulationConfig is not None and isinstance(cfg, SimulationConfig):
        try:
            cfg_hash = cfg.config_hash
        except Exception:
            cfg_hash = None
    meta: dict[str, Any] = {
        "iteration": iteration,
        "time": sim_time,
        "nx": state.nx,
        "ny": state.ny,
        "config_hash": _hash_config(cfg),
        "structured_config_hash": cfg_hash,
        "field_hashes": field_hashes,
        "seed": getattr(cfg, "seed", None),
        "schema_version": 1,
    }
    meta.update(_restart_meta(state, cfg))
    tmp_fd, tmp_path = tempfile.mkstemp(suffix=".tmp", dir=os.path.dirname(path) or ".")
    try:
        with os.fdopen(tmp_fd, "wb") as f:
            arrays = {f"field_{k}": v for k, v in fields.items()}
            cast(Any, np.savez_compressed)(f, meta=json.dumps(meta), **arrays)
            f.flush()
            try:
                os.fsync(f.fileno())
            except OSError:
                pass
        os.replace(tmp_path, path)
    finally:
        if os.path.exists(tmp_path):
            try:
                os.remove(tmp_path)
            except Exception:
                pass
    return path


def _state_from(data) -> tuple[State, dict[str, Any]]:
    meta = json.loads(str(data["meta"]))
    fields = {k.removeprefix("field_"): data[k].copy() for k in data.files if k.startswith("field_")}
    st = State(nx=meta.get("nx"), ny=meta.get("ny"), fields=fields)
    for k in _RESTART_META:
        if k in meta:
            v = meta[k]
            st.meta[k] = list(v) if isinstance(v, list) else (dict(v) if isinstance(v, dict) else v)
    return st, meta


def load_checkpoint(path: str) -> tuple[State, dict[str, Any]]:
    with np.load(path) as data:
        return _state_from(data)


def save_checkpoint_bytes(state: State, iteration: int, sim_time: float, cfg: Any) -> bytes:
    fields = _fields(state)
    field_hashes = {k: _hash_array(v) for k, v in fields.items()}
    meta = {"iteration": iteration, "time": sim_time, "nx": state.nx, "ny": state.ny,
            "config_hash": _hash_config(cfg), "field_hashes": field_hashes, "seed": getattr(cfg, "seed", None),
            "schema_version": 1}
    meta.update(_restart_meta(state, cfg))
    bio = io.BytesIO()
    arrays = {f"field_{k}": v for k, v in fields.items()}
    cast(Any, np.savez_compressed)(bio, meta=json.dumps(meta), **arrays)
    return bio.getvalue()


def load_checkpoint_bytes(buf: bytes) -> tuple[State, dict[str, Any]]:
    with np.load(io.BytesIO(buf)) as data:
        return _state_from(data)
