"""Simulation state (reference: core/ghost_fields.py:8-60).

``State`` keeps the reference's contract: ``fields`` maps names to float64
``(ny+2, nx+2)`` arrays with a ghost ring, ``interior_view`` is
``arr[1:-1, 1:-1]``, ``meta`` is a free dict, and ``step`` mutates the same
object in place.

The compute copy lives on the GPU. Between steps the host arrays are synced
lazily: a step leaves the fields on the device and marks the host copy stale;
any access through ``state.fields[...]`` (or iteration over it) downloads the
interiors into the existing arrays first, in place, so references already
handed out see the new values. A step uploads the host arrays again whenever
they may have been changed: after any access through ``fields`` or while any
array (or a view of it) is still referenced outside the state.
"""
from __future__ import annotations

import sys
from collections.abc import Mapping
from dataclasses import dataclass, field
from typing import Any

import numpy as np

__all__ = ["State", "allocate_state", "interior_view", "ghost_shapes"]


class _Fields(dict):
    """dict of field arrays that syncs from the device before handing them out."""

    __slots__ = ("_owner",)

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self._owner = None

    def _touch(self):
        st = self._owner
        if st is not None and st._dev is not None:
            st._dev.host_access()

    def __getitem__(self, key):
        self._touch()
        return dict.__getitem__(self, key)

    def get(self, key, default=None):
        self._touch()
        return dict.get(self, key, default)

    def __setitem__(self, key, value):
        dict.__setitem__(self, key, value)
        st = self._owner
        if st is not None and st._dev is not None:
            st._dev.host_replaced()

    def values(self):
        self._touch()
        return dict.values(self)

    def items(self):
        self._touch()
        return dict.items(self)

    def __iter__(self):
        return dict.__iter__(self)

    def raw(self, key):
        return dict.__getitem__(self, key)


@dataclass
class State:
    """Container for simulation fields and lightweight metadata."""
    nx: int
    ny: int
    fields: dict[str, np.ndarray]
    meta: dict[str, Any] = field(default_factory=dict)
    _dev: Any = field(default=None, repr=False, compare=False)

    def __post_init__(self):
        if not isinstance(self.fields, _Fields):
            self.fields = _Fields(self.fields)
        self.fields._owner = self

    def __setattr__(self, name, value):
        if name == "fields" and not isinstance(value, _Fields):
            value = _Fields(value)
            object.__setattr__(self, name, value)
            value._owner = self
            if self.__dict__.get("_dev") is not None:
                self._dev.host_replaced()
            return
        object.__setattr__(self, name, value)

    def arr(self, name: str) -> np.ndarray:
        return self.fields[name]

    def __getitem__(self, key: str) -> np.ndarray:  # type: ignore[override]
        return self.fields[key]

    def __setitem__(self, key: str, value: np.ndarray) -> None:  # type: ignore[override]
        self.fields[key] = value

    def get(self, key: str, default=None):
        if key in self.fields:
            return self.fields[key]
        return self.meta.get(key, default)

    def set_meta(self, key: str, value: Any) -> None:
        self.meta[key] = value

    def keys(self):
        return self.fields.keys()

    def items(self):
        return self.fields.items()

    def values(self):
        return self.fields.values()

    # -- device mirror support (not part of the reference API) -------------
    def _raw(self, name: str) -> np.ndarray:
        return self.fields.raw(name)

    def _host_readonly(self) -> dict[str, np.ndarray]:
        """Current host arrays without marking them as handed out (checkpoints)."""
        if self._dev is not None:
            self._dev.ensure_host()
        return dict(dict.items(self.fields))

    def _external_refs(self) -> bool:
        """True when any field array is referenced outside this state."""
        for k in dict.keys(self.fields):
            a = dict.__getitem__(self.fields, k)
            # one reference from the dict, one from getrefcount's argument,
            # one from the local `a`
            if sys.getrefcount(a) > 3:
                return True
        return False


def allocate_state(nx: int, ny: int, fields: tuple[str, ...] = ("u", "v", "p")) -> State:
    """Allocate a State with named float64 fields including ghost cells."""
    shape = (ny + 2, nx + 2)
    arrs = {k: np.zeros(shape, dtype=float) for k in fields}
    return State(nx=nx, ny=ny, fields=arrs)


def interior_view(arr: np.ndarray) -> np.ndarray:
    return arr[1:-1, 1:-1]


def ghost_shapes(state: Mapping[str, np.ndarray] | State) -> dict[str, tuple[int, int]]:
    if isinstance(state, State):
        items = dict.items(state.fields)
    else:
        items = state.items()
    return {k: v.shape for k, v in items if isinstance(v, np.ndarray)}
