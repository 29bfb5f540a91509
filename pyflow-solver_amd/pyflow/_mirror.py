"""Device mirror of a ``State``: one device cavity plus the host/device sync rules.

The arithmetic of a cavity is fixed by its precision and by the tiling the
library resolved when the cavity was created (region height, workgroups,
iterations per round, FP64 iterate, inner tolerance): the tiling sets the
order of the FP64 partial sums. Both are recorded in
``state.meta['kernel_config']`` the first time a state reaches the device, are
written into checkpoints, and are applied again to every later step of that
state, so a resumed run repeats the uninterrupted one bit for bit whatever
``PYFLOW_*`` environment or tuner pick the new process has.
"""
from __future__ import annotations

import weakref

import numpy as np

from . import _hip as H
from ._device import DeviceCavity, precision_code

__all__ = ["StateMirror", "mirror_for", "config_precision"]

_DEVICE_FIELDS = ("u", "v", "p")


_PINNED_INT = ("tile_rows", "max_workgroups", "pcg_per_round", "iterate64", "deflate")


def _precision_name(prec: int) -> str:
    return "mixed" if prec == H.PREC_MIXED else "fp64"


def config_precision(cfg, state=None) -> int:
    """The config's ``precision`` if it sets one, else the precision recorded
    with the state (a resumed checkpoint), else ``PYFLOW_PRECISION``."""
    explicit = getattr(cfg, "precision", None)
    if explicit is None and state is not None:
        explicit = (state.meta.get("kernel_config") or {}).get("precision")
    return precision_code(explicit)


def _pin(state, params: H.CavityParams, prec: int) -> None:
    """Apply the state's recorded kernel config to ``params`` (in place)."""
    kc = state.meta.get("kernel_config")
    if not kc or kc.get("precision") != _precision_name(prec):
        return  # nothing recorded, or another precision was asked for explicitly
    if (int(kc["nx"]), int(kc["ny"])) != (params.nx, params.ny):
        raise ValueError(f"state.meta['kernel_config'] was recorded for a {kc['nx']}x{kc['ny']} grid, "
                         f"not {params.nx}x{params.ny}")
    for k in _PINNED_INT:
        setattr(params, k, int(kc.get(k, 0)))  # a key missing from an older record: the library default
    params.inner_rtol = float(kc["inner_rtol"])


def _record(state, mirror: "StateMirror") -> None:
    r = mirror.cav.resolved_params()
    kc = {"precision": _precision_name(mirror.prec), "nx": int(r.nx), "ny": int(r.ny),
          **{k: int(getattr(r, k)) for k in _PINNED_INT}, "inner_rtol": float(r.inner_rtol)}
    old = state.meta.get("kernel_config")
    if old and old.get("precision") == kc["precision"] and any(old[k] != kc[k] for k in kc):
        raise RuntimeError(f"the device cavity runs with {kc}, not the recorded kernel config {old}")
    state.meta["kernel_config"] = kc


class StateMirror:
    """Keeps ``state.fields['u'|'v'|'p']`` interiors on the GPU between steps."""

    def __init__(self, state, params: H.CavityParams, prec: int):
        self._state = weakref.ref(state)
        self.cav = DeviceCavity(params, prec)
        self.prec = prec
        self.nx, self.ny = params.nx, params.ny
        self.host_valid = True     # host arrays hold the latest values
        self.dev_valid = False     # device holds the latest values
        self.exposed = False       # arrays handed out since the last upload

    # called by State/_Fields ---------------------------------------------
    def host_access(self):
        self.ensure_host()
        self.exposed = True

    def host_replaced(self):
        self.dev_valid = False
        self.host_valid = True

    def ensure_host(self):
        if self.host_valid:
            return
        st = self._state()
        if st is None:
            return
        arrs = []
        for k in _DEVICE_FIELDS:
            arrs.append(st.fields.raw(k) if k in st.fields else None)
        tmp = []
        for a in arrs:
            if a is None or (a.dtype == np.float64 and a.flags.c_contiguous and a.shape == (self.ny + 2, self.nx + 2)):
                tmp.append(a)
            else:
                tmp.append(np.zeros((self.ny + 2, self.nx + 2)))
        self.cav.download(*tmp, ghost=True)
        for a, t in zip(arrs, tmp):
            if a is not None and t is not a:
                a[1:-1, 1:-1] = t[1:-1, 1:-1]
        self.host_valid = True

    # called by step/solve ---------------------------------------------------
    def before(self, params: H.CavityParams):
        st = self._state()
        self.cav.set_params(params)
        if (not self.dev_valid) or self.exposed or st._external_refs():
            self.ensure_host()
            arrs = [st.fields.raw(k) if k in st.fields else None for k in _DEVICE_FIELDS]
            if arrs[0] is None or arrs[1] is None:
                raise KeyError("state needs 'u' and 'v' fields")
            for a in arrs:
                if a is not None and a.shape != (self.ny + 2, self.nx + 2):
                    raise ValueError(f"field shape {a.shape} != {(self.ny + 2, self.nx + 2)}")
            self.cav.upload(*arrs, ghost=True)
            self.dev_valid = True
            self.exposed = False

    def after(self):
        st = self._state()
        self.host_valid = False
        if st is not None and st._external_refs():
            self.ensure_host()


def mirror_for(state, params: H.CavityParams, prec: int) -> StateMirror:
    """The state's device cavity for ``params``; ``params`` gets the state's
    recorded kernel config (pass the same object to ``before``)."""
    _pin(state, params, prec)
    m = state._dev
    if m is not None and (m.nx, m.ny, m.prec) == (params.nx, params.ny, prec):
        return m
    if m is not None:
        m.ensure_host()
        m.cav.close()
        state._dev = None
    m = StateMirror(state, params, prec)
    state._dev = m
    _record(state, m)
    return m
