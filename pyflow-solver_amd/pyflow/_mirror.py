"""Device mirror of a ``State``: one device cavity plus the host/device sync rules."""
from __future__ import annotations

import weakref

import numpy as np

from . import _hip as H
from ._device import DeviceCavity, precision_code

__all__ = ["StateMirror", "mirror_for", "config_precision"]

_DEVICE_FIELDS = ("u", "v", "p")


def config_precision(cfg) -> int:
    return precision_code(getattr(cfg, "precision", None))


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
    m = state._dev
    if m is not None and (m.nx, m.ny, m.prec) == (params.nx, params.ny, prec):
        return m
    if m is not None:
        m.ensure_host()
        m.cav.close()
        state._dev = None
    m = StateMirror(state, params, prec)
    state._dev = m
    return m
