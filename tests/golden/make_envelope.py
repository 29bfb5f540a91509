"""Measure the reduction-order sensitivity of each golden step trajectory.

    python tests/golden/make_envelope.py

The FP64 oracle reproduces the reference bit for bit (tests/test_oracle_golden.py).
This script reruns every golden step case with the oracle's dot products,
norms, sums and means replaced by exactly rounded ones (math.fsum): an equally
valid summation order. The per-quantity relative change is the intrinsic
parity floor of the reference path (CG amplifies last-bit differences), and
the GPU parity tests bound the HIP path's deviation by a fixed multiple of it.
Writes tests/golden/golden_envelope.json. CPU only; no reference import.
"""
from __future__ import annotations

import json
import math
import os
import sys
import types
from types import SimpleNamespace

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyflow_oracle as O  # noqa: E402


def _exact_numpy():
    real = O.np

    class Proxy(types.ModuleType):
        def __getattr__(self, k):
            return getattr(real, k)

    px = Proxy("np_exact")
    px.dot = lambda a, b: math.fsum((np.asarray(a) * np.asarray(b)).ravel().tolist())
    px.mean = lambda a: math.fsum(np.asarray(a).ravel().tolist()) / np.asarray(a).size
    px.sum = lambda a: math.fsum(np.asarray(a).ravel().tolist())
    px.linalg = SimpleNamespace(norm=lambda a: math.sqrt(math.fsum((np.asarray(a).ravel() ** 2).tolist())))
    return px


def main():
    s = np.load(os.path.join(HERE, "golden_steps.npz"))
    keys = json.loads(str(s["step_keys"]))
    real = O.np
    O.np = _exact_numpy()
    out = {}
    try:
        for name in json.loads(str(s["step_cases"])):
            if name == "nan10":
                continue
            cfg = SimpleNamespace(**json.loads(str(s[f"st_{name}_cfg"])))
            u0 = s[f"st_{name}_u0"]
            ny, nx = u0.shape[0] - 2, u0.shape[1] - 2
            st = O.allocate_state(nx, ny)
            st.fields["u"][:] = u0
            st.fields["v"][:] = s[f"st_{name}_v0"]
            D = s[f"st_{name}_diag"]
            rows = []
            for it in range(D.shape[0]):
                res, diag = O.step(cfg, st, it)
                rows.append([float(diag[k]) if k in diag else float(res[k]) for k in keys])
            rows = np.array(rows)
            with np.errstate(all="ignore"):
                rel = np.where(D != 0, np.abs(rows - D) / np.abs(D), np.abs(rows - D))
            env = {k: float(np.max(rel[:, i])) for i, k in enumerate(keys)}
            for f in ("u", "v", "p"):
                ref = s[f"st_{name}_{f}"][-1]
                d = st.fields[f] - ref
                env[f"field_{f}_maxrel"] = float(np.max(np.abs(d)) / max(np.max(np.abs(ref)), 1e-300))
                env[f"field_{f}_l2rel"] = float(np.linalg.norm(d) / max(np.linalg.norm(ref), 1e-300))
            out[name] = env
            print(name, {k: f"{v:.1e}" for k, v in env.items() if v > 0}, flush=True)
    finally:
        O.np = real
    with open(os.path.join(HERE, "golden_envelope.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
