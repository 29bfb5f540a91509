"""Measure how far each golden step trajectory moves under other summation orders.

    python tests/golden/make_envelope.py

The FP64 oracle reproduces the reference bit for bit (tests/test_oracle_golden.py).
This script reruns every golden step case with the oracle's dot products,
norms, sums and means computed in other, equally valid summation orders:

* ``fsum``: exactly rounded (math.fsum);
* ``seq`` / ``rev``: one running sum, first to last / last to first;
* ``wave64``: running sums over 64-element blocks, block totals summed in order
  (the shape of a per-lane accumulation followed by a fold);
* ``perm1`` / ``perm2``: running sums in two fixed random orders.

Products are rounded before they are summed, as in the device's FP64 partials.
For every case and quantity the envelope is the largest relative change from
the reference's own values over these orders: the intrinsic parity floor of
the reference path (CG amplifies last-bit differences). The GPU parity tests
bound the HIP path's deviation by a small multiple of it. For the C1 case
(``c1_64``, the run of the reference's own C1 stream) the envelope also covers
the centerline ``u[:, nx//2]`` of every step. Also stored, per
step, ``lin_tol * ||b||`` of the reference-order run: the stopping threshold
that a converged solve's true residual ``Rp_residual`` lies below.

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


def _running(x: np.ndarray) -> float:
    return float(np.cumsum(x)[-1]) if x.size else 0.0


def _wave64(x: np.ndarray) -> float:
    if x.size == 0:
        return 0.0
    pad = np.zeros((-x.size) % 64)
    blocks = np.concatenate([x, pad]).reshape(-1, 64)
    return _running(np.cumsum(blocks, axis=1)[:, -1])


_PERM_CACHE: dict[tuple[int, int], np.ndarray] = {}


def _perm(seed: int):
    def total(x: np.ndarray) -> float:
        key = (seed, x.size)
        if key not in _PERM_CACHE:
            _PERM_CACHE[key] = np.random.default_rng(seed).permutation(x.size)
        return _running(x[_PERM_CACHE[key]])
    return total


ORDERS = {
    "fsum": lambda x: math.fsum(x.tolist()),
    "seq": _running,
    "rev": lambda x: _running(x[::-1]),
    "wave64": _wave64,
    "perm1": _perm(1),
    "perm2": _perm(2),
}


def _numpy_with(total):
    real = O.np

    class Proxy(types.ModuleType):
        def __getattr__(self, k):
            return getattr(real, k)

    flat = lambda a: np.asarray(a, dtype=np.float64).ravel()  # noqa: E731
    px = Proxy("np_reordered")
    px.dot = lambda a, b: total(flat(a) * flat(b))
    px.mean = lambda a: total(flat(a)) / flat(a).size
    px.sum = lambda a: total(flat(a))
    px.linalg = SimpleNamespace(norm=lambda a: math.sqrt(total(flat(a) * flat(a))))
    return px


def _run(s, name, keys):
    cfg = SimpleNamespace(**json.loads(str(s[f"st_{name}_cfg"])))
    u0 = s[f"st_{name}_u0"]
    ny, nx = u0.shape[0] - 2, u0.shape[1] - 2
    st = O.allocate_state(nx, ny)
    st.fields["u"][:] = u0
    st.fields["v"][:] = s[f"st_{name}_v0"]
    rows, bnorm, center = [], [], []
    for it in range(s[f"st_{name}_diag"].shape[0]):
        res, diag = O.step(cfg, st, it)
        rows.append([float(diag[k]) if k in diag else float(res[k]) for k in keys])
        bnorm.append(float(diag["Rp_rhs_norm"]))
        center.append(st.fields["u"][1:-1, 1 + nx // 2].copy())
    return np.array(rows), st, bnorm, float(getattr(cfg, "lin_tol", 1e-10)), center


def _cli_centerlines():
    """u[:, nx//2] of every step of the reference's own C1 stream."""
    with open(os.path.join(HERE, "golden_cli_c1.jsonl")) as f:
        return [np.array(json.loads(l)["diagnostics"]["u_centerline"][1]) for l in f if l.strip()]


def main():
    s = np.load(os.path.join(HERE, "golden_steps.npz"))
    cli_center = _cli_centerlines()
    keys = json.loads(str(s["step_keys"]))
    real = O.np
    out = {}
    for name in json.loads(str(s["step_cases"])):
        if name == "nan10":
            continue
        D = s[f"st_{name}_diag"]
        _, _, bnorm, tol, _ = _run(s, name, keys)  # reference order (the oracle is bitwise the reference)
        env: dict[str, float] = {}
        per_order = {}
        for oname, total in ORDERS.items():
            O.np = _numpy_with(total)
            try:
                rows, st, _, _, center = _run(s, name, keys)
            finally:
                O.np = real
            with np.errstate(all="ignore"):
                rel = np.where(D != 0, np.abs(rows - D) / np.abs(D), np.abs(rows - D))
            cur = {k: float(np.max(rel[:, i])) for i, k in enumerate(keys)}
            for f in ("u", "v", "p"):
                ref = s[f"st_{name}_{f}"][-1]
                d = st.fields[f] - ref
                cur[f"field_{f}_maxrel"] = float(np.max(np.abs(d)) / max(np.max(np.abs(ref)), 1e-300))
                cur[f"field_{f}_l2rel"] = float(np.linalg.norm(d) / max(np.linalg.norm(ref), 1e-300))
            if name == "c1_64":  # the same run as the reference's C1 stream: centerline of every step
                cur["centerline_maxrel"] = max(float(np.max(np.abs(c - e)) / max(np.max(np.abs(e)), 1e-300))
                                               for c, e in zip(center, cli_center))
            per_order[oname] = cur
            for k, v in cur.items():
                env[k] = max(env.get(k, 0.0), v)
        env["rp_abs_bound"] = [tol * b for b in bnorm]
        env["orders"] = sorted(ORDERS)
        out[name] = env
        print(name, {k: f"{v:.1e}" for k, v in env.items() if isinstance(v, float) and v > 0}, flush=True)
    with open(os.path.join(HERE, "golden_envelope.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
