"""Chunks of rounds between host checks. A run of one stepping cavity sizes
its chunks from the step ends it has seen (PYFLOW_TRACK, on by default);
other runs double their chunks from 16. The chunking never changes a result:
rounds queued after a cavity finished leave its state untouched."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 256


def _batch(track):
    from pyflow._device import DeviceBatch, cavity_params
    p = cavity_params(N, N, dx=1.0, dy=1.0, scheme="quick", lid_velocity=1.0, Re=1000.0, lin_tol=1e-6,
                      lin_maxiter=10000)
    old = os.environ.get("PYFLOW_TRACK")
    os.environ["PYFLOW_TRACK"] = "1" if track else "0"  # read when the batch is created
    try:
        return DeviceBatch([p], "mixed")
    finally:
        if old is None:
            del os.environ["PYFLOW_TRACK"]
        else:
            os.environ["PYFLOW_TRACK"] = old


def _fields(b):
    u = np.zeros((N + 2, N + 2)); v = np.zeros_like(u); p = np.zeros_like(u)
    b.download(0, u, v, p)
    return u, v, p


def test_tracked_chunks_same_bits_fewer_rounds():
    """Two warm-up steps, then 20 steps in one run, then one step: tracked and
    doubling chunks give the same bits after every run. The tracked 20-step
    run launches no more rounds than the doubling one and stays within 64
    rounds of what the cavity executed; the one-step run after it is sized
    from the rounds per step, not from the whole 20-step run."""
    res = {}
    for track in (True, False):
        b = _batch(track)
        try:
            out = []
            for k, n in enumerate((2, 20, 1)):
                print(f"[chunk] track={track} run {k}: {n} steps", flush=True)
                d = b.run(n)[0]
                assert d.steps_done == n and d.rp_converged and not d.nan_detected
                out.append((b.rounds(), _fields(b)))
            res[track] = out
        finally:
            b.close()
    for k in range(3):
        for f0, f1 in zip(res[True][k][1], res[False][k][1]):
            assert np.array_equal(f0, f1), k
    r_track, r_double = res[True][1][0], res[False][1][0]
    print(f"[chunk] 20-step run: tracked {r_track} rounds, doubling {r_double}", flush=True)
    assert r_track <= r_double
    # rounds per step of the 20-step run, from the tracked run's launch count
    per_step = r_track / 20.0
    r_one = res[True][2][0]
    print(f"[chunk] one-step run after it: {r_one} rounds ({per_step:.0f} per step above)", flush=True)
    assert r_one <= 1.5 * per_step + 64
    assert res[False][2][0] <= 1.5 * per_step + 64


def test_tracked_run_ends_close_to_the_last_round():
    """The tracked run's launch count exceeds the rounds the cavity needed by
    at most 80: a second batch runs the same steps in chunks of at most 16
    rounds after its first, so its count exceeds the need by at most 16."""
    b = _batch(True)
    try:
        b.run(2)
        b.run(12)
        tracked = b.rounds()
        u_t = _fields(b)
    finally:
        b.close()
    os.environ["PYFLOW_MAX_CHUNK"] = "16"
    try:
        c = _batch(False)
    finally:
        del os.environ["PYFLOW_MAX_CHUNK"]
    try:
        c.run(2)
        c.run(12)
        fine = c.rounds()  # at most 16 past the last round the cavity executed
        u_c = _fields(c)
    finally:
        c.close()
    print(f"[chunk] 12-step run: tracked {tracked} rounds, 16-round chunks {fine}", flush=True)
    for f0, f1 in zip(u_t, u_c):
        assert np.array_equal(f0, f1)
    assert tracked <= fine + 64
