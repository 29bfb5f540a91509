"""Host logic of the autotuner's pick (no GPU): candidates within 10% of the
fastest first timing are timed twice more and keep their best time; the
others are timed once; the pick is the fastest kept time; a tuner with no
finite timing raises."""
import pytest


def _script(monkeypatch, timings):
    """Replace the device timing with scripted values: timings maps a
    candidate key to the list of values its successive timings return."""
    from pyflow import autotune as A
    calls = {}

    def measure(nx, ny, precision, tile, *, warmup=16, rounds=64, seed=0):
        key = (tile.tile_rows, tile.max_workgroups, tile.pcg_per_round)
        k = calls.get(key, 0)
        calls[key] = k + 1
        seq = timings.get(key, [float("inf")])
        return A.Tile(tile.tile_rows, tile.max_workgroups, tile.pcg_per_round, seq[min(k, len(seq) - 1)])

    monkeypatch.setattr(A, "measure", measure)
    monkeypatch.setattr(A, "_key", lambda nx, ny, prec: f"{nx}x{ny}/test")
    monkeypatch.setattr(A, "_store_file", lambda key, tile: None)
    monkeypatch.setattr(A, "_cache", {})  # the process-wide cache is left as it was
    monkeypatch.delenv("PYFLOW_TUNE_S3", raising=False)
    return calls


def test_close_candidates_are_timed_again(monkeypatch):
    from pyflow import autotune as A
    n = 256
    cands = [(c.tile_rows, c.max_workgroups, c.pcg_per_round) for c in A.candidates(n, n, "mixed")]
    assert len(cands) >= 3
    a, b, c = cands[0], cands[1], cands[2]
    # a looks fastest once and then slower; b is steady and within 10%; c is 50% slower.
    timings = {a: [10.0, 12.0, 11.5], b: [10.8, 10.1, 10.9], c: [15.0, 9.0, 9.0]}
    timings.update({k: [100.0] for k in cands[3:]})
    calls = _script(monkeypatch, timings)
    best = A.tune(n, n, "mixed", force=True)
    assert calls[a] == 3 and calls[b] == 3  # within 10% of 10.0: timed twice more
    assert calls[c] == 1 and all(calls[k] == 1 for k in cands[3:])  # not close: once
    # kept times: a 10.0, b 10.1 -> a is picked with its best time
    assert (best.tile_rows, best.max_workgroups, best.pcg_per_round) == a
    assert best.us_per_iteration == 10.0


def test_retiming_can_change_the_pick(monkeypatch):
    from pyflow import autotune as A
    n = 256
    cands = [(c.tile_rows, c.max_workgroups, c.pcg_per_round) for c in A.candidates(n, n, "mixed")]
    a, b = cands[0], cands[1]
    timings = {a: [10.0, 10.4, 10.6], b: [10.9, 9.7, 10.5]}
    timings.update({k: [50.0] for k in cands[2:]})
    _script(monkeypatch, timings)
    best = A.tune(n, n, "mixed", force=True)
    assert (best.tile_rows, best.max_workgroups, best.pcg_per_round) == b
    assert best.us_per_iteration == 9.7


def test_no_finite_timing_raises(monkeypatch):
    from pyflow import autotune as A
    _script(monkeypatch, {})
    with pytest.raises(RuntimeError, match="no working candidate"):
        A.tune(64, 64, "mixed", force=True)
