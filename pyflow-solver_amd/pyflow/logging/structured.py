"""JSON-lines step records.

Interface kept from the reference's ``logging/structured.py:14-56`` (the
solver creates ``JsonlLogger(path=..., stream=...)`` at ``solvers/solver.py:175``
and the test suite calls ``close_all_jsonl_logs()`` after every test):

* ``JsonlLogger(path=None, stream=None)`` writes to ``stream`` when one is
  given, otherwise appends to ``path`` (its directory is created); with
  neither it raises ``ValueError``.
* ``log(record)`` writes ``record`` as one compact JSON object per line
  (separators ``","`` and ``":"``), adding ``"ts": time.time()`` to the record
  when it has no ``ts`` key, and flushes the sink.
* ``close()`` closes a file the logger opened itself (never a caller's
  stream); loggers are also context managers.
* ``close_all_jsonl_logs()`` closes every logger still alive.
"""
from __future__ import annotations

import json
import os
import threading
import time
import weakref
from typing import IO, Any

__all__ = ["JsonlLogger", "close_all_jsonl_logs"]

_to_line = json.JSONEncoder(separators=(",", ":")).encode

# One lock for every logger: two loggers may share a stream, and a line must
# never be split by another thread's line.
_emit_lock = threading.Lock()

# Live loggers by id; an entry removes itself when its logger is collected.
_live: dict[int, "weakref.ref[JsonlLogger]"] = {}


def _unregister(key: int):
    return lambda _dead: _live.pop(key, None)


class JsonlLogger:
    """Appends one JSON object per line to a file or a text stream."""

    def __init__(self, path: str | None = None, stream: IO[str] | None = None):
        if stream is None and path is None:
            raise ValueError("JsonlLogger requires either path or stream")
        self._own_file: IO[str] | None = None
        if stream is None:
            parent = os.path.dirname(path)
            os.makedirs(parent or ".", exist_ok=True)
            # line buffering: each record reaches the file when its line ends
            self._own_file = open(path, "a", encoding="utf-8", buffering=1)
        self._target: IO[str] = stream if stream is not None else self._own_file
        key = id(self)
        _live[key] = weakref.ref(self, _unregister(key))

    def log(self, record: dict[str, Any]) -> None:
        if "ts" not in record:
            record["ts"] = time.time()
        text = _to_line(record) + "\n"
        with _emit_lock:
            self._target.write(text)
            flush = getattr(self._target, "flush", None)
            if flush is not None:
                try:
                    flush()
                except (OSError, ValueError):
                    pass

    def close(self) -> None:
        f = self._own_file
        if f is not None and not f.closed:
            try:
                f.close()
            except OSError:
                pass

    def __enter__(self) -> "JsonlLogger":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def close_all_jsonl_logs() -> None:
    for ref in list(_live.values()):
        logger = ref()
        if logger is not None:
            logger.close()
