"""Command line front end: ``pyflow run | validate | show-config``.

Interface kept from the reference's ``cli.py:35-194``, which the dashboard and
the AI controller drive through a subprocess:

* Flags and defaults (all three commands): ``--nx 64 --ny 64 --re 100
  --lid-velocity 1 --scheme {quick,upwind}=quick --disable-advection --cfl 0.5
  --cfl-growth 1.05 --lin-tol 1e-10 --lin-maxiter 300 --checkpoint-interval 0
  (0 = none) --log-jsonl PATH --seed N --diagnostics``; ``run`` adds ``--steps
  100`` (an absolute bound on the iteration index), ``--checkpoint``,
  ``--restart``, ``--allow-hash-mismatch``, ``--json-stream``,
  ``--continuity-threshold``, ``--progress``, ``--assert-invariants``.
* Exit codes: 0 done, 2 invalid configuration, 3 restart refused because the
  checkpoint's config hash differs, 130 interrupted (after a best-effort final
  checkpoint).
* ``--json-stream`` line, one per step, compact separators::

      {"type":"step","iteration":..,"dt":..,"CFL":..,"continuity":..,
       "wall_time":..,"residuals":{"Ru","Rv","Rp","continuity"},
       "diagnostics":{every diagnostics key except iteration, dt, CFL,
       wall_time},"run_meta":{"seed":..}}   (run_meta only with --seed)

* A restart resumes at checkpoint iteration + 1; an unreadable checkpoint
  prints a message and starts from rest.

Added here: ``--precision {fp64,mixed}`` picks the GPU arithmetic of the step
(not part of the config hash).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
from typing import Any

from pydantic import ValidationError

from .config.model import ConfigError, SimulationConfig
from .core.ghost_fields import State, allocate_state
from .drivers.simulation_driver import SimulationDriver
from .residuals.manager import ResidualManager

__all__ = ["build_parser", "main", "stream_record"]

EXIT_OK, EXIT_BAD_CONFIG, EXIT_HASH_MISMATCH, EXIT_INTERRUPTED = 0, 2, 3, 130

# (flag, argparse keyword arguments): options every command accepts.
_MODEL_OPTIONS: tuple[tuple[str, dict[str, Any]], ...] = (
    ("--nx", {"type": int, "default": 64}),
    ("--ny", {"type": int, "default": 64}),
    ("--re", {"type": float, "default": 100.0}),
    ("--lid-velocity", {"type": float, "default": 1.0}),
    ("--scheme", {"choices": ["quick", "upwind"], "default": "quick"}),
    ("--disable-advection", {"action": "store_true"}),
    ("--cfl", {"type": float, "default": 0.5}),
    ("--cfl-growth", {"type": float, "default": 1.05}),
    ("--lin-tol", {"type": float, "default": 1e-10}),
    ("--lin-maxiter", {"type": int, "default": 300}),
    ("--checkpoint-interval", {"type": int, "default": 0}),
    ("--log-jsonl", {"type": str, "default": None}),
    ("--seed", {"type": int, "default": None}),
    ("--diagnostics", {"action": "store_true"}),
    ("--precision", {"choices": ["fp64", "mixed"], "default": None}),
)

# Options of ``run`` only.
_RUN_OPTIONS: tuple[tuple[str, dict[str, Any]], ...] = (
    ("--steps", {"type": int, "default": 100}),
    ("--checkpoint", {"type": str, "default": None}),
    ("--restart", {"type": str, "default": None}),
    ("--allow-hash-mismatch", {"action": "store_true"}),
    ("--json-stream", {"action": "store_true"}),
    ("--continuity-threshold", {"type": float, "default": None}),
    ("--progress", {"action": "store_true"}),
    ("--assert-invariants", {"action": "store_true"}),
)

_COMMANDS = (
    ("run", "Run a simulation", True),
    ("validate", "Validate and display normalized config", False),
    ("show-config", "Show normalized config (hash only)", False),
)

# Diagnostics keys that the stream line carries at its top level instead.
_TOP_LEVEL_KEYS = frozenset({"iteration", "dt", "CFL", "wall_time"})


def build_parser() -> argparse.ArgumentParser:
    parser = argparse.ArgumentParser(description="PyFlow CLI")
    commands = parser.add_subparsers(dest="command", required=True)
    for name, help_text, is_run in _COMMANDS:
        sp = commands.add_parser(name, help=help_text)
        for flag, spec in _MODEL_OPTIONS + (_RUN_OPTIONS if is_run else ()):
            sp.add_argument(flag, **spec)
    return parser


def _interval(ns: argparse.Namespace) -> int | None:
    """--checkpoint-interval with 0 meaning "no periodic checkpoints"."""
    return ns.checkpoint_interval or None


def _config(ns: argparse.Namespace) -> SimulationConfig:
    return SimulationConfig(
        nx=ns.nx, ny=ns.ny, Re=ns.re, lid_velocity=ns.lid_velocity,
        advection_scheme=ns.scheme, disable_advection=ns.disable_advection,
        cfl_target=ns.cfl, cfl_growth=ns.cfl_growth,
        lin_tol=ns.lin_tol, lin_maxiter=ns.lin_maxiter,
        checkpoint_interval=_interval(ns), log_path=ns.log_jsonl,
        diagnostics=ns.diagnostics, seed=ns.seed, turbulence_model=None,
    )


def stream_record(residuals: dict[str, Any], diag: dict[str, Any], seed: int | None = None) -> dict[str, Any]:
    """The ``--json-stream`` object of one step."""
    rec: dict[str, Any] = {"type": "step"}
    for key in ("iteration", "dt", "CFL"):
        rec[key] = diag.get(key)
    rec["continuity"] = residuals.get("continuity")
    rec["wall_time"] = diag.get("wall_time")
    rec["residuals"] = {key: residuals.get(key) for key in ("Ru", "Rv", "Rp", "continuity")}
    rec["diagnostics"] = {k: v for k, v in diag.items() if k not in _TOP_LEVEL_KEYS}
    if seed is not None:
        rec["run_meta"] = {"seed": seed}
    return rec


def _initial_state(ns: argparse.Namespace, cfg: SimulationConfig) -> tuple[State, int] | int:
    """(state, first iteration), or the exit code of a refused restart."""
    if not ns.restart:
        return allocate_state(cfg.nx, cfg.ny), 0
    from .io.checkpoint import load_checkpoint
    try:
        state, meta = load_checkpoint(ns.restart)
    except Exception as exc:  # unreadable: start from rest, as the reference does
        print(f"Failed to load checkpoint {ns.restart}: {exc}")
        return allocate_state(cfg.nx, cfg.ny), 0
    saved = meta.get("structured_config_hash")
    if saved and saved != cfg.config_hash and not ns.allow_hash_mismatch:
        print(f"Refusing restart: checkpoint hash {saved} != current {cfg.config_hash} "
              "(use --allow-hash-mismatch to override)")
        return EXIT_HASH_MISMATCH
    try:
        return state, int(meta.get("iteration", 0)) + 1
    except (TypeError, ValueError) as exc:
        print(f"Failed to load checkpoint {ns.restart}: {exc}")
        return allocate_state(cfg.nx, cfg.ny), 0


def _final_checkpoint(ns, cfg, last, quiet: bool) -> None:
    """On Ctrl-C: write the last yielded state (best effort)."""
    if not quiet:
        print("KeyboardInterrupt received – attempting graceful checkpoint...")
    if last is None:
        return
    state, diag = last
    target = ns.checkpoint or "interrupt_final_checkpoint.npz"
    try:
        from .io.checkpoint import save_checkpoint
        save_checkpoint(target, state, diag.get("iteration", 0), diag.get("wall_time", 0.0), cfg)
    except Exception as exc:  # pragma: no cover - disk errors
        if not quiet:
            print(f"Failed to write final checkpoint: {exc}")
        return
    if not quiet:
        print(f"Final checkpoint written: {target}")


def _run(ns: argparse.Namespace, cfg: SimulationConfig) -> int:
    if ns.precision:
        os.environ["PYFLOW_PRECISION"] = ns.precision
    start = _initial_state(ns, cfg)
    if isinstance(start, int):
        return start
    state, first_it = start
    quiet = bool(ns.json_stream)
    if quiet:
        cfg.force_quiet = True
    if cfg.seed is not None:
        import numpy as np
        np.random.seed(int(cfg.seed))
    driver = SimulationDriver(cfg, state, ResidualManager())
    steps = driver.run(max_steps=ns.steps, start_iteration=first_it, progress=ns.progress and not quiet,
                       checkpoint_path=ns.checkpoint, checkpoint_interval=_interval(ns))
    threshold = ns.continuity_threshold
    last = None
    out = sys.stdout
    try:
        for st, residuals, diag in steps:
            last = (st, diag)
            if quiet:
                out.write(_to_line(stream_record(residuals, diag, cfg.seed)) + "\n")
                out.flush()
            if threshold is None:
                continue
            cont = residuals.get("continuity", 1.0)
            if cont < threshold:
                if not quiet:
                    print(f"Stopping early: continuity {cont:.3e} < threshold {threshold:.3e}")
                break
    except KeyboardInterrupt:
        _final_checkpoint(ns, cfg, last, quiet)
        return EXIT_INTERRUPTED
    return EXIT_OK


_to_line = json.JSONEncoder(separators=(",", ":")).encode


def main(argv: list[str] | None = None) -> int:
    ns = build_parser().parse_args(argv)
    try:
        cfg = _config(ns)
    except (ValidationError, ConfigError) as exc:
        print("Configuration invalid:\n" + str(exc))
        return EXIT_BAD_CONFIG
    if ns.command == "validate":
        dump = cfg.model_dump()
        dump["config_hash"] = cfg.config_hash
        print(json.dumps(dump, indent=2, sort_keys=True))
        return EXIT_OK
    if ns.command == "show-config":
        print(json.dumps({"hash": cfg.config_hash, "brief": cfg.brief()}, indent=2))
        return EXIT_OK
    return _run(ns, cfg)


if __name__ == "__main__":  # pragma: no cover
    raise SystemExit(main())
