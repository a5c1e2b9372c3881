"""``aospectrum band CONFIG [--resume]`` on the MI355X.

The command-line contract is the reference's (src/aospectrum/cli.py): the
TOML keys of the band configuration (cli.py:123-177), the resumable work
directory next to the output with its ``run.json`` identity (cli.py:275-327),
per-k checkpoints, the atomic publish of ``bands.csv``/``summary.json``/
``band.png`` (cli.py:299-303, 368-436), device striping over ``[compute]
devices`` (cli.py:383-405) and exit status 2 on an ``AOSpectrumError``
(cli.py:595-611).  Here the configuration is read through a declarative key
schema, the work directory is one object that prepares, resumes and
publishes, and each device's worker hands ``[solver] k_batch`` points at a
time to the GPU.

Extra ``[solver]`` keys (defaults keep the reference's behaviour):
``precision``, ``refinement_steps``, ``refinement_passes``, ``k_batch``,
``conjugate_pairs``, ``refinement_policy``, ``residual_target_ev``,
``autotune``, ``resident_limit``, ``far_shift_threshold``.
"""

from __future__ import annotations

import argparse
import json
import multiprocessing
import os
import shutil
import sys
from concurrent.futures import ProcessPoolExecutor
from dataclasses import asdict, dataclass
from pathlib import Path
from time import perf_counter
from typing import Any, Callable, Sequence

try:
    import tomllib
except ModuleNotFoundError:  # Python 3.10
    import tomli as tomllib

from . import __version__, formats
from .band import (
    BandCalculation,
    BandResult,
    KPath,
    read_kpath,
    read_point,
    resolve_energy_reference,
    scope_warnings,
    write_band_output,
    write_point,
)
from .bundle import load_bundle
from .ciss import CissSettings
from .data import EnergyInterval, LocalizedOperatorBundle
from .errors import AOSpectrumError, InputError, SolverError


@dataclass(frozen=True, slots=True)
class BandConfig:
    source: Path
    bundle: Path
    kpath: KPath
    interval: EnergyInterval
    reference_mode: str
    reference_value_ev: float | None
    devices: tuple[int, ...]
    settings: CissSettings
    output: Path

    @property
    def work(self) -> Path:
        return self.output.with_name(f".{self.output.name}.work")


# --------------------------------------------------------------- TOML keys
class _Doc:
    """A parsed TOML document with its tables checked on first use."""

    def __init__(self, path: str | Path) -> None:
        self.source = Path(path).expanduser().resolve()
        try:
            with self.source.open("rb") as handle:
                self.values = tomllib.load(handle)
        except (OSError, tomllib.TOMLDecodeError) as exc:
            raise InputError(f"cannot read TOML configuration {self.source}: {exc}") from exc

    def table(self, name: str) -> dict[str, Any]:
        table = self.values.get(name, {})
        if not isinstance(table, dict):
            raise InputError(f"[{name}] must be a TOML table")
        return table

    def get(self, dotted: str, default: Any = None) -> Any:
        table, key = dotted.split(".")
        return self.table(table).get(key, default)

    def path(self, dotted: str) -> Path:
        value = self.get(dotted)
        if not isinstance(value, str) or not value.strip():
            raise InputError(f"{dotted} must be a path")
        p = Path(value).expanduser()
        return p.resolve() if p.is_absolute() else (self.source.parent / p).resolve()

    def pair(self, dotted: str) -> tuple[float, float]:
        value = self.get(dotted)
        try:
            if not isinstance(value, list) or len(value) != 2:
                raise TypeError
            return float(value[0]), float(value[1])
        except (TypeError, ValueError) as exc:
            raise InputError(f"{dotted} must contain two numbers") from exc


def _boolean(value: Any) -> bool:
    if not isinstance(value, bool):
        raise TypeError("expected a boolean")
    return value


# CissSettings field <- [solver] key (type, default).  contour_batch_size may
# also come from [resources], the reference's older location (cli.py:155-160).
_SOLVER_KEYS: tuple[tuple[str, Callable[[Any], Any], Any], ...] = (
    ("integration_points", int, 32),
    ("contour_batch_size", int, None),
    ("block_size", int, 32),
    ("moment_size", int, 4),
    ("precision", str, "fp32"),
    ("refinement_steps", int, 0),
    ("refinement_passes", int, 0),
    ("k_batch", int, 1),
    ("conjugate_pairs", _boolean, False),
    ("refinement_policy", str, "fixed"),
    ("residual_target_ev", float, 1.0e-4),
    ("autotune", _boolean, False),
    ("resident_limit", int, 0),
    ("far_shift_threshold", float, 0.9),
)


def _solver_settings(doc: _Doc) -> CissSettings:
    solver = doc.table("solver")
    fallback_batch = doc.table("resources").get("contour_batch_size", 1)
    values = {}
    for name, kind, default in _SOLVER_KEYS:
        raw = solver.get(name, fallback_batch if name == "contour_batch_size" else default)
        values[name] = kind(raw)
    return CissSettings(**values)


def _devices(doc: _Doc) -> tuple[int, ...]:
    devices = doc.get("compute.devices", [0])
    valid = (
        isinstance(devices, list)
        and bool(devices)
        and all(isinstance(d, int) and not isinstance(d, bool) and d >= 0 for d in devices)
        and len(set(devices)) == len(devices)
    )
    if not valid:
        raise InputError("compute.devices must contain unique GPU indices")
    return tuple(devices)


def read_band_config(path: str | Path) -> BandConfig:
    """The band configuration (cli.py:123-177) with the device-execution keys."""
    doc = _Doc(path)
    devices = _devices(doc)
    interval = EnergyInterval(*doc.pair("band.energy_interval_ev"))
    try:
        ref_value = doc.get("band.energy_reference_ev")
        ref_value = None if ref_value is None else float(ref_value)
        settings = _solver_settings(doc)
    except (TypeError, ValueError) as exc:
        raise InputError("invalid Band solver setting") from exc
    return BandConfig(
        source=doc.source,
        bundle=doc.path("input.bundle"),
        kpath=read_kpath(doc.path("band.kpath")),
        interval=interval,
        reference_mode=str(doc.get("band.energy_reference", "input_zero")),
        reference_value_ev=ref_value,
        devices=devices,
        settings=settings,
        output=doc.path("output.directory"),
    )


# ------------------------------------------------------------ work directory
class WorkDirectory:
    """The hidden ``.<output>.work`` directory: created with the run's
    identity, reopened only for the same identity on --resume, and replaced
    by the published output when every k point is done."""

    def __init__(self, output: Path) -> None:
        self.output = output
        self.root = output.with_name(f".{output.name}.work")
        self.staging = output.with_name(f".{output.name}.publish")

    def check_output_free(self) -> None:
        out = self.output
        if out.exists() and (not out.is_dir() or any(out.iterdir())):
            raise InputError(f"output directory is not empty: {out}")

    def open(self, identity: dict[str, Any], *, resume: bool) -> None:
        run = self.root / formats.RUN_DOCUMENT
        if not resume:
            if self.root.exists():
                raise InputError(f"work directory already exists: {self.root}; use --resume")
            self.root.mkdir(parents=True)
            run.write_text(formats.json_text(identity), encoding="utf-8")
            return
        try:
            recorded = json.loads(run.read_text(encoding="utf-8"))
        except (OSError, json.JSONDecodeError) as exc:
            raise InputError(f"cannot resume work directory {self.root}") from exc
        if recorded != identity:
            raise InputError("resume work belongs to another calculation")

    def checkpoint(self, index: int) -> Path:
        return self.root / formats.CHECKPOINT_PATTERN.format(index=index)

    def publish(self, write: Callable[[Path], Any]) -> None:
        if self.staging.exists():
            shutil.rmtree(self.staging)
        write(self.staging)
        if self.output.exists():
            self.output.rmdir()  # empty, checked before the run
        os.replace(self.staging, self.output)
        shutil.rmtree(self.root)


def _prepare_work(path: Path, spec: dict[str, Any], *, resume: bool) -> None:
    """Create or reopen the work directory ``path`` (the reference's
    _prepare_work contract, cli.py:275-296)."""
    work = WorkDirectory(path.with_name(path.name.removeprefix(".").removesuffix(".work")))
    work.root = path
    work.open(spec, resume=resume)


def _band_specification(config: BandConfig, bundle: LocalizedOperatorBundle) -> dict[str, Any]:
    """The identity a resumed run must match (run.json, cli.py:306-327)."""
    path = config.kpath
    return dict(
        kind="band",
        bundle=str(config.bundle),
        basis_identity=bundle.basis.basis_identity,
        n_orbitals=bundle.n_orbitals,
        interval_ev=[config.interval.lower_ev, config.interval.upper_ev],
        reference=[config.reference_mode, config.reference_value_ev],
        kpoints=path.fractional_points.tolist(),
        nodes=path.node_indices.tolist(),
        labels=list(path.node_labels),
        solver=asdict(config.settings),
    )


# --------------------------------------------------------------- execution
def _visible_token(index: int, inherited: str | None) -> str:
    """The HIP_VISIBLE_DEVICES entry of configured device ``index``: a
    position in an inherited list, else the ordinal itself (cli.py:259-267)."""
    if not inherited:
        return str(index)
    visible = [v.strip() for v in inherited.split(",") if v.strip()]
    if index >= len(visible):
        raise InputError(f"GPU index {index} exceeds HIP_VISIBLE_DEVICES={inherited}")
    return visible[index]


def _band_worker(config: BandConfig, reference: float, rank: int, world: int, token: str | None) -> int:
    """Solve the k points i = rank (mod world) that have no checkpoint yet,
    ``k_batch`` at a time on this worker's device; returns the points this
    stripe holds afterwards."""
    if token is not None:
        os.environ["HIP_VISIBLE_DEVICES"] = token
    work = WorkDirectory(config.output)
    stripe = range(rank, config.kpath.n_points, world)
    todo = [i for i in stripe
            if read_point(work.checkpoint(i), index=i, k_fractional=config.kpath.fractional_points[i]) is None]
    calc = BandCalculation(load_bundle(config.bundle), config.kpath, config.interval, reference, config.settings)
    try:
        step = max(1, config.settings.k_batch)
        for batch in (todo[j:j + step] for j in range(0, len(todo), step)):
            print(f"[gpu {rank}] k points {batch[0] + 1}..{batch[-1] + 1}/{config.kpath.n_points}", flush=True)
            for point in calc.solve_points(batch):
                write_point(work.checkpoint(point.index), point)
    finally:
        calc.close()
    return len(stripe)


def _run_workers(config: BandConfig, reference: float) -> None:
    inherited = os.environ.get("HIP_VISIBLE_DEVICES")
    if len(config.devices) == 1:
        only = config.devices[0]
        token = _visible_token(only, inherited) if (only != 0 or inherited) else None
        _band_worker(config, reference, 0, 1, token)
        return
    tokens = [_visible_token(i, inherited) for i in config.devices]
    with ProcessPoolExecutor(max_workers=len(tokens), mp_context=multiprocessing.get_context("spawn")) as pool:
        jobs = [pool.submit(_band_worker, config, reference, r, len(tokens), t) for r, t in enumerate(tokens)]
        for job in jobs:
            job.result()


def _collect(config: BandConfig, work: WorkDirectory) -> list:
    points = []
    for index, k in enumerate(config.kpath.fractional_points):
        point = read_point(work.checkpoint(index), index=index, k_fractional=k)
        if point is None:
            raise SolverError(f"Band point {index} is missing")
        points.append(point)
    return points


def run_band(path: str | Path, *, resume: bool = False) -> Path:
    """Solve every k point of the configured path, resumably, and publish."""
    config = read_band_config(path)
    work = WorkDirectory(config.output)
    work.check_output_free()
    bundle = load_bundle(config.bundle)
    reference = resolve_energy_reference(bundle, config.reference_mode, config.reference_value_ev)
    work.open(_band_specification(config, bundle), resume=resume)
    started = perf_counter()
    _run_workers(config, reference)
    result = BandResult(
        kpath=config.kpath,
        display_interval_ev=config.interval,
        solve_interval_ev=config.interval.expanded(0.02),
        reference_energy_ev=reference,
        cell_angstrom=bundle.structure.cell_angstrom,
        points=tuple(_collect(config, work)),
        elapsed_seconds=perf_counter() - started,
        warnings=scope_warnings(bundle),
    )
    work.publish(lambda staging: write_band_output(result, staging))
    print(config.output)
    return config.output


# --------------------------------------------------------------------- CLI
_COMMANDS: dict[str, Callable[[argparse.Namespace], Any]] = {
    "band": lambda args: run_band(args.config, resume=args.resume),
}


def _parser() -> argparse.ArgumentParser:
    parser = argparse.ArgumentParser(prog="aospectrum", description="Band path for localized AO Hamiltonians (MI355X).")
    parser.add_argument("--version", action="version", version=__version__)
    sub = parser.add_subparsers(dest="command", required=True)
    band = sub.add_parser("band", help="calculate an energy-window band structure")
    band.add_argument("config", type=Path)
    band.add_argument("--resume", action="store_true")
    return parser


def main(argv: Sequence[str] | None = None) -> int:
    args = _parser().parse_args(argv)
    try:
        _COMMANDS[args.command](args)
    except AOSpectrumError as exc:
        print(f"aospectrum: {exc}", file=sys.stderr)
        return 2
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
