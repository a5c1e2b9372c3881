"""``aospectrum band CONFIG [--resume]`` on the MI355X.

Mirrors the band half of src/aospectrum/cli.py: TOML parsing
(``read_band_config``, cli.py:123-177), the resumable work directory
(``_prepare_work``, cli.py:275-296, with ``run.json`` identity cli.py:306-327),
per-k checkpoints and the atomic publish of ``bands.csv``/``summary.json``/
``band.png`` (cli.py:299-303, 368-436), and exit status 2 on
``AOSpectrumError`` (cli.py:595-611).  The k loop hands ``[solver] k_batch``
points at a time to the device.  Listing several ``[compute] devices`` stripes
k points over one spawned process per GPU, as the reference does
(cli.py:383-405); every process sees only its own device.

Extra ``[solver]`` keys (defaults keep the reference's behaviour):
``precision``, ``refinement_steps``, ``refinement_passes``, ``k_batch``,
``conjugate_pairs``, ``refinement_policy``, ``residual_target_ev``, ``autotune``,
``resident_limit``.
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
from typing import Any, Sequence

try:
    import tomllib
except ModuleNotFoundError:  # Python 3.10
    import tomli as tomllib

from .band import (
    BandCalculation,
    KPath,
    read_kpath,
    read_point,
    resolve_energy_reference,
    write_band_output,
    write_point,
)
from .bundle import load_bundle
from .ciss import CissSettings
from .data import EnergyInterval, LocalizedOperatorBundle
from .errors import AOSpectrumError, InputError, SolverError

__version__ = "1.0.0rc2+mi355x"


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


def _toml(path: str | Path) -> tuple[Path, dict[str, Any]]:
    source = Path(path).expanduser().resolve()
    try:
        with source.open("rb") as handle:
            return source, tomllib.load(handle)
    except (OSError, tomllib.TOMLDecodeError) as exc:
        raise InputError(f"cannot read TOML configuration {source}: {exc}") from exc


def _section(values: dict[str, Any], name: str) -> dict[str, Any]:
    section = values.get(name, {})
    if not isinstance(section, dict):
        raise InputError(f"[{name}] must be a TOML table")
    return section


def _path(base: Path, value: Any, name: str) -> Path:
    if not isinstance(value, str) or not value.strip():
        raise InputError(f"{name} must be a path")
    p = Path(value).expanduser()
    return (base / p).resolve() if not p.is_absolute() else p.resolve()


def _pair(value: Any, name: str) -> tuple[float, float]:
    if not isinstance(value, list) or len(value) != 2:
        raise InputError(f"{name} must contain two numbers")
    try:
        return float(value[0]), float(value[1])
    except (TypeError, ValueError) as exc:
        raise InputError(f"{name} must contain two numbers") from exc


def read_band_config(path: str | Path) -> BandConfig:
    """(cli.py:123-177) plus the device-execution keys of ``[solver]``."""
    source, values = _toml(path)
    base = source.parent
    inp, band = _section(values, "input"), _section(values, "band")
    compute, solver = _section(values, "compute"), _section(values, "solver")
    resources, output = _section(values, "resources"), _section(values, "output")
    devices = compute.get("devices", [0])
    if (
        not isinstance(devices, list)
        or not devices
        or any(isinstance(v, bool) or not isinstance(v, int) or v < 0 for v in devices)
        or len(set(devices)) != len(devices)
    ):
        raise InputError("compute.devices must contain unique GPU indices")
    interval = EnergyInterval(*_pair(band.get("energy_interval_ev"), "band.energy_interval_ev"))
    mode = str(band.get("energy_reference", "input_zero"))
    ref_value = band.get("energy_reference_ev")
    try:
        ref_value = None if ref_value is None else float(ref_value)
        conj = solver.get("conjugate_pairs", False)
        tune = solver.get("autotune", False)
        if not isinstance(conj, bool) or not isinstance(tune, bool):
            raise TypeError("conjugate_pairs and autotune must be booleans")
        settings = CissSettings(
            integration_points=int(solver.get("integration_points", 32)),
            contour_batch_size=int(solver.get("contour_batch_size", resources.get("contour_batch_size", 1))),
            block_size=int(solver.get("block_size", 32)),
            moment_size=int(solver.get("moment_size", 4)),
            precision=str(solver.get("precision", "fp32")),
            refinement_steps=int(solver.get("refinement_steps", 0)),
            refinement_passes=int(solver.get("refinement_passes", 0)),
            k_batch=int(solver.get("k_batch", 1)),
            conjugate_pairs=conj,
            refinement_policy=str(solver.get("refinement_policy", "fixed")),
            residual_target_ev=float(solver.get("residual_target_ev", 1.0e-4)),
            autotune=tune,
            resident_limit=int(solver.get("resident_limit", 0)),
        )
    except (TypeError, ValueError) as exc:
        raise InputError("invalid Band solver setting") from exc
    return BandConfig(
        source,
        _path(base, inp.get("bundle"), "input.bundle"),
        read_kpath(_path(base, band.get("kpath"), "band.kpath")),
        interval,
        mode,
        ref_value,
        tuple(devices),
        settings,
        _path(base, output.get("directory"), "output.directory"),
    )


def _device_token(index: int, inherited: str | None) -> str:
    """(cli.py:259-267) for HIP_VISIBLE_DEVICES."""
    if inherited:
        visible = [v.strip() for v in inherited.split(",") if v.strip()]
        if index >= len(visible):
            raise InputError(f"GPU index {index} exceeds HIP_VISIBLE_DEVICES={inherited}")
        return visible[index]
    return str(index)


def _require_output(path: Path) -> None:
    if path.exists() and (not path.is_dir() or any(path.iterdir())):
        raise InputError(f"output directory is not empty: {path}")


def _prepare_work(path: Path, spec: dict[str, Any], *, resume: bool) -> None:
    """(cli.py:275-296)"""
    run = path / "run.json"
    if resume:
        try:
            observed = json.loads(run.read_text(encoding="utf-8"))
        except (OSError, json.JSONDecodeError) as exc:
            raise InputError(f"cannot resume work directory {path}") from exc
        if observed != spec:
            raise InputError("resume work belongs to another calculation")
        return
    if path.exists():
        raise InputError(f"work directory already exists: {path}; use --resume")
    path.mkdir(parents=True)
    run.write_text(json.dumps(spec, indent=2, sort_keys=True) + "\n", encoding="utf-8")


def _publish_directory(output: Path, publish: Path, work: Path) -> None:
    if output.exists():
        output.rmdir()
    os.replace(publish, output)
    shutil.rmtree(work)


def _band_specification(config: BandConfig, bundle: LocalizedOperatorBundle) -> dict[str, Any]:
    """run.json identity (cli.py:306-327)."""
    return {
        "kind": "band",
        "bundle": str(config.bundle),
        "basis_identity": bundle.basis.basis_identity,
        "n_orbitals": bundle.n_orbitals,
        "interval_ev": [config.interval.lower_ev, config.interval.upper_ev],
        "reference": [config.reference_mode, config.reference_value_ev],
        "kpoints": config.kpath.fractional_points.tolist(),
        "nodes": config.kpath.node_indices.tolist(),
        "labels": list(config.kpath.node_labels),
        "solver": asdict(config.settings),
    }


def _band_worker(config: BandConfig, reference: float, rank: int, world: int, token: str | None) -> int:
    """Solve this rank's stripe of k points (cli.py:330-365), k_batch at a time."""
    if token is not None:
        os.environ["HIP_VISIBLE_DEVICES"] = token
    bundle = load_bundle(config.bundle)
    calc = BandCalculation(bundle, config.kpath, config.interval, reference, config.settings)
    done = 0
    try:
        pending = []
        for index in range(rank, config.kpath.n_points, world):
            path = config.work / f"k-{index:06d}.npz"
            if read_point(path, index=index, k_fractional=config.kpath.fractional_points[index]) is None:
                pending.append(index)
            else:
                done += 1
        step = max(1, config.settings.k_batch)
        for start in range(0, len(pending), step):
            chunk = pending[start:start + step]
            print(
                f"[gpu {rank}] k points {chunk[0] + 1}..{chunk[-1] + 1}/{config.kpath.n_points}",
                flush=True,
            )
            for point in calc.solve_points(chunk):
                write_point(config.work / f"k-{point.index:06d}.npz", point)
                done += 1
    finally:
        calc.close()
    return done


def run_band(path: str | Path, *, resume: bool = False) -> Path:
    """(cli.py:368-436)"""
    config = read_band_config(path)
    _require_output(config.output)
    bundle = load_bundle(config.bundle)
    reference = resolve_energy_reference(bundle, config.reference_mode, config.reference_value_ev)
    _prepare_work(config.work, _band_specification(config, bundle), resume=resume)
    started = perf_counter()
    inherited = os.environ.get("HIP_VISIBLE_DEVICES")
    if len(config.devices) == 1:
        token = _device_token(config.devices[0], inherited) if config.devices[0] != 0 or inherited else None
        _band_worker(config, reference, 0, 1, token)
    else:
        tokens = [_device_token(i, inherited) for i in config.devices]
        ctx = multiprocessing.get_context("spawn")
        with ProcessPoolExecutor(max_workers=len(tokens), mp_context=ctx) as pool:
            futures = [
                pool.submit(_band_worker, config, reference, rank, len(tokens), token)
                for rank, token in enumerate(tokens)
            ]
            for f in futures:
                f.result()
    points = []
    for index, kpoint in enumerate(config.kpath.fractional_points):
        point = read_point(config.work / f"k-{index:06d}.npz", index=index, k_fractional=kpoint)
        if point is None:
            raise SolverError(f"Band point {index} is missing")
        points.append(point)
    warnings = []
    if bundle.basis.spinor_width != 1:
        warnings.append("spinor input is outside the current non-SOC Band scope; the result may be inaccurate")
    if not all(bundle.structure.periodic):
        warnings.append("non-periodic axes are outside the current Bloch scope; the result may be inaccurate")
    from .band import BandResult

    solve_interval = config.interval.expanded(0.02)
    result = BandResult(
        kpath=config.kpath,
        display_interval_ev=config.interval,
        solve_interval_ev=solve_interval,
        reference_energy_ev=reference,
        cell_angstrom=bundle.structure.cell_angstrom,
        points=tuple(points),
        elapsed_seconds=perf_counter() - started,
        warnings=tuple(warnings),
    )
    publish = config.output.with_name(f".{config.output.name}.publish")
    if publish.exists():
        shutil.rmtree(publish)
    write_band_output(result, publish)
    _publish_directory(config.output, publish, config.work)
    print(config.output)
    return config.output


def _parser() -> argparse.ArgumentParser:
    parser = argparse.ArgumentParser(prog="aospectrum", description="Band path for localized AO Hamiltonians (MI355X).")
    parser.add_argument("--version", action="version", version=__version__)
    commands = parser.add_subparsers(dest="command", required=True)
    band = commands.add_parser("band", help="calculate an energy-window band structure")
    band.add_argument("config", type=Path)
    band.add_argument("--resume", action="store_true")
    return parser


def main(argv: Sequence[str] | None = None) -> int:
    """(cli.py:595-611)"""
    args = _parser().parse_args(argv)
    try:
        if args.command == "band":
            run_band(args.config, resume=args.resume)
    except AOSpectrumError as exc:
        print(f"aospectrum: {exc}", file=sys.stderr)
        return 2
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
