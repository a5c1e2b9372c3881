"""Energy-window Band calculation, resume checkpoints and outputs.

Mirrors src/aospectrum/band.py: ``KPath`` (band.py:26-98), ``read_kpath``
(band.py:101-129), ``resolve_energy_reference`` (band.py:132-147),
``BandPoint``/``BandResult`` (band.py:150-185), ``BandCalculation``
(band.py:188-278), ``write_point``/``read_point`` (band.py:281-344),
``assign_track_ids`` (band.py:347-366) and ``write_band_output``
(band.py:396-499), with the same file formats.  The difference is where the
work runs: ``BandCalculation`` keeps the bundle resident on the GPU and solves
k points in batches through ``aosx_band_solve_v1`` (on-device assembly), unless
a custom ``solver`` is injected, in which case it assembles H(k), S(k) on the
GPU, downloads them and calls ``solver.solve_interval`` exactly like the
reference does.
"""

from __future__ import annotations

import csv
import json
import os
from dataclasses import dataclass, field
from pathlib import Path
from time import perf_counter
from typing import Any, Sequence

import numpy as np
from scipy.optimize import linear_sum_assignment

from .ciss import CissSettings, HipCissSolver, band_solve
from .data import EnergyInterval, LocalizedOperatorBundle
from .errors import InputError, SolverError
from .sparse import BlochPattern

CONTINUITY_GUARD_EV = 0.02
KPATH_SCHEMA = "aospectrum.kpath/v1"


@dataclass(frozen=True, slots=True)
class KPath:
    """Fractional k points with labelled nodes (band.py:26-98)."""

    fractional_points: np.ndarray
    node_indices: np.ndarray
    node_labels: tuple[str, ...]

    def __post_init__(self) -> None:
        pts = np.asarray(self.fractional_points, dtype=np.float64)
        nodes = np.asarray(self.node_indices, dtype=np.int64)
        labels = tuple(str(v) for v in self.node_labels)
        if pts.ndim != 2 or pts.shape[1:] != (3,) or pts.shape[0] < 2:
            raise InputError("k path must contain at least two 3D points")
        if (
            nodes.ndim != 1
            or nodes.size < 2
            or nodes[0] != 0
            or nodes[-1] != pts.shape[0] - 1
            or np.any(np.diff(nodes) <= 0)
            or len(labels) != nodes.size
        ):
            raise InputError("invalid k-path nodes or labels")
        pts.setflags(write=False)
        nodes.setflags(write=False)
        object.__setattr__(self, "fractional_points", pts)
        object.__setattr__(self, "node_indices", nodes)
        object.__setattr__(self, "node_labels", labels)

    @classmethod
    def interpolate(cls, nodes: Sequence[Sequence[float]], labels: Sequence[str], *, points_per_segment: int) -> "KPath":
        coords = np.asarray(nodes, dtype=np.float64)
        count = int(points_per_segment)
        if coords.ndim != 2 or coords.shape[1:] != (3,) or coords.shape[0] < 2 or count < 1:
            raise InputError("invalid interpolated k path")
        parts = []
        idx = [0]
        for i in range(coords.shape[0] - 1):
            seg = np.linspace(coords[i], coords[i + 1], count + 1, dtype=np.float64)
            parts.append(seg if i == 0 else seg[1:])
            idx.append(idx[-1] + count)
        return cls(np.concatenate(parts), np.asarray(idx), tuple(labels))

    @property
    def n_points(self) -> int:
        return int(self.fractional_points.shape[0])

    def distances(self, cell_angstrom: np.ndarray) -> np.ndarray:
        recip = 2.0 * np.pi * np.linalg.inv(cell_angstrom).T
        cart = self.fractional_points @ recip
        out = np.zeros(self.n_points, dtype=np.float64)
        out[1:] = np.cumsum(np.linalg.norm(np.diff(cart, axis=0), axis=1))
        return out


def ygxm_path(points_per_segment: int = 40) -> KPath:
    """Y(0,1/2,0) - G - X(1/2,0,0) - M(1/2,1/2,0), as examples/band/YGXM-121.json."""
    nodes = [[0.0, 0.5, 0.0], [0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.5, 0.5, 0.0]]
    return KPath.interpolate(nodes, ["Y", "G", "X", "M"], points_per_segment=points_per_segment)


def read_kpath(path: str | Path) -> KPath:
    """(band.py:101-129)"""
    source = Path(path)
    try:
        value = json.loads(source.read_text(encoding="utf-8"))
    except (OSError, UnicodeError, json.JSONDecodeError) as exc:
        raise InputError(f"cannot read k path {source}: {exc}") from exc
    if not isinstance(value, dict):
        raise InputError("k path must contain a JSON object")
    if value.get("schema", KPATH_SCHEMA) != KPATH_SCHEMA:
        raise InputError(f"k path schema must be {KPATH_SCHEMA}")
    labels = value.get("labels")
    if not isinstance(labels, list):
        raise InputError("k path labels must be a list")
    if "nodes_fractional" in value:
        return KPath.interpolate(value["nodes_fractional"], labels, points_per_segment=value.get("points_per_segment", 0))
    try:
        return KPath(value["fractional_points"], value["node_indices"], tuple(labels))
    except KeyError as exc:
        raise InputError("k path needs nodes_fractional or fractional_points") from exc


def resolve_energy_reference(bundle: LocalizedOperatorBundle, mode: str, value_ev: float | None = None) -> float:
    """(band.py:132-147)"""
    if mode in {"input_zero", "absolute"}:
        return 0.0
    if mode == "fixed":
        if value_ev is None:
            raise InputError("fixed energy reference requires a value")
        return float(value_ev)
    if mode == "bundle_fermi":
        if bundle.filling is None or bundle.filling.fermi_energy_ev is None:
            raise InputError("bundle has no Fermi energy")
        return bundle.filling.fermi_energy_ev
    raise InputError("energy reference must be input_zero, fixed, or bundle_fermi")


@dataclass(slots=True)
class BandPoint:
    """(band.py:150-166)"""

    index: int
    k_fractional: np.ndarray
    energies_ev: np.ndarray
    maximum_residual_ev: float
    stage_seconds: dict[str, float]
    assembly_seconds: float
    peak_gpu_memory_bytes: int | None
    warnings: tuple[str, ...] = ()

    def __post_init__(self) -> None:
        self.index = int(self.index)
        self.k_fractional = np.asarray(self.k_fractional, dtype=np.float64)
        self.energies_ev = np.asarray(self.energies_ev, dtype=np.float32)
        if self.index < 0 or self.k_fractional.shape != (3,):
            raise InputError("invalid Band point")


@dataclass(slots=True)
class BandResult:
    """(band.py:169-185)"""

    kpath: KPath
    display_interval_ev: EnergyInterval
    solve_interval_ev: EnergyInterval
    reference_energy_ev: float
    cell_angstrom: np.ndarray
    points: tuple[BandPoint, ...]
    elapsed_seconds: float
    warnings: tuple[str, ...] = ()

    @property
    def display_energies_ev(self) -> tuple[np.ndarray, ...]:
        return tuple(p.energies_ev - self.reference_energy_ev for p in self.points)


@dataclass(slots=True)
class BandCalculation:
    """Per-bundle Band driver (band.py:188-278) with batched k solves."""

    bundle: LocalizedOperatorBundle
    kpath: KPath
    display_interval_ev: EnergyInterval
    reference_energy_ev: float
    settings: CissSettings = field(default_factory=CissSettings)
    solver: Any | None = None
    device: int = 0
    solve_interval_ev: EnergyInterval = field(init=False)
    absolute_interval_ev: EnergyInterval = field(init=False)
    pattern: BlochPattern = field(init=False)
    workspace: Any = field(init=False)
    warnings: tuple[str, ...] = field(init=False)
    autotune_result: dict | None = field(init=False)

    def __post_init__(self) -> None:
        self.solve_interval_ev = self.display_interval_ev.expanded(CONTINUITY_GUARD_EV)
        self.absolute_interval_ev = EnergyInterval(
            self.solve_interval_ev.lower_ev + self.reference_energy_ev,
            self.solve_interval_ev.upper_ev + self.reference_energy_ev,
        )
        self.pattern = BlochPattern.from_bundle(self.bundle, device=self.device)
        self.workspace = self.pattern.workspace(complex_values=True)
        self.autotune_result = None
        if self.settings.autotune and self.solver is None:
            self.autotune_result = self.pattern.native.autotune(
                self.settings.factorizations_per_launch, self.settings.precision
            )
        warnings = []
        if self.bundle.basis.spinor_width != 1:
            warnings.append("spinor input is outside the current non-SOC Band scope; the result may be inaccurate")
        if not all(self.bundle.structure.periodic):
            warnings.append("non-periodic axes are outside the current Bloch scope; the result may be inaccurate")
        self.warnings = tuple(warnings)

    def _point(self, index: int, result, assembly_seconds: float) -> BandPoint:
        if not result.complete:
            raise SolverError("CISS search subspace is saturated; increase block_size or moment_size")
        return BandPoint(
            index=index,
            k_fractional=self.kpath.fractional_points[index],
            energies_ev=result.energies_ev,
            maximum_residual_ev=result.maximum_residual_ev,
            stage_seconds=result.stage_seconds,
            assembly_seconds=assembly_seconds,
            peak_gpu_memory_bytes=result.peak_gpu_memory_bytes,
            warnings=result.warnings,
        )

    def solve_points(self, indices: Sequence[int]) -> list[BandPoint]:
        """Solve several k points; with the built-in solver they are factored
        in groups of ``settings.k_batch`` on the device."""
        idx = [int(i) for i in indices]
        if any(not 0 <= i < self.kpath.n_points for i in idx):
            raise InputError("k-point index is out of range")
        if self.solver is not None:
            return [self.solve_point(i) for i in idx]
        k = self.kpath.fractional_points[idx]
        results, _ = band_solve(self.pattern.native, self.settings, self.absolute_interval_ev, kpoints=k)
        return [
            self._point(i, r, r.stage_seconds.get("native_assembly", 0.0)) for i, r in zip(idx, results)
        ]

    def solve_point(self, index: int) -> BandPoint:
        """(band.py:227-253)"""
        if not 0 <= int(index) < self.kpath.n_points:
            raise InputError("k-point index is out of range")
        index = int(index)
        if self.solver is None:
            return self.solve_points([index])[0]
        started = perf_counter()
        matrices = self.workspace.update(self.bundle, self.kpath.fractional_points[index])
        assembly = perf_counter() - started
        result = self.solver.solve_interval(matrices, self.absolute_interval_ev)
        return self._point(index, result, assembly)

    def build_result(self, points: Sequence[BandPoint], *, elapsed_seconds: float) -> BandResult:
        ordered = tuple(sorted(points, key=lambda p: p.index))
        if [p.index for p in ordered] != list(range(self.kpath.n_points)):
            raise SolverError("Band result does not contain every k point")
        return BandResult(
            kpath=self.kpath,
            display_interval_ev=self.display_interval_ev,
            solve_interval_ev=self.solve_interval_ev,
            reference_energy_ev=self.reference_energy_ev,
            cell_angstrom=self.bundle.structure.cell_angstrom,
            points=ordered,
            elapsed_seconds=float(elapsed_seconds),
            warnings=self.warnings,
        )

    def close(self) -> None:
        close = getattr(self.solver, "close", None)
        if callable(close):
            close()
        self.pattern.native.close()


def write_point(path: str | Path, point: BandPoint) -> None:
    """Atomic per-k checkpoint with the reference's npz keys (band.py:281-309)."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_name(f".{path.name}.tmp")
    with tmp.open("wb") as handle:
        np.savez(
            handle,
            index=np.asarray(point.index, dtype=np.int64),
            k_fractional=point.k_fractional,
            energies_ev=point.energies_ev,
            maximum_residual_ev=np.asarray(point.maximum_residual_ev, dtype=np.float64),
            stage_names=np.asarray(tuple(point.stage_seconds), dtype="U64"),
            stage_values=np.asarray(tuple(point.stage_seconds.values()), dtype=np.float64),
            assembly_seconds=np.asarray(point.assembly_seconds, dtype=np.float64),
            peak_gpu_memory_bytes=np.asarray(
                -1 if point.peak_gpu_memory_bytes is None else point.peak_gpu_memory_bytes, dtype=np.int64
            ),
            warnings=np.asarray(point.warnings, dtype="U256"),
        )
    os.replace(tmp, path)


def read_point(path: str | Path, *, index: int, k_fractional: np.ndarray) -> BandPoint | None:
    """(band.py:312-344)"""
    path = Path(path)
    if not path.is_file():
        return None
    try:
        with np.load(path, allow_pickle=False) as v:
            if int(v["index"]) != index or not np.array_equal(np.asarray(v["k_fractional"]), k_fractional):
                raise InputError(f"Band checkpoint does not match point {index}")
            peak = int(v["peak_gpu_memory_bytes"])
            return BandPoint(
                index=index,
                k_fractional=np.asarray(v["k_fractional"]),
                energies_ev=v["energies_ev"],
                maximum_residual_ev=float(v["maximum_residual_ev"]),
                stage_seconds=dict(zip(v["stage_names"].tolist(), v["stage_values"].tolist())),
                assembly_seconds=float(v["assembly_seconds"]),
                peak_gpu_memory_bytes=None if peak < 0 else peak,
                warnings=tuple(v["warnings"].tolist()),
            )
    except (OSError, ValueError, KeyError) as exc:
        raise InputError(f"cannot read Band checkpoint {path}: {exc}") from exc


def assign_track_ids(energies: Sequence[np.ndarray]) -> tuple[np.ndarray, ...]:
    """Band continuity by minimum-cost matching of |dE| (band.py:347-366)."""
    first = np.arange(energies[0].size, dtype=np.int64)
    tracks = [first]
    next_id = int(first.size)
    for prev, cur in zip(energies, energies[1:]):
        prev_ids = tracks[-1]
        cur_ids = np.full(cur.size, -1, dtype=np.int64)
        if prev.size and cur.size:
            rows, cols = linear_sum_assignment(np.abs(prev[:, None] - cur[None, :]))
            cur_ids[cols] = prev_ids[rows]
        free = np.flatnonzero(cur_ids < 0)
        cur_ids[free] = np.arange(next_id, next_id + free.size)
        next_id += int(free.size)
        tracks.append(cur_ids)
    return tuple(tracks)


def _segments(x: np.ndarray, energies, tracks) -> np.ndarray:
    segs = []
    for i in range(x.size - 1):
        right = {int(t): j for j, t in enumerate(tracks[i + 1])}
        for j, t in enumerate(tracks[i]):
            m = right.get(int(t))
            if m is not None:
                segs.append(np.asarray([[x[i], energies[i][j]], [x[i + 1], energies[i + 1][m]]]))
    return np.stack(segs) if segs else np.empty((0, 2, 2), dtype=np.float64)


def write_band_output(result: BandResult, root: str | Path, *, backend: str = "ciss-hip") -> Path:
    """bands.csv, summary.json and band.png (band.py:396-499)."""
    root = Path(root)
    root.mkdir(parents=True, exist_ok=True)
    dist = result.kpath.distances(result.cell_angstrom)
    energies = result.display_energies_ev
    tracks = assign_track_ids(energies)
    with (root / "bands.csv").open("w", encoding="utf-8", newline="") as handle:
        w = csv.writer(handle)
        w.writerow(("k_index", "k_distance_inv_angstrom", "track_id", "energy_ev"))
        for point, d, ids, vals in zip(result.points, dist, tracks, energies):
            for t, e in zip(ids, vals):
                w.writerow((point.index, f"{d:.12e}", int(t), f"{float(e):.12e}"))
    stages: dict[str, float] = {}
    warnings = list(result.warnings)
    for p in result.points:
        for name, sec in p.stage_seconds.items():
            stages[name] = stages.get(name, 0.0) + float(sec)
        warnings.extend(p.warnings)
    peaks = [p.peak_gpu_memory_bytes for p in result.points if p.peak_gpu_memory_bytes is not None]
    summary = {
        "backend": backend,
        "scalar_dtype": "complex64",
        "kpoint_count": result.kpath.n_points,
        "energy_interval_ev": [result.display_interval_ev.lower_ev, result.display_interval_ev.upper_ev],
        "reference_energy_ev": result.reference_energy_ev,
        "maximum_eigenpair_residual_ev": max((p.maximum_residual_ev for p in result.points), default=0.0),
        "elapsed_seconds": result.elapsed_seconds,
        "assembly_seconds": sum(p.assembly_seconds for p in result.points),
        "solver_stage_seconds": stages,
        "peak_gpu_memory_bytes": max(peaks) if peaks else None,
        "warnings": list(dict.fromkeys(warnings)),
    }
    (root / "summary.json").write_text(json.dumps(summary, indent=2, sort_keys=True) + "\n", encoding="utf-8")
    try:
        import matplotlib

        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
        from matplotlib.collections import LineCollection
    except ImportError:  # plotting is optional output
        return root
    segs = _segments(dist, energies, tracks)
    fig, ax = plt.subplots(figsize=(7.2, 5.2), constrained_layout=True)
    if segs.size:
        ax.add_collection(LineCollection(segs, colors="#17213a", linewidths=1.05, antialiased=True))
    ax.set_xlim(float(dist[0]), float(dist[-1]))
    ax.set_ylim(result.display_interval_ev.lower_ev, result.display_interval_ev.upper_ev)
    node_x = dist[result.kpath.node_indices]
    for pos in node_x[1:-1]:
        ax.axvline(pos, color="#9aa1ad", linewidth=0.7, zorder=0)
    ax.axhline(0.0, color="#b73b3b", linewidth=0.8, linestyle="--")
    ax.set_xticks(node_x, result.kpath.node_labels)
    ax.set_xlabel("K path")
    ax.set_ylabel("Energy relative to reference (eV)")
    ax.tick_params(direction="in")
    fig.savefig(root / "band.png", dpi=240)
    plt.close(fig)
    return root
