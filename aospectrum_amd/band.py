"""Energy-window Band calculation on the MI355X: k paths, the per-bundle
driver, resumable per-k checkpoints, band continuity and the published
outputs.

Interfaces kept from src/aospectrum/band.py so that drivers and outputs are
interchangeable: ``KPath`` (band.py:26-98), ``read_kpath`` (band.py:101-129),
``resolve_energy_reference`` (band.py:132-147), ``BandPoint``/``BandResult``
(band.py:150-185), ``BandCalculation`` (band.py:188-278), ``write_point``/
``read_point`` (band.py:281-344), ``assign_track_ids`` (band.py:347-366) and
``write_band_output`` (band.py:396-499).  The file layouts are declared in
``formats.py``.

The calculation itself differs from the reference: ``BandCalculation`` keeps
the bundle resident on the GPU and solves k points in batches through
``aosx_band_solve_v1`` (on-device assembly).  Only when a custom ``solver`` is
injected does it assemble H(k), S(k), hand them over as host CSR matrices and
call ``solver.solve_interval`` per k point, as the reference always does.
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path
from time import perf_counter
from typing import Any, Callable, Sequence

import numpy as np
from scipy.optimize import linear_sum_assignment

from . import formats
from .ciss import CissSettings, band_solve
from .data import EnergyInterval, LocalizedOperatorBundle
from .errors import InputError, SolverError
from .formats import KPATH_SCHEMA
from .sparse import BlochPattern

CONTINUITY_GUARD_EV = 0.02

__all__ = [
    "CONTINUITY_GUARD_EV",
    "KPATH_SCHEMA",
    "BandCalculation",
    "BandPoint",
    "BandResult",
    "KPath",
    "assign_track_ids",
    "read_kpath",
    "read_point",
    "resolve_energy_reference",
    "write_band_output",
    "write_point",
    "ygxm_path",
]


# ------------------------------------------------------------------ k paths
def _kpath_problems(points: np.ndarray, nodes: np.ndarray, labels: tuple[str, ...]) -> str | None:
    """The first violated rule of a k path, or None."""
    if points.ndim != 2 or points.shape[1:] != (3,) or points.shape[0] < 2:
        return "k path must contain at least two 3D points"
    node_rules = (
        nodes.ndim == 1 and nodes.size >= 2,
        nodes.size >= 2 and nodes[0] == 0 and nodes[-1] == points.shape[0] - 1,
        nodes.size < 2 or bool(np.all(np.diff(nodes) > 0)),
        len(labels) == nodes.size,
    )
    return None if all(node_rules) else "invalid k-path nodes or labels"


@dataclass(frozen=True, slots=True)
class KPath:
    """Fractional k points with labelled high-symmetry nodes."""

    fractional_points: np.ndarray
    node_indices: np.ndarray
    node_labels: tuple[str, ...]

    def __post_init__(self) -> None:
        points = np.array(self.fractional_points, dtype=np.float64)
        nodes = np.array(self.node_indices, dtype=np.int64)
        labels = tuple(map(str, self.node_labels))
        problem = _kpath_problems(points, nodes, labels)
        if problem:
            raise InputError(problem)
        for array in (points, nodes):
            array.flags.writeable = False
        for name, value in (("fractional_points", points), ("node_indices", nodes), ("node_labels", labels)):
            object.__setattr__(self, name, value)

    @classmethod
    def interpolate(cls, nodes: Sequence[Sequence[float]], labels: Sequence[str], *, points_per_segment: int) -> "KPath":
        """Uniform segments between consecutive nodes, each node listed once."""
        corners = np.asarray(nodes, dtype=np.float64)
        per = int(points_per_segment)
        if corners.ndim != 2 or corners.shape[1:] != (3,) or corners.shape[0] < 2 or per < 1:
            raise InputError("invalid interpolated k path")
        legs = [np.linspace(a, b, per + 1, dtype=np.float64)[1:] for a, b in zip(corners[:-1], corners[1:])]
        points = np.vstack([corners[:1], *legs])
        return cls(points, np.arange(corners.shape[0]) * per, tuple(labels))

    @property
    def n_points(self) -> int:
        return int(self.fractional_points.shape[0])

    def distances(self, cell_angstrom: np.ndarray) -> np.ndarray:
        """Cumulative Cartesian path length (1/Angstrom, 2 pi convention)."""
        reciprocal = 2.0 * np.pi * np.linalg.inv(cell_angstrom).T
        steps = np.linalg.norm(np.diff(self.fractional_points @ reciprocal, axis=0), axis=1)
        return np.concatenate(([0.0], np.cumsum(steps)))


def ygxm_path(points_per_segment: int = 40) -> KPath:
    """Y(0,1/2,0) - G - X(1/2,0,0) - M(1/2,1/2,0), the path of
    examples/band/YGXM-121.json."""
    corners = [[0.0, 0.5, 0.0], [0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.5, 0.5, 0.0]]
    return KPath.interpolate(corners, ["Y", "G", "X", "M"], points_per_segment=points_per_segment)


def read_kpath(path: str | Path) -> KPath:
    """A k-path JSON document: either ``nodes_fractional`` with
    ``points_per_segment`` (interpolated) or explicit ``fractional_points``
    with ``node_indices``; ``labels`` name the nodes."""
    source = Path(path)
    try:
        doc = json.loads(source.read_text(encoding="utf-8"))
    except (OSError, UnicodeError, json.JSONDecodeError) as exc:
        raise InputError(f"cannot read k path {source}: {exc}") from exc
    checks = (
        (lambda d: isinstance(d, dict), "k path must contain a JSON object"),
        (lambda d: d.get("schema", KPATH_SCHEMA) == KPATH_SCHEMA, f"k path schema must be {KPATH_SCHEMA}"),
        (lambda d: isinstance(d.get("labels"), list), "k path labels must be a list"),
    )
    for ok, message in checks:
        if not ok(doc):
            raise InputError(message)
    if "nodes_fractional" in doc:
        return KPath.interpolate(doc["nodes_fractional"], doc["labels"],
                                 points_per_segment=doc.get("points_per_segment", 0))
    if "fractional_points" not in doc or "node_indices" not in doc:
        raise InputError("k path needs nodes_fractional or fractional_points")
    return KPath(doc["fractional_points"], doc["node_indices"], tuple(doc["labels"]))


def _fermi_reference(bundle: LocalizedOperatorBundle, _value: float | None) -> float:
    fermi = None if bundle.filling is None else bundle.filling.fermi_energy_ev
    if fermi is None:
        raise InputError("bundle has no Fermi energy")
    return fermi


def _fixed_reference(_bundle: LocalizedOperatorBundle, value: float | None) -> float:
    if value is None:
        raise InputError("fixed energy reference requires a value")
    return float(value)


_REFERENCE_MODES: dict[str, Callable[[LocalizedOperatorBundle, float | None], float]] = {
    "input_zero": lambda _b, _v: 0.0,
    "absolute": lambda _b, _v: 0.0,
    "fixed": _fixed_reference,
    "bundle_fermi": _fermi_reference,
}


def resolve_energy_reference(bundle: LocalizedOperatorBundle, mode: str, value_ev: float | None = None) -> float:
    """The energy that bands.csv and band.png are shown relative to."""
    resolver = _REFERENCE_MODES.get(mode)
    if resolver is None:
        raise InputError("energy reference must be input_zero, fixed, or bundle_fermi")
    return resolver(bundle, value_ev)


# ------------------------------------------------------------------ results
@dataclass(slots=True)
class BandPoint:
    """One solved k point (the fields of band.py:150-166)."""

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
        if not (self.index >= 0 and self.k_fractional.shape == (3,)):
            raise InputError("invalid Band point")


@dataclass(slots=True)
class BandResult:
    """All k points of a path with their window and energy reference."""

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
        shift = self.reference_energy_ev
        return tuple(point.energies_ev - shift for point in self.points)


def scope_warnings(bundle: LocalizedOperatorBundle) -> tuple[str, ...]:
    """Inputs outside the non-SOC, fully periodic Bloch scope of the path."""
    notes = {
        "spinor input is outside the current non-SOC Band scope; the result may be inaccurate":
            bundle.basis.spinor_width != 1,
        "non-periodic axes are outside the current Bloch scope; the result may be inaccurate":
            not all(bundle.structure.periodic),
    }
    return tuple(text for text, applies in notes.items() if applies)


@dataclass(slots=True)
class BandCalculation:
    """Per-bundle Band driver: one resident bundle (symbolic analysis and
    block values on the GPU), k points solved ``settings.k_batch`` at a time."""

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
        # the solve interval widens the display window by the continuity
        # guard, and solves are posed in absolute energies
        self.solve_interval_ev = self.display_interval_ev.expanded(CONTINUITY_GUARD_EV)
        ref = self.reference_energy_ev
        self.absolute_interval_ev = EnergyInterval(self.solve_interval_ev.lower_ev + ref,
                                                   self.solve_interval_ev.upper_ev + ref)
        self.pattern = BlochPattern.from_bundle(self.bundle, device=self.device)
        self.workspace = self.pattern.workspace(complex_values=True)
        tune = self.settings.autotune and self.solver is None
        self.autotune_result = (
            self.pattern.native.autotune(self.settings.factorizations_per_launch, self.settings.precision)
            if tune else None
        )
        self.warnings = scope_warnings(self.bundle)

    def _point(self, index: int, result, assembly_seconds: float) -> BandPoint:
        # a saturated search space may have missed states (band.py:239-243)
        if not result.complete:
            raise SolverError("CISS search subspace is saturated; increase block_size or moment_size")
        return BandPoint(index, self.kpath.fractional_points[index], result.energies_ev,
                         result.maximum_residual_ev, result.stage_seconds, assembly_seconds,
                         result.peak_gpu_memory_bytes, result.warnings)

    def _checked(self, indices: Sequence[int]) -> list[int]:
        idx = [int(i) for i in indices]
        if not all(0 <= i < self.kpath.n_points for i in idx):
            raise InputError("k-point index is out of range")
        return idx

    def solve_points(self, indices: Sequence[int]) -> list[BandPoint]:
        """Solve several k points; with the built-in solver they are factored
        in groups of ``settings.k_batch`` on the device."""
        idx = self._checked(indices)
        if self.solver is not None:
            return [self._plugin_point(i) for i in idx]
        kpoints = self.kpath.fractional_points[idx]
        results, _ = band_solve(self.pattern.native, self.settings, self.absolute_interval_ev, kpoints=kpoints)
        return [self._point(i, r, r.stage_seconds.get("native_assembly", 0.0)) for i, r in zip(idx, results)]

    def _plugin_point(self, index: int) -> BandPoint:
        """The reference's per-k path through an injected interval solver."""
        started = perf_counter()
        matrices = self.workspace.update(self.bundle, self.kpath.fractional_points[index])
        assembly = perf_counter() - started
        return self._point(index, self.solver.solve_interval(matrices, self.absolute_interval_ev), assembly)

    def solve_point(self, index: int) -> BandPoint:
        (point,) = self.solve_points([index])
        return point

    def build_result(self, points: Sequence[BandPoint], *, elapsed_seconds: float) -> BandResult:
        by_index = {p.index: p for p in points}
        if sorted(by_index) != list(range(self.kpath.n_points)) or len(by_index) != len(points):
            raise SolverError("Band result does not contain every k point")
        return BandResult(self.kpath, self.display_interval_ev, self.solve_interval_ev, self.reference_energy_ev,
                          self.bundle.structure.cell_angstrom, tuple(by_index[i] for i in sorted(by_index)),
                          float(elapsed_seconds), self.warnings)

    def close(self) -> None:
        for closer in (getattr(self.solver, "close", None), self.pattern.native.close):
            if callable(closer):
                closer()


# -------------------------------------------------------------- checkpoints
def write_point(path: str | Path, point: BandPoint) -> None:
    """Per-k checkpoint, written atomically (fields in formats.CHECKPOINT_FIELDS)."""
    with formats.atomic_output(Path(path)) as handle:
        np.savez(handle, **formats.encode_checkpoint(point))


def read_point(path: str | Path, *, index: int, k_fractional: np.ndarray) -> BandPoint | None:
    """The checkpoint of point ``index``, None when absent; one written for
    another point or path is an input error."""
    path = Path(path)
    if not path.is_file():
        return None
    try:
        with np.load(path, allow_pickle=False) as arrays:
            fields = formats.decode_checkpoint(arrays)
    except (OSError, ValueError, KeyError) as exc:
        raise InputError(f"cannot read Band checkpoint {path}: {exc}") from exc
    if fields["index"] != index or not np.array_equal(fields["k_fractional"], k_fractional):
        raise InputError(f"Band checkpoint does not match point {index}")
    return BandPoint(
        index=index,
        k_fractional=fields["k_fractional"],
        energies_ev=fields["energies_ev"],
        maximum_residual_ev=fields["maximum_residual_ev"],
        stage_seconds=dict(zip(fields["stage_names"], fields["stage_values"])),
        assembly_seconds=fields["assembly_seconds"],
        peak_gpu_memory_bytes=fields["peak_gpu_memory_bytes"],
        warnings=fields["warnings"],
    )


# --------------------------------------------------------- band continuity
class _TrackBook:
    """Hands out band-track identifiers along the path: the states of each k
    point continue the tracks of the previous one by a minimum-cost matching
    of |dE| (scipy's linear_sum_assignment); unmatched states open new tracks
    numbered in order of appearance."""

    def __init__(self) -> None:
        self.issued = 0
        self.tracks: list[np.ndarray] = []

    def _fresh(self, count: int) -> np.ndarray:
        ids = np.arange(self.issued, self.issued + count, dtype=np.int64)
        self.issued += count
        return ids

    def extend(self, previous: np.ndarray | None, current: np.ndarray) -> None:
        ids = np.full(current.size, -1, dtype=np.int64)
        if previous is not None and previous.size and current.size:
            cost = np.abs(np.subtract.outer(previous, current))
            prev_pos, cur_pos = linear_sum_assignment(cost)
            ids[cur_pos] = self.tracks[-1][prev_pos]
        open_slots = ids < 0
        ids[open_slots] = self._fresh(int(open_slots.sum()))
        self.tracks.append(ids)


def assign_track_ids(energies: Sequence[np.ndarray]) -> tuple[np.ndarray, ...]:
    book = _TrackBook()
    previous = None
    for current in energies:
        book.extend(previous, current)
        previous = current
    return tuple(book.tracks)


# ------------------------------------------------------------------ outputs
def _bands_csv(path: Path, result: BandResult, distance: np.ndarray, energies, tracks) -> None:
    lines = [formats.csv_header()]
    for point, x, ids, values in zip(result.points, distance, tracks, energies):
        lines.extend(formats.csv_line((point.index, x, t, e)) for t, e in zip(ids, values))
    with formats.atomic_output(path, "w") as handle:
        handle.write("".join(lines))


def _summary(result: BandResult, backend: str) -> dict[str, Any]:
    points = result.points
    stage_totals: dict[str, float] = {}
    for point in points:
        for name, seconds in point.stage_seconds.items():
            stage_totals[name] = stage_totals.get(name, 0.0) + float(seconds)
    peaks = [p.peak_gpu_memory_bytes for p in points if p.peak_gpu_memory_bytes is not None]
    warnings = [*result.warnings, *(w for p in points for w in p.warnings)]
    values = {
        "backend": backend,
        "scalar_dtype": "complex64",
        "kpoint_count": result.kpath.n_points,
        "energy_interval_ev": [result.display_interval_ev.lower_ev, result.display_interval_ev.upper_ev],
        "reference_energy_ev": result.reference_energy_ev,
        "maximum_eigenpair_residual_ev": max((p.maximum_residual_ev for p in points), default=0.0),
        "elapsed_seconds": result.elapsed_seconds,
        "assembly_seconds": sum(p.assembly_seconds for p in points),
        "solver_stage_seconds": stage_totals,
        "peak_gpu_memory_bytes": max(peaks, default=None),
        "warnings": list(dict.fromkeys(warnings)),
    }
    return {key: values[key] for key in formats.SUMMARY_KEYS}


def _track_segments(x: np.ndarray, energies, tracks) -> np.ndarray:
    """Line pieces joining each track's states at consecutive k points."""
    pieces = []
    for i in range(x.size - 1):
        _, left, right = np.intersect1d(tracks[i], tracks[i + 1], assume_unique=True, return_indices=True)
        if left.size:
            start = np.column_stack((np.full(left.size, x[i]), energies[i][left]))
            stop = np.column_stack((np.full(right.size, x[i + 1]), energies[i + 1][right]))
            pieces.append(np.stack((start, stop), axis=1))
    return np.concatenate(pieces) if pieces else np.empty((0, 2, 2))


def _band_png(path: Path, result: BandResult, distance: np.ndarray, energies, tracks) -> None:
    try:
        import matplotlib

        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
        from matplotlib.collections import LineCollection
    except ImportError:  # the figure is optional output
        return
    style = formats.PLOT_STYLE
    fig, ax = plt.subplots(figsize=style["figsize"], constrained_layout=True)
    segments = _track_segments(distance, energies, tracks)
    if segments.size:
        ax.add_collection(LineCollection(segments, colors=style["line_color"], linewidths=style["line_width"],
                                         antialiased=True))
    node_x = distance[result.kpath.node_indices]
    ax.set_xlim(float(distance[0]), float(distance[-1]))
    ax.set_ylim(result.display_interval_ev.lower_ev, result.display_interval_ev.upper_ev)
    for x in node_x[1:-1]:
        ax.axvline(x, color=style["node_color"], linewidth=style["node_width"], zorder=0)
    ax.axhline(0.0, color=style["zero_color"], linewidth=style["zero_width"], linestyle="--")
    ax.set_xticks(node_x, result.kpath.node_labels)
    ax.set_xlabel(style["xlabel"])
    ax.set_ylabel(style["ylabel"])
    ax.tick_params(direction="in")
    fig.savefig(path, dpi=style["dpi"])
    plt.close(fig)


def write_band_output(result: BandResult, root: str | Path, *, backend: str = "ciss-hip") -> Path:
    """bands.csv, summary.json and band.png in ``root``."""
    root = Path(root)
    root.mkdir(parents=True, exist_ok=True)
    distance = result.kpath.distances(result.cell_angstrom)
    energies = result.display_energies_ev
    tracks = assign_track_ids(energies)
    _bands_csv(root / "bands.csv", result, distance, energies, tracks)
    with formats.atomic_output(root / "summary.json", "w") as handle:
        handle.write(formats.json_text(_summary(result, backend)))
    _band_png(root / "band.png", result, distance, energies, tracks)
    return root
