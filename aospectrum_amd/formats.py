"""The on-disk contract of ``aospectrum band`` as data.

Everything a reader of the reference's outputs depends on is declared here
once, and the writers and readers of ``band.py`` are driven by these tables:

* the per-k checkpoint ``k-%06d.npz`` -- field names and dtypes of
  src/aospectrum/band.py:281-309;
* ``bands.csv`` -- its header and number formats (band.py:403-431);
* ``summary.json`` -- its keys (band.py:433-468);
* ``band.png`` -- the figure geometry and colours (band.py:470-499);
* the k-path document schema (band.py:23, 101-129).

The codecs are generic: a checkpoint is a list of (field, encoder, decoder)
rows, a CSV row is a tuple formatted by a column table.
"""

from __future__ import annotations

import contextlib
import json
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Any, Callable, Iterator

import numpy as np

KPATH_SCHEMA = "aospectrum.kpath/v1"
CHECKPOINT_PATTERN = "k-{index:06d}.npz"
RUN_DOCUMENT = "run.json"


@contextlib.contextmanager
def atomic_output(path: Path, mode: str = "wb") -> Iterator[Any]:
    """Write to a hidden sibling and rename it into place when the block
    completes, so that a reader never sees a partial file."""
    path.parent.mkdir(parents=True, exist_ok=True)
    staging = path.with_name(f".{path.name}.tmp")
    handle = staging.open(mode, **({} if "b" in mode else {"encoding": "utf-8", "newline": ""}))
    try:
        yield handle
    except BaseException:
        handle.close()
        staging.unlink(missing_ok=True)
        raise
    handle.close()
    os.replace(staging, path)


def json_text(document: dict) -> str:
    """The JSON layout of run.json and summary.json: sorted keys, two-space
    indent, trailing newline."""
    return json.dumps(document, indent=2, sort_keys=True) + "\n"


# ------------------------------------------------------------ checkpoints
@dataclass(frozen=True)
class NpzField:
    key: str
    encode: Callable[[Any], np.ndarray]   # point -> array
    decode: Callable[[Any], Any]          # stored array -> python value


def _scalar(dtype):
    return lambda value: np.asarray(value, dtype=dtype)


# The checkpoint of one k point: (npz key, value of the point, stored form).
# Stage timings are stored as two parallel arrays of names and seconds; a
# missing peak-memory figure is stored as -1.
CHECKPOINT_FIELDS: tuple[NpzField, ...] = (
    NpzField("index", lambda p: _scalar(np.int64)(p.index), lambda a: int(a)),
    NpzField("k_fractional", lambda p: p.k_fractional, lambda a: np.asarray(a)),
    NpzField("energies_ev", lambda p: p.energies_ev, lambda a: a),
    NpzField("maximum_residual_ev", lambda p: _scalar(np.float64)(p.maximum_residual_ev), lambda a: float(a)),
    NpzField("stage_names", lambda p: np.asarray(tuple(p.stage_seconds), dtype="U64"), lambda a: a.tolist()),
    NpzField("stage_values", lambda p: np.asarray(tuple(p.stage_seconds.values()), dtype=np.float64),
             lambda a: a.tolist()),
    NpzField("assembly_seconds", lambda p: _scalar(np.float64)(p.assembly_seconds), lambda a: float(a)),
    NpzField("peak_gpu_memory_bytes",
             lambda p: _scalar(np.int64)(-1 if p.peak_gpu_memory_bytes is None else p.peak_gpu_memory_bytes),
             lambda a: None if int(a) < 0 else int(a)),
    NpzField("warnings", lambda p: np.asarray(p.warnings, dtype="U256"), lambda a: tuple(a.tolist())),
)


def encode_checkpoint(point) -> dict[str, np.ndarray]:
    return {f.key: f.encode(point) for f in CHECKPOINT_FIELDS}


def decode_checkpoint(arrays) -> dict[str, Any]:
    return {f.key: f.decode(arrays[f.key]) for f in CHECKPOINT_FIELDS}


# --------------------------------------------------------------- bands.csv
# (column, formatter); rows end in CRLF, the csv module's default dialect
BANDS_COLUMNS: tuple[tuple[str, Callable[[Any], str]], ...] = (
    ("k_index", lambda v: str(int(v))),
    ("k_distance_inv_angstrom", lambda v: f"{float(v):.12e}"),
    ("track_id", lambda v: str(int(v))),
    ("energy_ev", lambda v: f"{float(v):.12e}"),
)
CSV_LINE_END = "\r\n"


def csv_line(values) -> str:
    return ",".join(fmt(v) for (_, fmt), v in zip(BANDS_COLUMNS, values)) + CSV_LINE_END


def csv_header() -> str:
    return ",".join(name for name, _ in BANDS_COLUMNS) + CSV_LINE_END


# ------------------------------------------------------------ summary.json
SUMMARY_KEYS = (
    "backend",
    "scalar_dtype",
    "kpoint_count",
    "energy_interval_ev",
    "reference_energy_ev",
    "maximum_eigenpair_residual_ev",
    "elapsed_seconds",
    "assembly_seconds",
    "solver_stage_seconds",
    "peak_gpu_memory_bytes",
    "warnings",
)

# ---------------------------------------------------------------- band.png
PLOT_STYLE = {
    "figsize": (7.2, 5.2),
    "dpi": 240,
    "line_color": "#17213a",
    "line_width": 1.05,
    "node_color": "#9aa1ad",
    "node_width": 0.7,
    "zero_color": "#b73b3b",
    "zero_width": 0.8,
    "xlabel": "K path",
    "ylabel": "Energy relative to reference (eV)",
}
