"""Reader and writer of the ``aospectrum.operator-bundle/v1`` directory layout.

The on-disk contract is the reference's (src/aospectrum/bundle.py:21-191 and
README.md:57-80): one ``.npy`` file per array under ``structure/``, ``basis/``
and ``operators/``, plus ``manifest.json``, ``basis/layout.json``, optional
``filling.json`` and ``provenance.json``.  File names, JSON keys and the
"destination must be empty" rule are kept identical so bundles are
interchangeable between the two implementations.  Arrays load memory-mapped
and never through pickle.
"""

from __future__ import annotations

import json
from pathlib import Path
from typing import Any

import numpy as np

from .data import (
    AtomicStructure,
    ElectronicFilling,
    LocalizedOperatorBlocks,
    LocalizedOperatorBundle,
    OrbitalBasisLayout,
)
from .errors import InputError

SCHEMA = "aospectrum.operator-bundle/v1"

# relative path -> attribute getter, in the order the layout defines them
_ARRAYS = (
    ("structure/cell.npy", lambda b: b.structure.cell_angstrom),
    ("structure/positions.npy", lambda b: b.structure.positions_angstrom),
    ("structure/atomic_numbers.npy", lambda b: b.structure.atomic_numbers),
    ("basis/orbital_atom.npy", lambda b: b.basis.orbital_atom),
    ("basis/local_index.npy", lambda b: b.basis.local_index),
    ("operators/row_atom.npy", lambda b: b.operators.row_atom),
    ("operators/col_atom.npy", lambda b: b.operators.col_atom),
    ("operators/lattice_shift.npy", lambda b: b.operators.lattice_shift),
    ("operators/block_offsets.npy", lambda b: b.operators.block_offsets),
    ("operators/block_shapes.npy", lambda b: b.operators.block_shapes),
    ("operators/h_values.npy", lambda b: b.operators.h_values),
    ("operators/s_values.npy", lambda b: b.operators.s_values),
)
_DESCRIPTORS = (
    ("basis/angular_momentum.npy", lambda b: b.basis.angular_momentum),
    ("basis/radial_index.npy", lambda b: b.basis.radial_index),
    ("basis/real_harmonic_index.npy", lambda b: b.basis.real_harmonic_index),
)


def _read_json(path: Path) -> dict[str, Any]:
    try:
        value = json.loads(path.read_text(encoding="utf-8"))
    except (OSError, UnicodeError, json.JSONDecodeError) as exc:
        raise InputError(f"cannot read {path}: {exc}") from exc
    if not isinstance(value, dict):
        raise InputError(f"{path} must contain a JSON object")
    return value


def _write_json(path: Path, value: dict[str, Any]) -> None:
    path.write_text(json.dumps(value, indent=2, sort_keys=True) + "\n", encoding="utf-8")


def load_bundle(root: str | Path, *, mmap_mode: str | None = "r") -> LocalizedOperatorBundle:
    """Load a bundle directory (bundle.py:41-114); arrays are memory-mapped."""

    root = Path(root).expanduser().resolve()
    manifest = _read_json(root / "manifest.json")
    if manifest.get("schema") != SCHEMA:
        raise InputError(f"unsupported operator bundle: {manifest.get('schema')}")
    layout = _read_json(root / "basis" / "layout.json")

    def arr(rel: str) -> np.ndarray:
        try:
            return np.load(root / rel, allow_pickle=False, mmap_mode=mmap_mode)
        except (OSError, ValueError) as exc:
            raise InputError(f"cannot read bundle array {rel}: {exc}") from exc

    filling = None
    if (root / "filling.json").is_file():
        raw = _read_json(root / "filling.json")
        filling = ElectronicFilling(
            electron_count=raw["electron_count"],
            spin_degeneracy=raw.get("spin_degeneracy", 2),
            fermi_energy_ev=raw.get("fermi_energy_ev"),
        )
    provenance = _read_json(root / "provenance.json") if (root / "provenance.json").is_file() else {}
    described = bool(layout.get("has_real_space_descriptors", False))
    return LocalizedOperatorBundle(
        structure=AtomicStructure(
            cell_angstrom=arr("structure/cell.npy"),
            positions_angstrom=arr("structure/positions.npy"),
            atomic_numbers=arr("structure/atomic_numbers.npy"),
            periodic=tuple(manifest.get("periodic", (True, True, True))),
            label=manifest.get("structure_label"),
        ),
        basis=OrbitalBasisLayout(
            n_atoms=layout["n_atoms"],
            orbital_atom=arr("basis/orbital_atom.npy"),
            local_index=arr("basis/local_index.npy"),
            basis_family=str(layout["basis_family"]),
            basis_identity=str(layout["basis_identity"]),
            orbital_labels=tuple(layout.get("orbital_labels", ())),
            angular_momentum=arr("basis/angular_momentum.npy") if described else None,
            radial_index=arr("basis/radial_index.npy") if described else None,
            real_harmonic_index=arr("basis/real_harmonic_index.npy") if described else None,
            spinor_width=layout.get("spinor_width", 1),
            metadata=dict(layout.get("metadata", {})),
        ),
        operators=LocalizedOperatorBlocks(
            row_atom=arr("operators/row_atom.npy"),
            col_atom=arr("operators/col_atom.npy"),
            lattice_shift=arr("operators/lattice_shift.npy"),
            block_offsets=arr("operators/block_offsets.npy"),
            block_shapes=arr("operators/block_shapes.npy"),
            h_values=arr("operators/h_values.npy"),
            s_values=arr("operators/s_values.npy"),
        ),
        filling=filling,
        provenance=provenance,
    )


def write_bundle(bundle: LocalizedOperatorBundle, root: str | Path) -> Path:
    """Write one bundle into an empty or absent directory (bundle.py:117-191)."""

    root = Path(root).expanduser().resolve()
    if root.exists() and (not root.is_dir() or any(root.iterdir())):
        raise InputError(f"bundle destination is not empty: {root}")
    for sub in ("structure", "basis", "operators"):
        (root / sub).mkdir(parents=True, exist_ok=True)
    entries = list(_ARRAYS)
    if bundle.basis.has_real_space_descriptors:
        entries += list(_DESCRIPTORS)
    for rel, getter in entries:
        np.save(root / rel, getter(bundle), allow_pickle=False)
    _write_json(
        root / "manifest.json",
        {
            "schema": SCHEMA,
            "periodic": list(bundle.structure.periodic),
            "structure_label": bundle.structure.label,
            "hamiltonian_unit": "eV",
            "length_unit": "angstrom",
        },
    )
    _write_json(
        root / "basis" / "layout.json",
        {
            "n_atoms": bundle.basis.n_atoms,
            "basis_family": bundle.basis.basis_family,
            "basis_identity": bundle.basis.basis_identity,
            "orbital_labels": list(bundle.basis.orbital_labels),
            "has_real_space_descriptors": bundle.basis.has_real_space_descriptors,
            "spinor_width": bundle.basis.spinor_width,
            "metadata": bundle.basis.metadata,
        },
    )
    if bundle.filling is not None:
        _write_json(
            root / "filling.json",
            {
                "mode": "closed_shell",
                "electron_count": bundle.filling.electron_count,
                "spin_degeneracy": bundle.filling.spin_degeneracy,
                "fermi_energy_ev": bundle.filling.fermi_energy_ev,
            },
        )
    _write_json(root / "provenance.json", bundle.provenance)
    return root
