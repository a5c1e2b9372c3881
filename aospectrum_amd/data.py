"""Validated, read-only data model of a localized-orbital operator bundle.

The classes keep the names, fields and validation rules of the reference data
model (src/aospectrum/data.py:27-352) so that bundles produced for the reference
load unchanged.  Every array is converted once to its canonical dtype and frozen
(``writeable = False``); the Band path never mutates bundle data.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .errors import InputError


def _frozen(value: Any, name: str, dtype: Any = None, ndim: int | None = None) -> np.ndarray:
    out = np.asarray(value, dtype=dtype)
    if ndim is not None and out.ndim != ndim:
        raise InputError(f"{name} must be {ndim}-dimensional")
    out.setflags(write=False)
    return out


@dataclass(frozen=True, slots=True)
class AtomicStructure:
    """Cell, Cartesian positions (Angstrom) and atomic numbers (data.py:27-68)."""

    cell_angstrom: np.ndarray
    positions_angstrom: np.ndarray
    atomic_numbers: np.ndarray
    periodic: tuple[bool, bool, bool] = (True, True, True)
    label: str | None = None

    def __post_init__(self) -> None:
        cell = _frozen(self.cell_angstrom, "cell_angstrom", np.float64, 2)
        pos = _frozen(self.positions_angstrom, "positions_angstrom", np.float64, 2)
        z = _frozen(self.atomic_numbers, "atomic_numbers", np.int32, 1)
        if cell.shape != (3, 3) or pos.shape != (z.size, 3):
            raise InputError("structure arrays have incompatible shapes")
        if z.size == 0:
            raise InputError("structure has no atoms")
        periodic = tuple(bool(v) for v in self.periodic)
        if len(periodic) != 3:
            raise InputError("periodic must contain three values")
        object.__setattr__(self, "cell_angstrom", cell)
        object.__setattr__(self, "positions_angstrom", pos)
        object.__setattr__(self, "atomic_numbers", z)
        object.__setattr__(self, "periodic", periodic)

    @property
    def n_atoms(self) -> int:
        return int(self.atomic_numbers.size)


@dataclass(frozen=True, slots=True)
class OrbitalBasisLayout:
    """AO -> atom map with contiguous per-atom AO ranges (data.py:71-180)."""

    n_atoms: int
    orbital_atom: np.ndarray
    local_index: np.ndarray
    basis_family: str
    basis_identity: str
    orbital_labels: tuple[str, ...] = ()
    angular_momentum: np.ndarray | None = None
    radial_index: np.ndarray | None = None
    real_harmonic_index: np.ndarray | None = None
    spinor_width: int = 1
    metadata: dict[str, Any] = field(default_factory=dict)
    _atom_offsets: np.ndarray = field(init=False, repr=False)

    def __post_init__(self) -> None:
        atoms = _frozen(self.orbital_atom, "orbital_atom", np.int32, 1)
        local = _frozen(self.local_index, "local_index", np.int32, 1)
        n_atoms = int(self.n_atoms)
        if (
            n_atoms <= 0
            or atoms.size == 0
            or atoms.shape != local.shape
            or atoms.min() < 0
            or atoms.max() >= n_atoms
            or np.any(atoms[1:] < atoms[:-1])
        ):
            raise InputError("invalid AO-to-atom layout")
        counts = np.bincount(atoms, minlength=n_atoms)
        if np.any(counts == 0):
            raise InputError("each atom must own at least one AO")
        offsets = np.concatenate(([0], np.cumsum(counts))).astype(np.int64)
        if not np.array_equal(local, np.arange(atoms.size) - offsets[atoms]):
            raise InputError("local AO indices must be contiguous within each atom")
        extras = (self.angular_momentum, self.radial_index, self.real_harmonic_index)
        present = [v is not None for v in extras]
        if any(present) and not all(present):
            raise InputError("orbital descriptors must be supplied together")
        if all(present):
            extras = tuple(
                _frozen(v, name, np.int32, 1)
                for v, name in zip(
                    extras, ("angular_momentum", "radial_index", "real_harmonic_index")
                )
            )
            if any(v.shape != atoms.shape for v in extras):
                raise InputError("orbital descriptor lengths do not match AO layout")
        else:
            extras = (None, None, None)
        labels = tuple(str(v) for v in self.orbital_labels)
        if labels and len(labels) != atoms.size:
            raise InputError("orbital_labels must contain one value per AO")
        if int(self.spinor_width) < 1:
            raise InputError("spinor_width must be positive")
        offsets.setflags(write=False)
        for name, value in (
            ("n_atoms", n_atoms),
            ("orbital_atom", atoms),
            ("local_index", local),
            ("orbital_labels", labels),
            ("angular_momentum", extras[0]),
            ("radial_index", extras[1]),
            ("real_harmonic_index", extras[2]),
            ("spinor_width", int(self.spinor_width)),
            ("metadata", dict(self.metadata)),
            ("_atom_offsets", offsets),
        ):
            object.__setattr__(self, name, value)

    @property
    def n_orbitals(self) -> int:
        return int(self.orbital_atom.size)

    @property
    def atom_offsets(self) -> np.ndarray:
        return self._atom_offsets

    @property
    def has_real_space_descriptors(self) -> bool:
        return self.angular_momentum is not None

    def orbital_count(self, atom: int) -> int:
        return int(self._atom_offsets[atom + 1] - self._atom_offsets[atom])


@dataclass(frozen=True, slots=True)
class ElectronicFilling:
    """Closed-shell filling (data.py:183-209)."""

    electron_count: int
    spin_degeneracy: int = 2
    fermi_energy_ev: float | None = None

    def __post_init__(self) -> None:
        electrons, degeneracy = int(self.electron_count), int(self.spin_degeneracy)
        if electrons <= 0 or degeneracy <= 0 or electrons % degeneracy:
            raise InputError("invalid closed-shell filling")
        object.__setattr__(self, "electron_count", electrons)
        object.__setattr__(self, "spin_degeneracy", degeneracy)
        if self.fermi_energy_ev is not None:
            object.__setattr__(self, "fermi_energy_ev", float(self.fermi_energy_ev))

    @property
    def vbm_state_number(self) -> int:
        return self.electron_count // self.spin_degeneracy

    @property
    def cbm_state_number(self) -> int:
        return self.vbm_state_number + 1


@dataclass(frozen=True, slots=True)
class LocalizedOperatorBlocks:
    """Atom-pair block table of H and S (data.py:212-285).

    Block ``b`` couples ``row_atom[b]`` in the home cell with ``col_atom[b]`` in
    the cell displaced by ``lattice_shift[b]``; its values are stored row-major
    in ``h_values[block_offsets[b]:block_offsets[b+1]]`` (and likewise S).
    """

    row_atom: np.ndarray
    col_atom: np.ndarray
    lattice_shift: np.ndarray
    block_offsets: np.ndarray
    block_shapes: np.ndarray
    h_values: np.ndarray
    s_values: np.ndarray

    def __post_init__(self) -> None:
        rows = _frozen(self.row_atom, "row_atom", np.int32, 1)
        cols = _frozen(self.col_atom, "col_atom", np.int32, 1)
        shifts = _frozen(self.lattice_shift, "lattice_shift", np.int32, 2)
        offsets = _frozen(self.block_offsets, "block_offsets", np.int64, 1)
        shapes = _frozen(self.block_shapes, "block_shapes", np.int32, 2)
        h = _frozen(self.h_values, "h_values", None, 1)
        s = _frozen(self.s_values, "s_values", None, 1)
        nb = rows.size
        if (
            nb == 0
            or cols.shape != rows.shape
            or rows.min() < 0
            or cols.min() < 0
            or shifts.shape != (nb, 3)
            or shapes.shape != (nb, 2)
            or offsets.shape != (nb + 1,)
            or offsets[0] != 0
        ):
            raise InputError("invalid operator block table")
        if not np.array_equal(np.diff(offsets), shapes[:, 0].astype(np.int64) * shapes[:, 1]):
            raise InputError("operator block offsets do not match block shapes")
        if h.shape != s.shape or h.size != offsets[-1]:
            raise InputError("H/S values do not match the operator blocks")

        def canonical(values: np.ndarray, name: str) -> np.ndarray:
            if values.dtype in (np.float32, np.complex64):
                return values
            target = np.complex64 if np.iscomplexobj(values) else np.float32
            return _frozen(values, name, target, 1)

        for name, value in (
            ("row_atom", rows),
            ("col_atom", cols),
            ("lattice_shift", shifts),
            ("block_offsets", offsets),
            ("block_shapes", shapes),
            ("h_values", canonical(h, "h_values")),
            ("s_values", canonical(s, "s_values")),
        ):
            object.__setattr__(self, name, value)

    @property
    def n_blocks(self) -> int:
        return int(self.row_atom.size)


@dataclass(frozen=True, slots=True)
class LocalizedOperatorBundle:
    """Structure + basis + operator blocks (data.py:288-314)."""

    structure: AtomicStructure
    basis: OrbitalBasisLayout
    operators: LocalizedOperatorBlocks
    filling: ElectronicFilling | None = None
    provenance: dict[str, Any] = field(default_factory=dict)

    def __post_init__(self) -> None:
        if self.structure.n_atoms != self.basis.n_atoms:
            raise InputError("structure and basis atom counts differ")
        ops = self.operators
        if ops.row_atom.max() >= self.structure.n_atoms or ops.col_atom.max() >= self.structure.n_atoms:
            raise InputError("operator block references an unknown atom")
        sizes = np.diff(self.basis.atom_offsets)
        expected = np.stack((sizes[ops.row_atom], sizes[ops.col_atom]), axis=1)
        if not np.array_equal(ops.block_shapes, expected):
            raise InputError("operator block shapes do not match the AO layout")
        object.__setattr__(self, "provenance", dict(self.provenance))

    @property
    def n_orbitals(self) -> int:
        return self.basis.n_orbitals


@dataclass(frozen=True, slots=True)
class EnergyInterval:
    """Open energy interval (eV) with ``lower < upper`` (data.py:317-334)."""

    lower_ev: float
    upper_ev: float

    def __post_init__(self) -> None:
        lo, hi = float(self.lower_ev), float(self.upper_ev)
        if not (np.isfinite(lo) and np.isfinite(hi) and lo < hi):
            raise InputError("energy interval requires lower < upper")
        object.__setattr__(self, "lower_ev", lo)
        object.__setattr__(self, "upper_ev", hi)

    def expanded(self, guard_ev: float) -> "EnergyInterval":
        return EnergyInterval(self.lower_ev - float(guard_ev), self.upper_ev + float(guard_ev))
