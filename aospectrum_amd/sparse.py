"""Bloch assembly of H(k), S(k) on one shared CSR pattern -- HIP-backed.

Mirrors the reference's ``SparsePair`` / ``BlochPattern`` / ``BlochWorkspace``
interface (src/aospectrum/sparse.py:39-258).  The symbolic part
(``BlochPattern.from_bundle``) runs in the native library once per bundle
together with the ordering and tile analysis of the shifted solver; the
numeric part (``BlochWorkspace.update``) runs the HIP assembly kernel and
hands back host CSR matrices for callers that want them.  The Band path itself
never downloads H(k): ``aosx_band_solve_v1`` assembles on the device.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Iterable

import numpy as np
from scipy import sparse

from . import _native as N
from .data import LocalizedOperatorBundle
from .errors import InputError, SolverError


@dataclass(slots=True)
class SparsePair:
    """H(k) and S(k) on one CSR pattern (sparse.py:39-47)."""

    hamiltonian: sparse.csr_matrix
    overlap: sparse.csr_matrix
    k_fractional: tuple[float, float, float]

    @property
    def n_orbitals(self) -> int:
        return int(self.hamiltonian.shape[0])


class NativeBundle:
    """Owner of one ``aosx_bundle`` handle: pattern, ordering, tile structure
    and (with ``device >= 0``) the device-resident block data."""

    def __init__(
        self,
        bundle: LocalizedOperatorBundle,
        *,
        device: int = 0,
        ordering: str = "auto",
    ) -> None:
        ops = bundle.operators
        self.n = bundle.n_orbitals
        real = ops.h_values.dtype == np.float32 and ops.s_values.dtype == np.float32
        kind = np.float32 if real else np.complex64
        self._arrays = dict(
            atom_offsets=np.ascontiguousarray(bundle.basis.atom_offsets, dtype=np.int64),
            row_atom=np.ascontiguousarray(ops.row_atom, dtype=np.int32),
            col_atom=np.ascontiguousarray(ops.col_atom, dtype=np.int32),
            lattice_shift=np.ascontiguousarray(ops.lattice_shift, dtype=np.int32),
            block_offsets=np.ascontiguousarray(ops.block_offsets, dtype=np.int64),
            h=np.ascontiguousarray(ops.h_values, dtype=kind),
            s=np.ascontiguousarray(ops.s_values, dtype=kind),
            positions=np.ascontiguousarray(bundle.structure.positions_angstrom, dtype=np.float64),
        )
        self._create(bundle.basis.n_atoms, ops.n_blocks, 0 if real else 1, device, ordering)

    @classmethod
    def from_csr(cls, indptr: np.ndarray, indices: np.ndarray, n: int, *, device: int = 0) -> "NativeBundle":
        """A bundle whose 'atoms' are single orbitals and whose blocks are the
        CSR entries (R = 0, unit values): gives the shifted solver the
        symbolic analysis of an arbitrary shared H/S pattern."""
        self = cls.__new__(cls)
        self.n = int(n)
        indptr = np.asarray(indptr, dtype=np.int64)
        indices = np.asarray(indices, dtype=np.int64)
        nnz = int(indices.size)
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
        self._arrays = dict(
            atom_offsets=np.arange(n + 1, dtype=np.int64),
            row_atom=rows.astype(np.int32),
            col_atom=indices.astype(np.int32),
            lattice_shift=np.zeros((nnz, 3), dtype=np.int32),
            block_offsets=np.arange(nnz + 1, dtype=np.int64),
            h=np.ones(nnz, dtype=np.float32),
            s=np.ones(nnz, dtype=np.float32),
            positions=None,
        )
        self._create(n, nnz, 0, device, "minimum_degree")
        return self

    def _create(self, n_atoms: int, n_blocks: int, kind: int, device: int, ordering: str) -> None:
        lib = N.load()
        a = self._arrays
        req = N.new(N.BundleRequest)
        req.n_orbitals = self.n
        req.n_atoms = int(n_atoms)
        req.n_blocks = int(n_blocks)
        req.atom_offsets = N.ptr(a["atom_offsets"], C.c_int64)
        req.row_atom = N.ptr(a["row_atom"], C.c_int32)
        req.col_atom = N.ptr(a["col_atom"], C.c_int32)
        req.lattice_shift = N.ptr(a["lattice_shift"], C.c_int32)
        req.block_offsets = N.ptr(a["block_offsets"], C.c_int64)
        req.value_kind = kind
        req.device = int(device)
        req.h_values = a["h"].ctypes.data
        req.s_values = a["s"].ctypes.data
        req.tile_size = 0
        req.ordering = {"auto": 0, "natural": 1, "minimum_degree": 2, "nested_dissection": 3}[ordering]
        if a["positions"] is not None:
            req.atom_positions = N.ptr(a["positions"], C.c_double)
        res = N.BundleResult()
        handle = C.c_void_p()
        rc = lib.aosx_bundle_create_v1(C.byref(req), C.byref(res), C.byref(handle))
        N.check(rc, res, "bundle analysis")
        self.handle = handle
        self.info = {name: getattr(res, name) for name, _ in N.BundleResult._fields_ if name not in ("error_message",)}
        self.nnz = int(res.nnz)
        self.device = int(device)
        # the host copies are no longer needed once resident (or analysed)
        self._arrays = None

    def pattern(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        indptr = np.empty(self.n + 1, dtype=np.int32)
        indices = np.empty(self.nnz, dtype=np.int32)
        perm = np.empty(self.n, dtype=np.int32)
        res = N.StatusResult()
        rc = N.load().aosx_bundle_pattern_v1(
            self.handle, N.ptr(indptr, C.c_int32), N.ptr(indices, C.c_int32), N.ptr(perm, C.c_int32), C.byref(res)
        )
        N.check(rc, res, "pattern query")
        return indptr, indices, perm

    def assemble(self, kpoints: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        k = np.ascontiguousarray(np.asarray(kpoints, dtype=np.float64).reshape(-1, 3))
        h = np.empty((k.shape[0], self.nnz), dtype=np.complex64)
        s = np.empty((k.shape[0], self.nnz), dtype=np.complex64)
        req = N.new(N.AssembleRequest)
        req.n_kpoints = k.shape[0]
        req.k_fractional = N.ptr(k, C.c_double)
        req.h_data = N.ptr(h, N.Complex64)
        req.s_data = N.ptr(s, N.Complex64)
        res = N.StatusResult()
        N.check(N.load().aosx_assemble_v1(self.handle, C.byref(req), C.byref(res)), res, "Bloch assembly")
        return h, s

    def trim(self) -> int:
        """Free the device workspace; returns the bytes still resident."""
        held = C.c_uint64(0)
        if N.load().aosx_bundle_trim_v1(self.handle, C.byref(held)) != 0:
            raise SolverError("cannot release the bundle workspace")
        return int(held.value)

    def autotune(
        self, matrices: int, precision: str = "bf16x3_ir", candidates: tuple[int, ...] = (), repeats: int = 1
    ) -> dict:
        """Time factorisations of ``matrices`` shifted matrices for each
        step-group size (default {4, 8, 12, 16, 24, 32}) and keep the fastest
        for this bundle.  Returns the chosen size and the per-candidate
        milliseconds per factorisation launch."""
        req = N.new(N.AutotuneRequest)
        req.matrices = int(matrices)
        req.precision = N.PRECISION[precision]
        req.repeats = int(repeats)
        cand = np.ascontiguousarray(np.asarray(candidates, dtype=np.int32))
        req.n_candidates = cand.size
        if cand.size:
            req.candidates = N.ptr(cand, C.c_int32)
        res = N.AutotuneResult()
        rc = N.load().aosx_bundle_autotune_v1(self.handle, C.byref(req), C.byref(res))
        N.check(rc, res, "step-group autotuning")
        n = int(res.n_candidates)
        return {
            "chosen_group": int(res.chosen_group),
            "factor_ms": {int(res.candidates[i]): float(res.factor_ms[i]) for i in range(n)},
            "matrices": int(matrices),
            "precision": precision,
        }

    def kernel_stats(self, enable: bool) -> dict:
        out = N.KernelStats()
        N.load().aosx_kernel_stats_read_v1(self.handle, 1 if enable else 0, C.byref(out))
        return {name: getattr(out, name) for name, _ in N.KernelStats._fields_[2:]}

    def close(self) -> None:
        handle = getattr(self, "handle", None)
        if handle is not None and handle.value:
            N.load().aosx_bundle_destroy_v1(handle)
        self.handle = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


@dataclass(frozen=True, slots=True)
class BlochPattern:
    """Shared CSR pattern of H(k), S(k) (sparse.py:50-144)."""

    n_orbitals: int
    csr_indices: np.ndarray
    csr_indptr: np.ndarray
    native: NativeBundle

    @classmethod
    def from_bundle(cls, bundle: LocalizedOperatorBundle, *, device: int = 0) -> "BlochPattern":
        native = NativeBundle(bundle, device=device)
        indptr, indices, _ = native.pattern()
        indptr.setflags(write=False)
        indices.setflags(write=False)
        return cls(bundle.n_orbitals, indices, indptr, native)

    @property
    def nnz(self) -> int:
        return int(self.csr_indices.size)

    def workspace(self, *, complex_values: bool) -> "BlochWorkspace":
        if not complex_values:
            raise InputError("the Band path assembles complex64 H(k), S(k) only")
        return BlochWorkspace(self)


@dataclass(slots=True)
class BlochWorkspace:
    """Per-k assembly into host CSR matrices (sparse.py:175-245)."""

    pattern: BlochPattern

    def update(self, bundle: LocalizedOperatorBundle, k_fractional: Iterable[float]) -> SparsePair:
        kpoint = tuple(float(v) for v in k_fractional)
        if len(kpoint) != 3:
            raise InputError("k point must contain three values")
        h, s = self.pattern.native.assemble(np.asarray(kpoint))
        shape = (self.pattern.n_orbitals, self.pattern.n_orbitals)
        H = sparse.csr_matrix((h[0], self.pattern.csr_indices, self.pattern.csr_indptr), shape=shape, copy=False)
        S = sparse.csr_matrix((s[0], self.pattern.csr_indices, self.pattern.csr_indptr), shape=shape, copy=False)
        return SparsePair(H, S, kpoint)


def assemble_bloch(bundle: LocalizedOperatorBundle, k_fractional: Iterable[float], *, complex_values: bool = True) -> SparsePair:
    """One-shot assembly (sparse.py:248-258)."""
    pattern = BlochPattern.from_bundle(bundle)
    return pattern.workspace(complex_values=complex_values).update(bundle, k_fractional)
