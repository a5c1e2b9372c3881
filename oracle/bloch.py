"""Oracle: host Bloch-sum assembly of H(k), S(k) on the shared CSR pattern.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates src/aospectrum/sparse.py:
* ``Pattern.build``   follows ``BlochPattern.from_bundle`` (sparse.py:63-140):
  every block entry becomes a (row, col) key, keys are stably sorted, equal keys
  form one CSR entry (duplicates summed), the mirror of every key must exist
  (Hermitian topology, sparse.py:115-121), CSR rows come out column-sorted.
* ``Pattern.assemble`` follows ``BlochWorkspace.update``/``_fill``
  (sparse.py:181-245): phases exp(2*pi*i R.k) computed in float32/complex64 as
  the reference does (sparse.py:207-213), gather, multiply, segmented sum, then
  the Hermitian projection of ``_project_hermitian`` (sparse.py:25-36).
``dense_bloch`` is the brute-force closed form used to pin the restatement.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp


class TopologyError(ValueError):
    pass


@dataclass(frozen=True)
class Pattern:
    n: int
    indptr: np.ndarray
    indices: np.ndarray
    source: np.ndarray       # block-value index of every sorted entry
    group: np.ndarray        # unique-shift index of every sorted entry
    starts: np.ndarray       # first sorted entry of every CSR entry
    shifts: np.ndarray       # unique lattice shifts, int32 [G, 3]
    upper: np.ndarray        # CSR positions with row < col
    lower: np.ndarray        # their mirrors
    diagonal: np.ndarray     # CSR positions with row == col

    @property
    def nnz(self) -> int:
        return int(self.indices.size)

    @classmethod
    def build(cls, ops, atom_offsets: np.ndarray, n: int) -> "Pattern":
        shifts, inverse = np.unique(np.asarray(ops.lattice_shift), axis=0, return_inverse=True)
        inverse = inverse.reshape(-1)
        shapes = np.asarray(ops.block_shapes)
        offs = np.asarray(ops.block_offsets)
        counts = shapes[:, 0].astype(np.int64) * shapes[:, 1]
        total = int(counts.sum())
        block_of = np.repeat(np.arange(counts.size), counts)
        local = np.arange(total, dtype=np.int64) - offs[block_of]
        ncols = shapes[block_of, 1].astype(np.int64)
        rows = atom_offsets[np.asarray(ops.row_atom)[block_of]] + local // ncols
        cols = atom_offsets[np.asarray(ops.col_atom)[block_of]] + local % ncols
        keys = rows * n + cols
        order = np.argsort(keys, kind="stable")
        skeys = keys[order]
        first = np.ones(skeys.size, dtype=bool)
        first[1:] = skeys[1:] != skeys[:-1]
        starts = np.flatnonzero(first)
        uniq = skeys[starts]
        urow, ucol = uniq // n, uniq % n
        mkeys = ucol * n + urow
        mirror = np.searchsorted(uniq, mkeys)
        if np.any(mirror >= uniq.size) or not np.array_equal(uniq[np.minimum(mirror, uniq.size - 1)], mkeys):
            raise TopologyError("operator topology is not Hermitian")
        indptr = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(np.bincount(urow, minlength=n), out=indptr[1:])
        upper = np.flatnonzero(urow < ucol)
        return cls(
            n=n,
            indptr=indptr,
            indices=ucol,
            source=order,                      # values are contiguous: index == offset
            group=inverse[block_of[order]],
            starts=starts,
            shifts=shifts.astype(np.int32),
            upper=upper,
            lower=mirror[upper],
            diagonal=np.flatnonzero(urow == ucol),
        )

    def phases(self, k) -> np.ndarray:
        # float32 R.k and a complex64 exponential, as sparse.py:207-213
        arg = self.shifts.astype(np.float32) @ np.asarray(k, dtype=np.float32)
        return np.exp(np.asarray(2.0j * np.pi, dtype=np.complex64) * arg).astype(np.complex64)

    def fill(self, values: np.ndarray, phases: np.ndarray) -> np.ndarray:
        gathered = np.asarray(values)[self.source].astype(np.complex64) * phases[self.group]
        return np.add.reduceat(gathered, self.starts).astype(np.complex64)

    def project(self, data: np.ndarray) -> np.ndarray:
        if self.upper.size:
            mean = (data[self.upper] + np.conj(data[self.lower])) * np.float32(0.5)
            data[self.upper] = mean
            data[self.lower] = np.conj(mean)
        data[self.diagonal] = data[self.diagonal].real
        return data

    def assemble(self, ops, k) -> tuple[np.ndarray, np.ndarray]:
        ph = self.phases(k)
        h = self.project(self.fill(ops.h_values, ph))
        s = self.project(self.fill(ops.s_values, ph))
        return h, s

    def csr(self, data: np.ndarray) -> sp.csr_matrix:
        return sp.csr_matrix((data, self.indices, self.indptr), shape=(self.n, self.n))


def dense_bloch(ops, atom_offsets: np.ndarray, n: int, k, dtype=np.complex128):
    """Closed form: H(k) = sum_b exp(2 pi i R_b.k) H_b placed at the atom block,
    followed by (X + X^H)/2 -- the brute-force statement of the assembly."""
    H = np.zeros((n, n), dtype=dtype)
    S = np.zeros((n, n), dtype=dtype)
    shifts = np.asarray(ops.lattice_shift, dtype=np.float64)
    phase = np.exp(2j * np.pi * (shifts @ np.asarray(k, dtype=np.float64)))
    offs = np.asarray(ops.block_offsets)
    shapes = np.asarray(ops.block_shapes)
    hv = np.asarray(ops.h_values)
    sv = np.asarray(ops.s_values)
    for b in range(shapes.shape[0]):
        r0 = atom_offsets[ops.row_atom[b]]
        c0 = atom_offsets[ops.col_atom[b]]
        nr, nc = int(shapes[b, 0]), int(shapes[b, 1])
        blk = slice(offs[b], offs[b + 1])
        H[r0:r0 + nr, c0:c0 + nc] += hv[blk].reshape(nr, nc) * phase[b]
        S[r0:r0 + nr, c0:c0 + nc] += sv[blk].reshape(nr, nc) * phase[b]
    H = 0.5 * (H + H.conj().T)
    S = 0.5 * (S + S.conj().T)
    return H, S
