"""Seeded synthetic operator bundles for the Band-path workloads.

The Band path has no public bundles, so every workload is generated here from
a geometric recipe (jittered atom grid, 13 atomic orbitals per atom, 6 Angstrom
block cutoff over lattice shifts R in {-1,0,1}^d):

* H off-site block entries ``N(0,1) * exp(-d / 1.5)`` eV, onsite diagonal
  ``U[-12, 4]`` eV plus symmetric onsite couplings ``N(0, 0.1)``;
* S onsite identity, off-site entries ``a_S * N(0,1) * exp(-d / 1.5)`` with
  ``a_S = 0.08``, then rescaled globally so that the absolute off-diagonal row
  sum (a bound on every S(k) row sum) is at most 0.5 -- S(k) is then
  Gershgorin-positive-definite for every k;
* the block of (j, i, -R) is the transpose of (i, j, R) (real values), so H(k)
  and S(k) are Hermitian for every k.

One calibration knob is added: onsite diagonal energies are drawn from
``U[-12, 4]`` with the band ``(gap_center - gap, gap_center + gap)`` removed.
With the literal recipe the CISS windows would hold hundreds to thousands of
states, far beyond a 128-column search space; the gap half-widths below were
chosen by measuring the in-window count with an FP64 dense eigensolver or
Sylvester inertia: small2d 40-41 states in [-2.02, 2.02] eV, stress2d 64 in
[-1.52, 1.52] eV, bulk3d 66-67 in [-1.02, 1.02] eV.  The slabs are too large
for a dense or inertia count on the host; their gaps were chosen with the GPU
CISS itself (block 64, moment 4, two filter passes) at Gamma, keeping the
smallest gap whose count stays well below the search dimension with an FP64
eigenpair residual below 1e-4 eV: slab52k 68 states in [-0.82, 0.82] eV
(gap 2.75), slab26k 53 states in [-2.02, 2.02] eV (gap 4.5), the widest
serving window.  Every spec records its seed and knob in the bundle
provenance.

The precision-stress bundle (stress2d) clusters its onsite levels on 96
centres with 1e-3 eV spread and scales the off-site overlap so that the
smallest eigenvalue of S(k) is 1.0e-3 along the path (no row-sum rescale).
"""

from __future__ import annotations

from dataclasses import asdict, dataclass, replace

import numpy as np
from scipy.spatial import cKDTree

from .data import (
    AtomicStructure,
    ElectronicFilling,
    LocalizedOperatorBlocks,
    LocalizedOperatorBundle,
    OrbitalBasisLayout,
)
from .errors import InputError


@dataclass(frozen=True)
class SynthSpec:
    """Geometry and value recipe of one synthetic bundle."""

    name: str
    grid: tuple[int, int, int]
    spacing: float
    cell: tuple[float, float, float]
    shift_axes: tuple[bool, bool, bool]
    seed: int
    aos_per_atom: int = 13
    cutoff: float = 6.0
    jitter: float = 0.1
    decay: float = 1.5
    onsite_low: float = -12.0
    onsite_high: float = 4.0
    onsite_coupling: float = 0.1
    s_amplitude: float = 0.08
    s_row_bound: float = 0.5
    gap_center: float = 0.0
    gap: float = 0.0
    gap_fill: float = 0.0
    cluster_spread: float | None = None
    cluster_count: int = 0

    @property
    def n_atoms(self) -> int:
        return int(np.prod(self.grid))

    @property
    def n_orbitals(self) -> int:
        return self.n_atoms * self.aos_per_atom


# Workload recipes.  Sizes follow the five Band-path workloads: a 2D 832-orbital
# system, a 3,328-orbital precision-stress system, a 13,000-orbital 3D bulk, a
# 52,000-orbital slab and 26,000-orbital slab variants for serving load.
SPECS: dict[str, SynthSpec] = {
    "small2d": SynthSpec(
        "small2d", (8, 8, 1), 2.5, (20.0, 20.0, 30.0), (True, True, False), 1000, gap=2.05
    ),
    "stress2d": SynthSpec(
        "stress2d",
        (16, 16, 1),
        2.5,
        (40.0, 40.0, 30.0),
        (True, True, False),
        1001,
        gap=2.2,
        # no row-sum rescale; the amplitude puts min eig S(k) at 1.0e-3 on the
        # Y-G-X-M path (measured: unscaled off-site overlap has min eig -3.2555)
        s_amplitude=0.30686,
        s_row_bound=1.0e30,
        cluster_spread=1.0e-3,
        cluster_count=96,
    ),
    "bulk3d": SynthSpec(
        "bulk3d", (10, 10, 10), 2.4, (24.0, 24.0, 24.0), (True, True, True), 1002, gap=2.85
    ),
    "slab52k": SynthSpec(
        "slab52k", (20, 20, 10), 2.4, (48.0, 48.0, 44.0), (True, True, False), 1003, gap=2.75
    ),
    # serving variants: a wide gap with a thin uniform population of levels
    # inside it, so that every window of half-width 0.5..2 eV around 0 holds
    # states (about 14 per eV) and none saturates a block-32 search space
    "slab26k_a": SynthSpec(
        "slab26k_a", (20, 10, 10), 2.4, (48.0, 24.0, 44.0), (True, True, False), 2001, gap=6.0, gap_fill=0.0062
    ),
    "slab26k_b": SynthSpec(
        "slab26k_b", (20, 10, 10), 2.4, (48.0, 24.0, 44.0), (True, True, False), 2002, gap=6.0, gap_fill=0.0062
    ),
}


def scaled_spec(name: str, **changes) -> SynthSpec:
    """A registered spec with some fields replaced (used for small test bundles)."""
    return replace(SPECS[name], **changes)


def _positions(spec: SynthSpec, rng: np.random.Generator) -> np.ndarray:
    ix, iy, iz = np.meshgrid(*(np.arange(n) for n in spec.grid), indexing="ij")
    base = np.stack((ix.ravel(), iy.ravel(), iz.ravel()), axis=1) * spec.spacing
    return base + rng.uniform(-spec.jitter, spec.jitter, size=base.shape)


def _neighbour_blocks(spec: SynthSpec, pos: np.ndarray, cell: np.ndarray):
    """All (i, j, R, d) with |r_j + R.cell - r_i| <= cutoff, R over the shift set."""
    ranges = [(-1, 0, 1) if on else (0,) for on in spec.shift_axes]
    tree = cKDTree(pos)
    rows, cols, shifts, dists = [], [], [], []
    for r0 in ranges[0]:
        for r1 in ranges[1]:
            for r2 in ranges[2]:
                shift = np.array((r0, r1, r2), dtype=np.int64)
                other = pos + shift @ cell
                pairs = tree.sparse_distance_matrix(
                    cKDTree(other), spec.cutoff, output_type="ndarray"
                )
                if pairs.size == 0:
                    continue
                rows.append(pairs["i"].astype(np.int64))
                cols.append(pairs["j"].astype(np.int64))
                shifts.append(np.broadcast_to(shift, (pairs.size, 3)))
                dists.append(pairs["v"])
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    shifts = np.concatenate(shifts)
    dists = np.concatenate(dists)
    # sparse_distance_matrix drops exact zeros: re-add the onsite blocks.
    onsite = np.arange(pos.shape[0], dtype=np.int64)
    keep = ~((rows == cols) & np.all(shifts == 0, axis=1))
    rows = np.concatenate((onsite, rows[keep]))
    cols = np.concatenate((onsite, cols[keep]))
    shifts = np.concatenate((np.zeros((onsite.size, 3), np.int64), shifts[keep]))
    dists = np.concatenate((np.zeros(onsite.size), dists[keep]))
    order = np.lexsort((shifts[:, 2], shifts[:, 1], shifts[:, 0], cols, rows))
    return rows[order], cols[order], shifts[order], dists[order]


def _onsite_levels(spec: SynthSpec, rng: np.random.Generator, count: int) -> np.ndarray:
    lo, hi = spec.onsite_low, spec.onsite_high
    a, b = spec.gap_center - spec.gap, spec.gap_center + spec.gap
    a, b = max(a, lo), min(b, hi)
    removed = max(b - a, 0.0)
    if removed >= hi - lo:
        raise InputError("onsite gap removes the whole onsite range")
    if spec.cluster_spread is not None and spec.cluster_count > 0:
        u = rng.uniform(0.0, (hi - lo) - removed, size=spec.cluster_count) + lo
        centres = np.where(u < a, u, u + removed)
        pick = rng.integers(0, spec.cluster_count, size=count)
        return centres[pick] + rng.uniform(-0.5, 0.5, size=count) * spec.cluster_spread
    u = rng.uniform(0.0, (hi - lo) - removed, size=count) + lo
    levels = np.where(u < a, u, u + removed)
    if spec.gap_fill > 0.0 and removed > 0.0:
        # a thin, uniform population inside the removed band (drawn after the
        # other levels, so specs without it are unchanged)
        inside = rng.uniform(size=count) < spec.gap_fill
        levels = np.where(inside, rng.uniform(a, b, size=count), levels)
    return levels


def generate(spec: SynthSpec) -> LocalizedOperatorBundle:
    """Build the bundle of ``spec`` (deterministic in ``spec.seed``)."""
    rng = np.random.default_rng(spec.seed)
    cell = np.diag(np.asarray(spec.cell, dtype=np.float64))
    pos = _positions(spec, rng)
    rows, cols, shifts, dists = _neighbour_blocks(spec, pos, cell)
    nb = rows.size
    m = spec.aos_per_atom
    bsz = m * m

    # canonical representative of every Hermitian pair: i < j, or i == j with
    # R lexicographically positive; onsite blocks (i, i, 0) are their own pair
    lex = shifts[:, 0] * 9 + shifts[:, 1] * 3 + shifts[:, 2]
    onsite = (rows == cols) & (lex == 0)
    canonical = (rows < cols) | ((rows == cols) & (lex > 0))
    n_atoms = spec.n_atoms
    code = (shifts[:, 0] + 1) * 9 + (shifts[:, 1] + 1) * 3 + (shifts[:, 2] + 1)
    keys = (rows * n_atoms + cols) * 27 + code  # ascending: blocks are lexsorted
    mirror_keys = (cols * n_atoms + rows) * 27 + (26 - code)
    mirror = np.searchsorted(keys, mirror_keys)
    if np.any(mirror >= nb) or not np.array_equal(keys[np.minimum(mirror, nb - 1)], mirror_keys):
        raise InputError("neighbour list is not closed under (i, j, R) -> (j, i, -R)")

    h = np.zeros((nb, m, m), dtype=np.float64)
    s = np.zeros((nb, m, m), dtype=np.float64)
    canon = np.flatnonzero(canonical)
    env = np.exp(-dists[canon] / spec.decay)[:, None, None]
    h[canon] = rng.standard_normal((canon.size, m, m)) * env
    s[canon] = spec.s_amplitude * rng.standard_normal((canon.size, m, m)) * env
    h[mirror[canon]] = np.transpose(h[canon], (0, 2, 1))
    s[mirror[canon]] = np.transpose(s[canon], (0, 2, 1))

    # global S rescale so that every absolute off-diagonal row sum is bounded
    row_sum = np.zeros(spec.n_orbitals)
    offsite = ~onsite
    np.add.at(row_sum, (rows[offsite][:, None] * m + np.arange(m)[None, :]).ravel(),
              np.abs(s[offsite]).sum(axis=2).ravel())
    peak = float(row_sum.max(initial=0.0))
    if peak > spec.s_row_bound:
        s *= spec.s_row_bound / peak

    onsite_ids = np.flatnonzero(onsite)
    atom_of_block = rows[onsite_ids]
    levels = _onsite_levels(spec, rng, spec.n_orbitals).reshape(spec.n_atoms, m)
    coupling = rng.standard_normal((onsite_ids.size, m, m)) * spec.onsite_coupling
    coupling = np.triu(coupling, 1)
    coupling = coupling + np.transpose(coupling, (0, 2, 1))
    eye = np.eye(m)
    h[onsite_ids] = coupling + levels[atom_of_block][:, :, None] * eye
    s[onsite_ids] = eye

    orbital_atom = np.repeat(np.arange(n_atoms, dtype=np.int32), m)
    local = np.tile(np.arange(m, dtype=np.int32), n_atoms)
    ops = LocalizedOperatorBlocks(
        row_atom=rows.astype(np.int32),
        col_atom=cols.astype(np.int32),
        lattice_shift=shifts.astype(np.int32),
        block_offsets=np.arange(nb + 1, dtype=np.int64) * bsz,
        block_shapes=np.full((nb, 2), m, dtype=np.int32),
        h_values=h.reshape(-1).astype(np.float32),
        s_values=s.reshape(-1).astype(np.float32),
    )
    return LocalizedOperatorBundle(
        structure=AtomicStructure(
            cell_angstrom=cell,
            positions_angstrom=pos,
            atomic_numbers=np.full(n_atoms, 6, dtype=np.int32),
            periodic=tuple(bool(v) for v in spec.shift_axes),
            label=spec.name,
        ),
        basis=OrbitalBasisLayout(
            n_atoms=n_atoms,
            orbital_atom=orbital_atom,
            local_index=local,
            basis_family="synthetic",
            basis_identity=f"synthetic-{m}ao",
        ),
        operators=ops,
        filling=ElectronicFilling(electron_count=2 * (spec.n_orbitals // 2)),
        provenance={"generator": "aospectrum_amd.synth", "spec": asdict(spec)},
    )
