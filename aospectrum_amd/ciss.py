"""Complex64 CISS interval solver on the MI355X (HIP through libaosx_hip.so).

Keeps the reference's solver interface (src/aospectrum/ciss.py:18-391):
``CissSettings``, ``CissResult``, ``ShiftedResult``, a shifted-solver plugin
with ``solve_shifted(matrices, shifts, rhs)`` and an interval solver with
``solve_interval(matrices, interval)``.  Everything numerical -- assembly,
shifted factorisations and solves, moments, compression, Rayleigh-Ritz, ghost
filter and residuals -- runs in the native library; this module validates
settings, builds the probe block and converts results.

Settings beyond the reference's (all default to the reference's behaviour):
``precision`` (``fp32`` | ``fp32_ir`` | ``bf16x3_ir`` | ``bf16_ir`` |
``per_shift_ir``, with ``far_shift_threshold``),
``refinement_steps`` (iterative refinement per shifted solve),
``refinement_passes`` (extra contour-filter passes over the Ritz block),
``k_batch`` (k points factored together), ``conjugate_pairs`` (solve the
mirrored shift with the conjugate-transposed factors), ``refinement_policy``
(``fixed`` | ``on_demand``), ``residual_target_ev``, ``autotune`` and
``resident_limit``.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace
from time import perf_counter
from typing import Any

import numpy as np
from scipy import sparse

from . import _native as N
from .data import EnergyInterval
from .errors import InputError, SolverError
from .sparse import NativeBundle, SparsePair


@dataclass(slots=True)
class CissResult:
    """Per-k result (ciss.py:18-25)."""

    energies_ev: np.ndarray
    maximum_residual_ev: float
    stage_seconds: dict[str, float]
    peak_gpu_memory_bytes: int | None
    warnings: tuple[str, ...]
    complete: bool
    subspace_rank: int = 0
    pass_rank: int = 0
    pass_width: int = 0
    # the selected Ritz vectors [n, m] (orbital order, S-normalised) when
    # requested with return_vectors=True
    eigenvectors: np.ndarray | None = None


@dataclass(slots=True)
class ShiftedResult:
    """Shifted-solve result (ciss.py:28-32): solutions (batch, n, L)."""

    solutions: np.ndarray
    stage_seconds: dict[str, float]
    peak_gpu_memory_bytes: int | None


@dataclass(frozen=True, slots=True)
class CissSettings:
    """Contour and subspace settings (ciss.py:35-62) plus device knobs."""

    integration_points: int = 32
    contour_batch_size: int = 1
    block_size: int = 32
    moment_size: int = 4
    ellipse_vertical_scale: float = 0.1
    endpoint_guard_ev: float = 1.0e-5
    rank_relative_tolerance: float = 1.0e-5
    spurious_relative_threshold: float = 1.0e-4
    random_seed: int = 20260727
    precision: str = "fp32"
    refinement_steps: int = 0
    refinement_passes: int = 0
    k_batch: int = 1
    conjugate_pairs: bool = False
    # "fixed": every shifted solve takes `refinement_steps` FP64-residual IR
    # steps.  "on_demand": solve without refinement first and re-solve, with
    # `refinement_steps`, only the k points whose FP64 eigenpair residual
    # exceeds `residual_target_ev`.
    refinement_policy: str = "fixed"
    residual_target_ev: float = 1.0e-4
    # time the step-group sizes of the tile LU on the bundle before the first
    # solve and keep the fastest (NativeBundle.autotune)
    autotune: bool = False
    # filter passes keep the factorisations of a batch resident when at most
    # this many (0: no limit) and they fit in device memory; otherwise every
    # pass factors each shift group again (streamed shifts)
    resident_limit: int = 0
    # precision "per_shift_ir": contour point j (angle theta_j) is factored
    # with one BF16 product per Schur update when |sin theta_j| >= this
    # threshold -- its distance rv |sin theta_j| from the real axis, where the
    # spectrum lies, bounds the conditioning of z_j S - H -- and with three
    # (BF16x3) otherwise; every solve is refined on FP64 residuals
    far_shift_threshold: float = 0.9

    @property
    def factorizations_per_launch(self) -> int:
        per_k = self.integration_points // 2 if self.conjugate_pairs else self.contour_batch_size
        return self.k_batch * per_k

    def __post_init__(self) -> None:
        if (
            self.integration_points < 1
            or self.contour_batch_size < 1
            or self.integration_points % self.contour_batch_size
            or self.block_size < 1
            or self.moment_size < 1
        ):
            raise InputError("CISS sizes must be positive and contour_batch_size must divide integration_points")
        if self.precision not in N.PRECISION:
            raise InputError(f"precision must be one of {sorted(N.PRECISION)}")
        if self.refinement_steps < 0 or self.refinement_passes < 0 or self.k_batch < 1 or self.resident_limit < 0:
            raise InputError("refinement counts must be >= 0 and k_batch >= 1")
        if self.conjugate_pairs and self.integration_points % 2:
            raise InputError("conjugate_pairs needs an even number of integration points")
        if self.refinement_policy not in ("fixed", "on_demand"):
            raise InputError("refinement_policy must be 'fixed' or 'on_demand'")
        if not self.residual_target_ev > 0.0:
            raise InputError("residual_target_ev must be positive")
        if not 0.0 <= self.far_shift_threshold <= 2.0:
            raise InputError("far_shift_threshold must lie in [0, 2]")

    @property
    def search_dimension(self) -> int:
        return self.block_size * self.moment_size


def contour_geometry(interval: EnergyInterval, settings: CissSettings):
    """z_j, t_j, w_j of the flat ellipse (ciss.py:81-104), FP64 -> complex64.
    The native library forms the same nodes itself; this copy serves callers
    that weight solutions of ``solve_shifted`` on the host."""
    lo = interval.lower_ev - 2.0 * settings.endpoint_guard_ev
    hi = interval.upper_ev + 2.0 * settings.endpoint_guard_ev
    c, rh = 0.5 * (lo + hi), 0.5 * (hi - lo)
    rv = rh * settings.ellipse_vertical_scale
    theta = 2.0 * np.pi * (np.arange(settings.integration_points) + 0.5) / settings.integration_points
    t = np.cos(theta) + 1j * settings.ellipse_vertical_scale * np.sin(theta)
    w = (rv * np.cos(theta) + 1j * rh * np.sin(theta)) / settings.integration_points
    return (c + rh * t).astype(np.complex64), t.astype(np.complex64), w.astype(np.complex64)


_PROBE_CACHE: dict[tuple[int, int, int], np.ndarray] = {}


def probe_block(n: int, settings: CissSettings, *, seed_offset: int = 0) -> np.ndarray:
    """Orthonormal complex probes from the fixed seed, generated exactly as the
    reference does (ciss.py:284-292): standard normals in float32, /sqrt(2),
    reduced QR.  The block is k-independent, so it is built once per (n, L).
    ``seed_offset = 1`` gives the independent block of 2L columns that
    augments the filter passes (a pass subspace of about 1.5-2x the in-window
    count keeps the subspace iteration contracting fast)."""
    width = settings.block_size if seed_offset == 0 else 2 * settings.block_size
    key = (int(n), width, settings.random_seed + seed_offset)
    if key not in _PROBE_CACHE:
        rng = np.random.default_rng(settings.random_seed + seed_offset)
        shape = (n, width)
        v = rng.standard_normal(shape, dtype=np.float32) + 1j * rng.standard_normal(shape, dtype=np.float32)
        v = np.asarray(v / np.asarray(math.sqrt(2.0), dtype=np.float32), dtype=np.complex64)
        q, _ = np.linalg.qr(v, mode="reduced")
        _PROBE_CACHE[key] = np.ascontiguousarray(q, dtype=np.complex64)
    return _PROBE_CACHE[key]


def _stage_dict(res: "N.BandResult", share: float) -> dict[str, float]:
    return {f"native_{name}": float(res.stage_seconds[i]) * share for i, name in enumerate(N.STAGE_NAMES)}


def _warnings(residual: float, saturated: bool) -> tuple[str, ...]:
    out = []
    if residual > 1.0e-3:
        out.append(f"maximum eigenpair residual is {residual:.3e} eV")
    if saturated:
        out.append("CISS search subspace is saturated; the returned interval spectrum may be incomplete")
    return tuple(out)


def band_solve(
    native: NativeBundle,
    settings: CissSettings,
    interval: EnergyInterval,
    *,
    kpoints: np.ndarray | None = None,
    h_data: np.ndarray | None = None,
    s_data: np.ndarray | None = None,
    intervals: np.ndarray | None = None,
    return_vectors: bool = False,
) -> tuple[list[CissResult], dict[str, Any]]:
    """Run the CISS Band path for a batch of k points (on-device assembly) or
    of explicit H/S value arrays on the bundle's pattern.  ``intervals``
    ([nk, 2], absolute eV) gives every k point its own window; otherwise all
    use ``interval``.

    With ``refinement_policy = "on_demand"`` the batch is first solved without
    iterative refinement; the k points whose FP64 eigenpair residual exceeds
    ``residual_target_ev`` are solved again with ``refinement_steps`` steps and
    their results replace the first ones (``info["refined"]`` lists them).
    ``return_vectors`` also returns each k point's selected Ritz vectors."""
    if settings.refinement_policy == "fixed" or settings.refinement_steps == 0:
        results, info = _band_call(native, settings, interval, kpoints, h_data, s_data, intervals, return_vectors)
        info["refined"] = []
        return results, info
    first = replace(settings, refinement_steps=0, refinement_policy="fixed")
    results, info = _band_call(native, first, interval, kpoints, h_data, s_data, intervals, return_vectors)
    redo = [i for i, r in enumerate(results) if r.maximum_residual_ev > settings.residual_target_ev]
    info["refined"] = redo
    if redo:
        fixed = replace(settings, refinement_policy="fixed")
        sel = np.asarray(redo)
        again, info2 = _band_call(
            native,
            fixed,
            interval,
            None if kpoints is None else np.asarray(kpoints, dtype=np.float64).reshape(-1, 3)[sel],
            None if h_data is None else np.asarray(h_data).reshape(len(results), -1)[sel],
            None if s_data is None else np.asarray(s_data).reshape(len(results), -1)[sel],
            None if intervals is None else np.asarray(intervals, dtype=np.float64).reshape(-1, 2)[sel],
            return_vectors,
        )
        for i, r in zip(redo, again):
            results[i] = r
        info["factorizations"] += info2["factorizations"]
        info["replaced_pivots"] += info2["replaced_pivots"]
        info["wall_seconds"] += info2["wall_seconds"]
        info["peak_device_bytes"] = max(info["peak_device_bytes"], info2["peak_device_bytes"])
        for name, v in info2["stage_seconds"].items():
            info["stage_seconds"][name] += v
    return results, info


def _band_call(native, settings, interval, kpoints, h_data, s_data, intervals, return_vectors):
    lib = N.load()
    if (kpoints is None) == (h_data is None):
        raise InputError("give either k points or explicit H/S values")
    if settings.search_dimension >= native.n:
        raise SolverError("CISS search dimension must be smaller than matrix dimension")
    started = perf_counter()
    probes = probe_block(native.n, settings)
    augmentation = probe_block(native.n, settings, seed_offset=1) if settings.refinement_passes else None
    setup = perf_counter() - started
    req = N.new(N.BandRequest)
    keep: list[np.ndarray] = [probes]
    if augmentation is not None:
        keep.append(augmentation)
        req.augmentation = N.ptr(augmentation, N.Complex64)
        req.augmentation_width = augmentation.shape[1]
    if kpoints is not None:
        k = np.ascontiguousarray(np.asarray(kpoints, dtype=np.float64).reshape(-1, 3))
        keep.append(k)
        nk = k.shape[0]
        req.k_fractional = N.ptr(k, C.c_double)
    else:
        h = np.ascontiguousarray(np.asarray(h_data, dtype=np.complex64).reshape(-1, native.nnz))
        s = np.ascontiguousarray(np.asarray(s_data, dtype=np.complex64).reshape(-1, native.nnz))
        keep += [h, s]
        nk = h.shape[0]
        req.h_data = N.ptr(h, N.Complex64)
        req.s_data = N.ptr(s, N.Complex64)
    if intervals is not None:
        iv = np.ascontiguousarray(np.asarray(intervals, dtype=np.float64).reshape(-1, 2))
        if iv.shape[0] != nk:
            raise InputError("one interval per k point is required")
        keep.append(iv)
        req.intervals = N.ptr(iv, C.c_double)
    cap = settings.search_dimension
    energies = np.zeros((nk, cap), dtype=np.float32)
    counts = np.zeros(nk, dtype=np.int32)
    resid = np.zeros(nk, dtype=np.float64)
    rank = np.zeros(nk, dtype=np.int32)
    sat = np.zeros(nk, dtype=np.int32)
    prank = np.zeros(nk, dtype=np.int32)
    pwidth = np.zeros(nk, dtype=np.int32)
    vecs = np.zeros((nk, native.n, cap), dtype=np.complex64) if return_vectors else None
    req.n_kpoints = nk
    req.lower_ev = interval.lower_ev
    req.upper_ev = interval.upper_ev
    req.integration_points = settings.integration_points
    req.contour_batch_size = settings.contour_batch_size
    req.block_size = settings.block_size
    req.moment_size = settings.moment_size
    req.ellipse_vertical_scale = settings.ellipse_vertical_scale
    req.endpoint_guard_ev = settings.endpoint_guard_ev
    req.rank_relative_tolerance = settings.rank_relative_tolerance
    req.spurious_relative_threshold = settings.spurious_relative_threshold
    req.probes = N.ptr(probes, N.Complex64)
    req.precision = N.PRECISION[settings.precision]
    req.refinement_steps = settings.refinement_steps
    req.refinement_passes = settings.refinement_passes
    req.k_batch = settings.k_batch
    req.conjugate_pairs = 1 if settings.conjugate_pairs else 0
    req.resident_limit = settings.resident_limit
    req.far_threshold = settings.far_shift_threshold
    req.energies_ev = N.ptr(energies, C.c_float)
    req.energy_capacity = cap
    req.energy_counts = N.ptr(counts, C.c_int32)
    req.maximum_residual_ev = N.ptr(resid, C.c_double)
    req.subspace_rank = N.ptr(rank, C.c_int32)
    req.saturated = N.ptr(sat, C.c_int32)
    req.pass_rank = N.ptr(prank, C.c_int32)
    req.pass_width = N.ptr(pwidth, C.c_int32)
    if vecs is not None:
        req.eigenvectors = N.ptr(vecs, N.Complex64)
    res = N.BandResult()
    rc = lib.aosx_band_solve_v1(native.handle, C.byref(req), C.byref(res))
    N.check(rc, res, "CISS band solve")
    wall = perf_counter() - started
    share = 1.0 / nk
    results = []
    for i in range(nk):
        stages = {"probe_setup": setup * share, **_stage_dict(res, share), "solver_total": wall * share}
        results.append(
            CissResult(
                energies_ev=energies[i, : counts[i]].copy(),
                maximum_residual_ev=float(resid[i]),
                stage_seconds=stages,
                peak_gpu_memory_bytes=int(res.peak_device_bytes),
                warnings=_warnings(float(resid[i]), bool(sat[i])),
                complete=not bool(sat[i]),
                subspace_rank=int(rank[i]),
                pass_rank=int(prank[i]),
                pass_width=int(pwidth[i]),
                eigenvectors=None if vecs is None else vecs[i, :, : counts[i]].copy(),
            )
        )
    info = {
        "replaced_pivots": int(res.replaced_pivots),
        "factorizations": int(res.factorizations),
        "stage_seconds": {name: float(res.stage_seconds[i]) for i, name in enumerate(N.STAGE_NAMES)},
        "peak_device_bytes": int(res.peak_device_bytes),
        "wall_seconds": wall,
    }
    return results, info


def eig_residual(
    native: NativeBundle,
    energies_ev: np.ndarray,
    vectors: np.ndarray,
    *,
    kpoints: np.ndarray | None = None,
    h_data: np.ndarray | None = None,
    s_data: np.ndarray | None = None,
) -> np.ndarray:
    """||H(k) v_j - e_j S(k) v_j||_2 in FP64 on the device, the residual of
    ciss.py:362-372 for caller-given pairs: ``energies_ev`` [nk, m],
    ``vectors`` [nk, n, m] (converted to complex128)."""
    lib = N.load()
    if (kpoints is None) == (h_data is None):
        raise InputError("give either k points or explicit H/S values")
    e = np.ascontiguousarray(np.asarray(energies_ev, dtype=np.float64))
    if e.ndim == 1:
        e = e[None, :]
    nk, m = e.shape
    v = np.ascontiguousarray(np.asarray(vectors, dtype=np.complex128).reshape(nk, native.n, m))
    out = np.zeros((nk, m), dtype=np.float64)
    req = N.new(N.ResidualRequest)
    keep = [e, v]
    if kpoints is not None:
        k = np.ascontiguousarray(np.asarray(kpoints, dtype=np.float64).reshape(nk, 3))
        keep.append(k)
        req.k_fractional = N.ptr(k, C.c_double)
    else:
        h = np.ascontiguousarray(np.asarray(h_data, dtype=np.complex64).reshape(nk, native.nnz))
        s = np.ascontiguousarray(np.asarray(s_data, dtype=np.complex64).reshape(nk, native.nnz))
        keep += [h, s]
        req.h_data = N.ptr(h, N.Complex64)
        req.s_data = N.ptr(s, N.Complex64)
    req.n_kpoints = nk
    req.count = m
    req.energies_ev = N.ptr(e, C.c_double)
    req.vectors = v.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
    req.residual_ev = N.ptr(out, C.c_double)
    res = N.StatusResult()
    rc = lib.aosx_eig_residual_v1(native.handle, C.byref(req), C.byref(res))
    del keep
    N.check(rc, res, "eigenpair residual")
    return out


def _require_shared_pattern(matrices: SparsePair) -> None:
    """(ciss.py:65-78)"""
    h, s = matrices.hamiltonian, matrices.overlap
    if not sparse.isspmatrix_csr(h) or not sparse.isspmatrix_csr(s):
        raise InputError("CISS requires CSR H and S")
    if not np.array_equal(h.indptr, s.indptr) or not np.array_equal(h.indices, s.indices):
        raise InputError("CISS requires one shared H/S CSR pattern")
    if h.dtype != np.complex64 or s.dtype != np.complex64:
        raise InputError("CISS requires complex64 H/S")


class _PatternSession:
    """Caches the native analysis of one CSR pattern, pinned by a token of
    the pattern arrays as the reference's shift session is (ciss.py:185-193)."""

    def __init__(self, device: int) -> None:
        self.device = device
        self.native: NativeBundle | None = None
        self.token: tuple | None = None
        self.perm: np.ndarray | None = None

    @staticmethod
    def _token(m: SparsePair) -> tuple:
        h = m.hamiltonian
        return (h.shape, h.indptr.__array_interface__["data"][0], h.indices.__array_interface__["data"][0],
                h.indptr.size, h.indices.size)

    def bind(self, m: SparsePair) -> NativeBundle:
        token = self._token(m)
        if self.native is not None and token != self.token:
            raise SolverError("shift session cannot change its CSR pattern")
        if self.native is None:
            h = m.hamiltonian
            if not h.has_sorted_indices:
                raise InputError("CSR columns must be sorted and duplicate-free")
            self.native = NativeBundle.from_csr(h.indptr, h.indices, h.shape[0], device=self.device)
            indptr, indices, _ = self.native.pattern()
            if not (np.array_equal(indptr, h.indptr) and np.array_equal(indices, h.indices)):
                raise InputError("CSR pattern must be Hermitian with a full diagonal")
            self.token = token
        return self.native

    def close(self) -> None:
        if self.native is not None:
            self.native.close()
        self.native = None
        self.token = None


class HipShiftSolver:
    """Shifted-solve plugin (the seam of ciss.py:180-250): (H - z_b S) X_b = RHS."""

    def __init__(self, *, device: int = 0, precision: str = "fp32", refinement_steps: int = 0) -> None:
        self._session = _PatternSession(device)
        self.precision = precision
        self.refinement_steps = refinement_steps
        self.last_relative_residual: float | None = None

    def solve_shifted(self, matrices: SparsePair, shifts_ev: np.ndarray, rhs: np.ndarray) -> ShiftedResult:
        native = self._session.bind(matrices)
        shifts = np.ascontiguousarray(np.asarray(shifts_ev, dtype=np.complex64).reshape(-1))
        rhs = np.ascontiguousarray(np.asarray(rhs, dtype=np.complex64))
        if rhs.ndim != 2 or rhs.shape[0] != native.n:
            raise InputError("RHS must have shape (n, 1..256)")
        h = np.ascontiguousarray(matrices.hamiltonian.data, dtype=np.complex64)
        s = np.ascontiguousarray(matrices.overlap.data, dtype=np.complex64)
        out = np.empty((shifts.size, native.n, rhs.shape[1]), dtype=np.complex64)
        req = N.new(N.ShiftRequest)
        req.batch_count = shifts.size
        req.rhs_count = rhs.shape[1]
        req.h_data = N.ptr(h, N.Complex64)
        req.s_data = N.ptr(s, N.Complex64)
        req.shifts = N.ptr(shifts, N.Complex64)
        req.rhs = N.ptr(rhs, N.Complex64)
        req.solution = N.ptr(out, N.Complex64)
        req.precision = N.PRECISION[self.precision]
        req.refinement_steps = self.refinement_steps
        res = N.ShiftResult()
        rc = N.load().aosx_shift_solve_v1(native.handle, C.byref(req), C.byref(res))
        N.check(rc, res, "shifted solve")
        self.last_relative_residual = float(res.relative_residual)
        stages = {
            "device_upload": res.upload_seconds,
            "numeric_factorization": res.factorization_seconds,
            "multi_rhs_solve": res.solve_seconds,
            "device_download": res.download_seconds,
            "native_total": res.total_seconds,
        }
        return ShiftedResult(out, stages, int(res.peak_device_bytes))

    def close(self) -> None:
        self._session.close()


@dataclass(slots=True)
class HipCissSolver:
    """Interval solver (ciss.py:253-391) running the whole CISS on the GPU."""

    settings: CissSettings = field(default_factory=CissSettings)
    device: int = 0
    _session: _PatternSession | None = None

    def __post_init__(self) -> None:
        self._session = _PatternSession(self.device)

    def solve_interval(self, matrices: SparsePair, interval: EnergyInterval) -> CissResult:
        _require_shared_pattern(matrices)
        if self.settings.search_dimension >= matrices.n_orbitals:
            raise SolverError("CISS search dimension must be smaller than matrix dimension")
        native = self._session.bind(matrices)
        results, _ = band_solve(
            native,
            self.settings,
            interval,
            h_data=matrices.hamiltonian.data[None, :],
            s_data=matrices.overlap.data[None, :],
        )
        return results[0]

    def close(self) -> None:
        if self._session is not None:
            self._session.close()
