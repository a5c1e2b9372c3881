"""Oracle: complex64 CISS interval solver with a SciPy SuperLU shifted solver.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates src/aospectrum/ciss.py step by step:
* ``contour``          -- ``_contour_geometry`` (ciss.py:81-104)
* ``probes``           -- probe block of ``solve_interval`` (ciss.py:284-295)
* ``solve_interval``   -- moment loop (ciss.py:305-333), ``_compress_moments``
  (ciss.py:107-126), ``_rayleigh_ritz`` (ciss.py:129-150),
  ``_nonspurious_mask`` (ciss.py:153-177), selection and residual
  (ciss.py:355-372)
* shifted systems use the sign convention of bindings.cpp:145-150:
  ``(H - z S) X = S V``.
The shifted solver is ``scipy.sparse.linalg.splu`` (complex64, COLAMD) or a
dense complex128 solve, chosen by the caller.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sl
import scipy.sparse as sp
import scipy.sparse.linalg as spla


@dataclass(frozen=True)
class Settings:
    integration_points: int = 32
    contour_batch_size: int = 1
    block_size: int = 32
    moment_size: int = 4
    ellipse_vertical_scale: float = 0.1
    endpoint_guard_ev: float = 1.0e-5
    rank_relative_tolerance: float = 1.0e-5
    spurious_relative_threshold: float = 1.0e-4
    random_seed: int = 20260727

    @property
    def search_dimension(self) -> int:
        return self.block_size * self.moment_size


@dataclass
class Result:
    energies_ev: np.ndarray
    maximum_residual_ev: float
    rank: int
    saturated: bool
    ritz_values: np.ndarray = field(default_factory=lambda: np.zeros(0))
    singular_values: np.ndarray = field(default_factory=lambda: np.zeros(0))


class SolverFailure(RuntimeError):
    pass


def contour(lower: float, upper: float, st: Settings):
    """z_j, t_j, w_j of the flat ellipse (ciss.py:81-104), float32 arithmetic."""
    lo = lower - 2.0 * st.endpoint_guard_ev
    hi = upper + 2.0 * st.endpoint_guard_ev
    center = 0.5 * (lo + hi)
    rh = 0.5 * (hi - lo)
    rv = rh * st.ellipse_vertical_scale
    idx = np.arange(st.integration_points, dtype=np.float32)
    theta = 2.0 * np.pi * (idx + 0.5) / st.integration_points
    t = np.cos(theta) + 1j * st.ellipse_vertical_scale * np.sin(theta)
    z = center + rh * t
    w = (rv * np.cos(theta) + 1j * rh * np.sin(theta)) / st.integration_points
    return z.astype(np.complex64), t.astype(np.complex64), w.astype(np.complex64)


def probes(n: int, st: Settings) -> np.ndarray:
    """Orthonormal complex probe block from the fixed seed (ciss.py:284-292)."""
    rng = np.random.default_rng(st.random_seed)
    shape = (n, st.block_size)
    v = (rng.standard_normal(shape, dtype=np.float32) + 1j * rng.standard_normal(shape, dtype=np.float32))
    v = np.asarray(v / np.asarray(math.sqrt(2.0), dtype=np.float32), dtype=np.complex64)
    q, _ = np.linalg.qr(v, mode="reduced")
    return np.asarray(q, dtype=np.complex64)


def superlu_solver(H: sp.csr_matrix, S: sp.csr_matrix, z: complex, rhs: np.ndarray) -> np.ndarray:
    a = (H - np.complex64(z) * S).astype(np.complex64).tocsc()
    lu = spla.splu(a, permc_spec="COLAMD")
    return np.asarray(lu.solve(rhs), dtype=np.complex64)


def lapack_solver(H, S, z, rhs):
    """Dense complex64 LAPACK getrf/getrs of (H - z S) -- the CPU-baseline
    shifted solver for near-dense fill (cheaper than SuperLU there)."""
    a = H.toarray() if hasattr(H, "toarray") else np.asarray(H)
    b = S.toarray() if hasattr(S, "toarray") else np.asarray(S)
    a = np.asarray(a, dtype=np.complex64) - np.complex64(z) * np.asarray(b, dtype=np.complex64)
    lu = sl.lu_factor(a, overwrite_a=True, check_finite=False)
    return np.asarray(sl.lu_solve(lu, rhs.astype(np.complex64), check_finite=False), dtype=np.complex64)


def dense_solver(H, S, z, rhs):
    a = (H.toarray().astype(np.complex128) - complex(z) * S.toarray().astype(np.complex128))
    return np.linalg.solve(a, rhs.astype(np.complex128)).astype(np.complex64)


def moments_from_solutions(solutions, t, w, st: Settings, n: int) -> np.ndarray:
    """M[:, mL:(m+1)L] += -w_j t_j^m X_j (ciss.py:324-333), complex64."""
    L = st.block_size
    M = np.zeros((n, st.search_dimension), dtype=np.complex64)
    for j, x in enumerate(solutions):
        for m in range(st.moment_size):
            M[:, m * L:(m + 1) * L] += -w[j] * t[j] ** m * x
    return M


def compress(S, M, tol):
    """Rank-revealing S-orthonormalisation (ciss.py:107-126)."""
    G = M.conj().T @ (S @ M)
    G = 0.5 * (G + G.conj().T)
    lam, U = np.linalg.eigh(G)
    scale = float(np.max(np.abs(lam), initial=0.0))
    if scale == 0.0 or not math.isfinite(scale):
        raise SolverFailure("CISS contour moments have zero numerical rank")
    keep = lam > scale * tol
    if not np.any(keep):
        raise SolverFailure("CISS contour moments have no positive stable subspace")
    pos = lam[keep]
    Q = np.asarray(M @ (U[:, keep] / np.sqrt(pos)[None, :]), dtype=np.complex64)
    return Q, np.sqrt(pos)


def rayleigh_ritz(H, S, Q):
    """Projected generalized eigenproblem (ciss.py:129-150)."""
    Hh = Q.conj().T @ (H @ Q)
    Sh = Q.conj().T @ (S @ Q)
    Hh = 0.5 * (Hh + Hh.conj().T)
    Sh = 0.5 * (Sh + Sh.conj().T)
    try:
        E, C = sl.eigh(Hh, Sh, check_finite=False, driver="gvd")
    except sl.LinAlgError as exc:
        raise SolverFailure("CISS reduced overlap is not positive definite") from exc
    return np.asarray(E, np.float32), np.asarray(Q @ C, np.complex64), np.asarray(C, np.complex64)


def nonspurious(C, sigma, threshold):
    """Sigma-weighted ghost filter (ciss.py:153-177)."""
    a = np.abs(C) ** 2
    num = a.sum(axis=0)
    den = (a / sigma[:, None]).sum(axis=0)
    score = np.divide(num, den, out=np.zeros(num.shape, dtype=num.dtype), where=den > 0.0)
    top = float(np.max(score, initial=0.0))
    if top == 0.0 or not math.isfinite(top):
        raise SolverFailure("CISS ghost filter produced no positive finite score")
    return np.isfinite(score) & (score >= threshold * top)


def solve_interval(H: sp.csr_matrix, S: sp.csr_matrix, lower: float, upper: float,
                   st: Settings = Settings(), shifted=superlu_solver, V=None) -> Result:
    """One k point of the reference CISS (ciss.py:266-391)."""
    n = H.shape[0]
    if st.search_dimension >= n:
        raise SolverFailure("CISS search dimension must be smaller than matrix dimension")
    z, t, w = contour(lower, upper, st)
    if V is None:
        V = probes(n, st)
    rhs = np.asarray(S @ V, dtype=np.complex64)
    sols = [shifted(H, S, zj, rhs) for zj in z]
    M = moments_from_solutions(sols, t, w, st, n)
    Q, sigma = compress(S, M, st.rank_relative_tolerance)
    saturated = Q.shape[1] == st.search_dimension
    E, X, C = rayleigh_ritz(H, S, Q)
    keep = nonspurious(C, sigma, st.spurious_relative_threshold) & (E > lower) & (E < upper)
    sel_E = E[keep]
    sel_X = X[:, keep]
    if sel_E.size:
        res = H @ sel_X - (S @ sel_X) * sel_E[None, :]
        max_res = float(np.max(np.linalg.norm(res, axis=0)))
    else:
        max_res = 0.0
    return Result(sel_E, max_res, int(Q.shape[1]), bool(saturated), E, sigma)
