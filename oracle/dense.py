"""Oracle: FP64 dense generalized eigenvalues and inertia counts.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``window_spectrum`` is the correctness reference named for the Band path: the
eigenvalues of the FP64 dense pencil ``(H(k), S(k))`` strictly inside the
absolute solve interval (the selection rule of src/aospectrum/ciss.py:355-361).
``inertia_count`` counts the same eigenvalues by Sylvester's law of inertia
for sizes where a full eigensolve is too slow: for S positive definite the
number of eigenvalues of the pencil below sigma equals the number of negative
eigenvalues of ``H - sigma S``, read off the block-diagonal D of its
Bunch-Kaufman factorisation ``P (H - sigma S) P^T = L D L^H`` (LAPACK zhetrf,
blocked with the optimal workspace).  ``eigenpair_residuals`` is the host FP64
statement of the residual ``max_i ||H v_i - e_i S v_i||_2`` of
src/aospectrum/ciss.py:362-372.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sl
import scipy.linalg.lapack as lapack


def window_spectrum(H: np.ndarray, S: np.ndarray, lower: float, upper: float) -> np.ndarray:
    w = sl.eigh(H.astype(np.complex128), S.astype(np.complex128), eigvals_only=True)
    return w[(w > lower) & (w < upper)]


def negatives_of_factor(ldu: np.ndarray, ipiv: np.ndarray) -> int:
    """Negative eigenvalues of D from zhetrf's lower output (ipiv 1-based;
    ipiv[i] = ipiv[i+1] < 0 marks a 2 x 2 block at rows i, i+1)."""
    n = ldu.shape[0]
    diag = np.real(np.diagonal(ldu)).copy()
    sub = np.diagonal(ldu, -1).copy() if n > 1 else np.zeros(0, ldu.dtype)
    neg, i = 0, 0
    while i < n:
        if ipiv[i] < 0:
            a, c, b = diag[i], diag[i + 1], sub[i]
            det = a * c - abs(b) ** 2
            if det < 0.0:
                neg += 1
            elif det > 0.0:
                neg += 2 if a + c < 0.0 else 0
            else:
                neg += int(a < 0.0) + int(c < 0.0)
            i += 2
        else:
            neg += int(diag[i] < 0.0)
            i += 1
    return neg


def negatives(A: np.ndarray, overwrite: bool = False) -> int:
    """Negative eigenvalues of Hermitian complex128 ``A``.  With
    ``overwrite`` a C- or Fortran-contiguous ``A`` is factored in place: the
    Fortran view of a C array is conj(A), which has the same inertia."""
    if A.flags.c_contiguous and not A.flags.f_contiguous:
        F = A.T
    else:
        F = A
    if not overwrite:
        F = np.array(F, order="F", copy=True)
    n = F.shape[0]
    lw, info = lapack.zhetrf_lwork(n, lower=1)
    if info != 0:
        raise RuntimeError(f"zhetrf_lwork failed: {info}")
    ldu, ipiv, info = lapack.zhetrf(F, lower=1, lwork=int(np.real(lw)), overwrite_a=1)
    if info < 0:
        raise RuntimeError(f"zhetrf argument {-info} invalid")
    # info > 0: an exactly zero pivot -- sigma is an eigenvalue; the count is
    # still defined with that zero counted as non-negative
    return negatives_of_factor(ldu, ipiv)


def negatives_below(H: np.ndarray, S: np.ndarray, sigma: float) -> int:
    """Eigenvalues of the pencil (H, S) strictly below sigma (S > 0)."""
    A = H.astype(np.complex128, copy=True)
    A -= sigma * S
    return negatives(A, overwrite=True)


def inertia_count(H: np.ndarray, S: np.ndarray, lower: float, upper: float) -> int:
    return negatives_below(H, S, upper) - negatives_below(H, S, lower)


def eigenpair_residuals(H, S, vectors: np.ndarray, energies: np.ndarray) -> np.ndarray:
    """||H v_i - e_i S v_i||_2 per column, in FP64 (H, S dense or sparse)."""
    V = np.asarray(vectors, dtype=np.complex128)
    e = np.asarray(energies, dtype=np.float64)
    R = (H @ V) - (S @ V) * e[None, :]
    return np.sqrt(np.sum(np.abs(R) ** 2, axis=0))
