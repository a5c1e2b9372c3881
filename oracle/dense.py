"""Oracle: FP64 dense generalized eigenvalues and inertia counts.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``window_spectrum`` is the correctness reference named for the Band path: the
eigenvalues of the FP64 dense pencil ``(H(k), S(k))`` strictly inside the
absolute solve interval (the selection rule of src/aospectrum/ciss.py:355-361).
``inertia_count`` counts the same eigenvalues by Sylvester's law of inertia
(LDL^T of ``H - sigma S`` at both interval ends) for sizes where a full
eigensolve is too slow.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sl


def window_spectrum(H: np.ndarray, S: np.ndarray, lower: float, upper: float) -> np.ndarray:
    w = sl.eigh(H.astype(np.complex128), S.astype(np.complex128), eigvals_only=True)
    return w[(w > lower) & (w < upper)]


def _negatives(A: np.ndarray) -> int:
    """Negative eigenvalues of Hermitian ``A`` from its Bunch-Kaufman LDL^T."""
    _, d, _ = sl.ldl(A, lower=True, hermitian=True)
    n = d.shape[0]
    sub = np.abs(np.diag(d, -1)) > 0.0
    diag = np.real(np.diag(d))
    neg, i = 0, 0
    while i < n:
        if i + 1 < n and sub[i]:
            neg += int(np.sum(np.linalg.eigvalsh(d[i:i + 2, i:i + 2]) < 0.0))
            i += 2
        else:
            neg += int(diag[i] < 0.0)
            i += 1
    return neg


def inertia_count(H: np.ndarray, S: np.ndarray, lower: float, upper: float) -> int:
    H = H.astype(np.complex128)
    S = S.astype(np.complex128)
    return _negatives(H - upper * S) - _negatives(H - lower * S)
