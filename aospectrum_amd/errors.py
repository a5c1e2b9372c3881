"""Exception set shared by every layer of the Band path.

Mirrors the reference's user-facing hierarchy (src/aospectrum/errors.py:4-13):
an ``InputError`` for malformed bundles, k paths, settings and C-ABI requests,
and a ``SolverError`` for numerical or device failures.  The CLI turns any
``AOSpectrumError`` into exit status 2 (src/aospectrum/cli.py:608-611).
"""


class AOSpectrumError(RuntimeError):
    """Base class of every error this package raises on purpose."""


class InputError(AOSpectrumError):
    """The caller supplied data or settings the Band path cannot accept."""


class SolverError(AOSpectrumError):
    """A numerical stage or the HIP runtime failed while solving."""
