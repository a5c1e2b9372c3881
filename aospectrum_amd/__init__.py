"""MI355X-native AOSpectrum Band path.

Bloch assembly of H(k), S(k), complex64 CISS with a project-written tile LU
(static pivoting, FP32 / BF16-split MFMA Schur updates, FP64-residual
refinement) and FP64 Rayleigh-Ritz, all on one GPU through libaosx_hip.so.
The public names mirror the reference package (src/aospectrum/__init__.py).
"""

from .bundle import load_bundle, write_bundle
from .data import (
    AtomicStructure,
    ElectronicFilling,
    EnergyInterval,
    LocalizedOperatorBlocks,
    LocalizedOperatorBundle,
    OrbitalBasisLayout,
)
from .errors import AOSpectrumError, InputError, SolverError

__all__ = [
    "AOSpectrumError",
    "AtomicStructure",
    "ElectronicFilling",
    "EnergyInterval",
    "InputError",
    "LocalizedOperatorBlocks",
    "LocalizedOperatorBundle",
    "OrbitalBasisLayout",
    "SolverError",
    "load_bundle",
    "write_bundle",
]

__version__ = "1.0.0rc2+mi355x"
