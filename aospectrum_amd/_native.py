"""ctypes binding of libaosx_hip.so, the C-ABI declared in include/aosx.h.

The structs below restate the header field by field; ``load()`` compares their
sizes with ``aosx_struct_sizes_v1`` so a layout drift fails at import time
instead of corrupting a call.  The library is built in-tree
(``aospectrum_amd/csrc/libaosx_hip.so``); there is no fallback: a missing
library raises ``SolverError``.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from .errors import InputError, SolverError

ABI_VERSION = 1
ERROR_CAPACITY = 512
STAGE_COUNT = 12
STAGE_NAMES = (
    "assembly",
    "probe_setup",
    "numeric_factorization",
    "multi_rhs_solve",
    "refinement",
    "moments",
    "subspace_compression",
    "rayleigh_ritz",
    "residual",
    "filter_passes",
    "host_sync",
    "native_total",
)

PRECISION = {"fp32": 0, "fp32_ir": 1, "bf16x3_ir": 2, "bf16_ir": 3, "per_shift_ir": 4}

# status codes that describe caller input rather than a device failure
_INPUT_STATUSES = {10, 11, 12, 13, 14, 30, 31, 32, 33, 34, 35, 36, 50, 51, 52, 64}

_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_F64P = C.POINTER(C.c_double)
_F32P = C.POINTER(C.c_float)


class Complex64(C.Structure):
    _fields_ = [("real", C.c_float), ("imag", C.c_float)]


_C64P = C.POINTER(Complex64)


class BundleRequest(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("n_orbitals", C.c_int64),
        ("n_atoms", C.c_int64),
        ("n_blocks", C.c_int64),
        ("atom_offsets", _I64P),
        ("row_atom", _I32P),
        ("col_atom", _I32P),
        ("lattice_shift", _I32P),
        ("block_offsets", _I64P),
        ("value_kind", C.c_int32),
        ("device", C.c_int32),
        ("h_values", _P),
        ("s_values", _P),
        ("tile_size", C.c_int32),
        ("ordering", C.c_int32),
        ("atom_positions", _F64P),
    ]


class BundleResult(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("status", C.c_int32),
        ("hip_status", C.c_int32),
        ("n", C.c_int64),
        ("nnz", C.c_int64),
        ("n_phase_shifts", C.c_int64),
        ("n_diagonal_tiles", C.c_int64),
        ("n_factor_tiles", C.c_int64),
        ("atom_fill_entries", C.c_int64),
        ("tile_size", C.c_int64),
        ("factor_flops", C.c_double),
        ("schur_flops", C.c_double),
        ("factor_bytes", C.c_uint64),
        ("resident_device_bytes", C.c_uint64),
        ("pattern_seconds", C.c_double),
        ("ordering_seconds", C.c_double),
        ("symbolic_seconds", C.c_double),
        ("upload_seconds", C.c_double),
        ("error_message", C.c_char * ERROR_CAPACITY),
    ]


class StatusResult(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("status", C.c_int32),
        ("hip_status", C.c_int32),
        ("seconds", C.c_double),
        ("error_message", C.c_char * ERROR_CAPACITY),
    ]


class AssembleRequest(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("n_kpoints", C.c_int64),
        ("k_fractional", _F64P),
        ("h_data", _C64P),
        ("s_data", _C64P),
    ]


class ShiftRequest(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("batch_count", C.c_int32),
        ("rhs_count", C.c_int32),
        ("h_data", _C64P),
        ("s_data", _C64P),
        ("shifts", _C64P),
        ("rhs", _C64P),
        ("solution", _C64P),
        ("precision", C.c_int32),
        ("refinement_steps", C.c_int32),
    ]


class ShiftResult(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("status", C.c_int32),
        ("hip_status", C.c_int32),
        ("replaced_pivots", C.c_int64),
        ("refinement_steps", C.c_int32),
        ("reserved", C.c_int32),
        ("relative_residual", C.c_double),
        ("upload_seconds", C.c_double),
        ("factorization_seconds", C.c_double),
        ("solve_seconds", C.c_double),
        ("download_seconds", C.c_double),
        ("total_seconds", C.c_double),
        ("peak_device_bytes", C.c_uint64),
        ("error_message", C.c_char * ERROR_CAPACITY),
    ]


class BandRequest(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("n_kpoints", C.c_int64),
        ("k_fractional", _F64P),
        ("h_data", _C64P),
        ("s_data", _C64P),
        ("lower_ev", C.c_double),
        ("upper_ev", C.c_double),
        ("intervals", _F64P),
        ("integration_points", C.c_int32),
        ("contour_batch_size", C.c_int32),
        ("block_size", C.c_int32),
        ("moment_size", C.c_int32),
        ("ellipse_vertical_scale", C.c_double),
        ("endpoint_guard_ev", C.c_double),
        ("rank_relative_tolerance", C.c_double),
        ("spurious_relative_threshold", C.c_double),
        ("probes", _C64P),
        ("augmentation", _C64P),
        ("augmentation_width", C.c_int32),
        ("reserved_aug", C.c_int32),
        ("precision", C.c_int32),
        ("refinement_steps", C.c_int32),
        ("refinement_passes", C.c_int32),
        ("k_batch", C.c_int32),
        ("conjugate_pairs", C.c_int32),
        ("resident_limit", C.c_int32),
        ("energies_ev", _F32P),
        ("energy_capacity", C.c_int32),
        ("energy_counts", _I32P),
        ("maximum_residual_ev", _F64P),
        ("subspace_rank", _I32P),
        ("saturated", _I32P),
        ("eigenvectors", _C64P),
        ("pass_rank", _I32P),
        ("pass_width", _I32P),
        ("far_threshold", C.c_double),
    ]


class BandResult(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("status", C.c_int32),
        ("hip_status", C.c_int32),
        ("replaced_pivots", C.c_int64),
        ("factorizations", C.c_int64),
        ("stage_seconds", C.c_double * STAGE_COUNT),
        ("peak_device_bytes", C.c_uint64),
        ("error_message", C.c_char * ERROR_CAPACITY),
    ]


class KernelStats(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("schur_launches", C.c_int64),
        ("schur_flops", C.c_double),
        ("schur_ms", C.c_double),
        ("panel_launches", C.c_int64),
        ("panel_ms", C.c_double),
        ("getrf_launches", C.c_int64),
        ("getrf_ms", C.c_double),
        ("solve_launches", C.c_int64),
        ("solve_ms", C.c_double),
        ("solve_bytes", C.c_double),
        ("assemble_launches", C.c_int64),
        ("assemble_ms", C.c_double),
        ("assemble_bytes", C.c_double),
        ("inner_launches", C.c_int64),
        ("inner_flops", C.c_double),
        ("inner_ms", C.c_double),
        ("schur_bytes", C.c_double),
    ]


AUTOTUNE_MAX_CANDIDATES = 16


class AutotuneRequest(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("matrices", C.c_int32),
        ("precision", C.c_int32),
        ("n_candidates", C.c_int32),
        ("repeats", C.c_int32),
        ("candidates", _I32P),
    ]


class AutotuneResult(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("status", C.c_int32),
        ("hip_status", C.c_int32),
        ("chosen_group", C.c_int32),
        ("n_candidates", C.c_int32),
        ("candidates", C.c_int32 * AUTOTUNE_MAX_CANDIDATES),
        ("factor_ms", C.c_double * AUTOTUNE_MAX_CANDIDATES),
        ("error_message", C.c_char * ERROR_CAPACITY),
    ]


class ResidualRequest(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("n_kpoints", C.c_int64),
        ("k_fractional", _F64P),
        ("h_data", _C64P),
        ("s_data", _C64P),
        ("count", C.c_int32),
        ("reserved", C.c_int32),
        ("energies_ev", _F64P),
        ("vectors", _F64P),
        ("residual_ev", _F64P),
    ]


class AssembleNoallocRequest(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("struct_size", C.c_uint32),
        ("n_kpoints", C.c_int64),
        ("k_fractional", _F64P),
        ("h_out", C.c_void_p),
        ("s_out", C.c_void_p),
    ]


_STRUCTS = (
    BundleRequest,
    BundleResult,
    StatusResult,
    AssembleRequest,
    ShiftRequest,
    ShiftResult,
    BandRequest,
    BandResult,
    KernelStats,
    AutotuneRequest,
    AutotuneResult,
    ResidualRequest,
    AssembleNoallocRequest,
)

LIBRARY_PATH = Path(__file__).resolve().parent / "csrc" / "libaosx_hip.so"
_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load the in-tree library once and verify the struct layouts."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("AOSX_LIBRARY", LIBRARY_PATH))
    if not path.is_file():
        raise SolverError(
            f"the HIP Band-path library is not built ({path}); run __graft_entry__.build() "
            "or make -C aospectrum_amd/csrc"
        )
    try:
        lib = C.CDLL(str(path))
    except OSError as exc:
        raise SolverError(f"cannot load {path}: {exc}") from exc
    lib.aosx_struct_sizes_v1.argtypes = [_I64P, C.c_int32]
    lib.aosx_struct_sizes_v1.restype = C.c_int
    sizes = (C.c_int64 * 16)()
    count = lib.aosx_struct_sizes_v1(sizes, 16)
    if count != len(_STRUCTS):
        raise SolverError("libaosx_hip.so exports an unexpected struct table")
    for struct, size in zip(_STRUCTS, sizes[:count]):
        if C.sizeof(struct) != size:
            raise SolverError(
                f"{struct.__name__} layout differs from libaosx_hip.so ({C.sizeof(struct)} != {size})"
            )
    sig = {
        "aosx_bundle_create_v1": [C.POINTER(BundleRequest), C.POINTER(BundleResult), C.POINTER(_P)],
        "aosx_bundle_pattern_v1": [_P, _I32P, _I32P, _I32P, C.POINTER(StatusResult)],
        "aosx_bundle_destroy_v1": [_P],
        "aosx_assemble_v1": [_P, C.POINTER(AssembleRequest), C.POINTER(StatusResult)],
        "aosx_shift_solve_v1": [_P, C.POINTER(ShiftRequest), C.POINTER(ShiftResult)],
        "aosx_band_solve_v1": [_P, C.POINTER(BandRequest), C.POINTER(BandResult)],
        "aosx_kernel_stats_read_v1": [_P, C.c_int32, C.POINTER(KernelStats)],
        "aosx_bundle_trim_v1": [_P, C.POINTER(C.c_uint64)],
        "aosx_bundle_autotune_v1": [_P, C.POINTER(AutotuneRequest), C.POINTER(AutotuneResult)],
        "aosx_eig_residual_v1": [_P, C.POINTER(ResidualRequest), C.POINTER(StatusResult)],
        "aosx_assemble_noalloc_v1": [_P, C.POINTER(AssembleNoallocRequest), C.POINTER(StatusResult)],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = lib
    return lib


EXPORTED_SYMBOLS = (
    "aosx_bundle_create_v1",
    "aosx_bundle_pattern_v1",
    "aosx_bundle_destroy_v1",
    "aosx_assemble_v1",
    "aosx_shift_solve_v1",
    "aosx_band_solve_v1",
    "aosx_kernel_stats_read_v1",
    "aosx_struct_sizes_v1",
    "aosx_bundle_trim_v1",
    "aosx_bundle_autotune_v1",
    "aosx_eig_residual_v1",
    "aosx_assemble_noalloc_v1",
)


def new(struct):
    """A zeroed request struct with its ABI header filled in."""
    value = struct()
    value.abi_version = ABI_VERSION
    value.struct_size = C.sizeof(struct)
    return value


def check(rc: int, result, what: str) -> None:
    """Turn a failed call into InputError (caller data) or SolverError."""
    if rc == 0 and result.status == 0:
        return
    message = bytes(result.error_message).split(b"\0", 1)[0].decode("utf-8", "replace")
    text = f"{what} failed: status={result.status}, message={message}"
    if result.status in _INPUT_STATUSES:
        raise InputError(text)
    raise SolverError(text)


def ptr(array, ctype):
    """ctypes pointer to the data of a C-contiguous NumPy array."""
    if not array.flags["C_CONTIGUOUS"]:
        raise InputError("native arrays must be C-contiguous")
    return array.ctypes.data_as(C.POINTER(ctype))
