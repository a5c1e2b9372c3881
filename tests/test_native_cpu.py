"""The C-ABI library on the CPU: it loads, exports every declared symbol,
its struct layouts match the bindings, the host symbolic analysis matches the
oracle, and every validation status can be forced from the host.  No compute
call reaches a GPU here (bundles are created with device = -1)."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from aospectrum_amd import _native as N
from aospectrum_amd import synth
from aospectrum_amd.errors import InputError, SolverError
from aospectrum_amd.sparse import NativeBundle
from oracle import bloch

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    lib = N.load()
    header = (ROOT / "include" / "aosx.h").read_text()
    declared = set(re.findall(r"^int (aosx_\w+)\(", header, flags=re.M))
    assert declared == set(N.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name


def test_struct_layouts_match():
    lib = N.load()
    sizes = (C.c_int64 * 16)()
    count = lib.aosx_struct_sizes_v1(sizes, 16)
    assert count == len(N._STRUCTS)
    assert [C.sizeof(s) for s in N._STRUCTS] == list(sizes[:count])


@pytest.mark.parametrize("ordering", ["minimum_degree", "nested_dissection", "natural"])
def test_host_pattern_matches_oracle(small2d, small2d_pattern, ordering):
    nb = NativeBundle(small2d, device=-1, ordering=ordering)
    indptr, indices, perm = nb.pattern()
    assert np.array_equal(indptr, small2d_pattern.indptr)
    assert np.array_equal(indices, small2d_pattern.indices)
    assert np.array_equal(np.sort(perm), np.arange(small2d.n_orbitals))
    # atoms stay contiguous in the factor order (13 orbitals each)
    atoms = perm.reshape(-1, 13) // 13
    assert np.all(atoms == atoms[:, :1])
    info = nb.info
    assert info["nnz"] == small2d_pattern.nnz
    assert info["n_phase_shifts"] == small2d_pattern.shifts.shape[0]
    assert info["n_diagonal_tiles"] == 13
    nb.close()


def test_host_pattern_with_duplicate_shifts(tiny3d):
    pat = bloch.Pattern.build(tiny3d.operators, tiny3d.basis.atom_offsets, tiny3d.n_orbitals)
    nb = NativeBundle(tiny3d, device=-1)
    indptr, indices, _ = nb.pattern()
    assert np.array_equal(indptr, pat.indptr) and np.array_equal(indices, pat.indices)
    nb.close()


def test_nested_dissection_reduces_tile_work():
    b = synth.generate(synth.SPECS["stress2d"])
    md = NativeBundle(b, device=-1, ordering="minimum_degree").info
    nd = NativeBundle(b, device=-1, ordering="nested_dissection").info
    assert nd["factor_flops"] < 0.6 * md["factor_flops"]
    assert nd["n_factor_tiles"] < md["n_factor_tiles"]


def _request(bundle):
    ops = bundle.operators
    arrays = dict(
        off=np.array(bundle.basis.atom_offsets, dtype=np.int64),
        ra=np.array(ops.row_atom, dtype=np.int32),
        ca=np.array(ops.col_atom, dtype=np.int32),
        ls=np.array(ops.lattice_shift, dtype=np.int32),
        bo=np.array(ops.block_offsets, dtype=np.int64),
        h=np.array(ops.h_values, dtype=np.float32),
        s=np.array(ops.s_values, dtype=np.float32),
    )
    req = N.new(N.BundleRequest)
    req.n_orbitals = bundle.n_orbitals
    req.n_atoms = bundle.basis.n_atoms
    req.n_blocks = ops.n_blocks
    req.atom_offsets = N.ptr(arrays["off"], C.c_int64)
    req.row_atom = N.ptr(arrays["ra"], C.c_int32)
    req.col_atom = N.ptr(arrays["ca"], C.c_int32)
    req.lattice_shift = N.ptr(arrays["ls"], C.c_int32)
    req.block_offsets = N.ptr(arrays["bo"], C.c_int64)
    req.value_kind = 0
    req.device = -1
    req.h_values = arrays["h"].ctypes.data
    req.s_values = arrays["s"].ctypes.data
    req.ordering = 2
    return req, arrays


def _create(req):
    res = N.BundleResult()
    handle = C.c_void_p()
    rc = N.load().aosx_bundle_create_v1(C.byref(req), C.byref(res), C.byref(handle))
    if handle.value:
        N.load().aosx_bundle_destroy_v1(handle)
    return rc, res


def test_status_abi_mismatch(small2d):
    req, keep = _request(small2d)
    req.struct_size -= 8
    rc, res = _create(req)
    assert rc == 1 and res.status == 10 and b"ABI" in bytes(res.error_message)


def test_status_null_pointer(small2d):
    req, keep = _request(small2d)
    req.row_atom = None
    rc, res = _create(req)
    assert rc == 1 and res.status == 13


def test_status_non_finite(small2d):
    req, keep = _request(small2d)
    keep["h"][5] = np.nan
    rc, res = _create(req)
    assert rc == 1 and res.status == 35


def test_status_non_hermitian_topology(small2d):
    req, keep = _request(small2d)
    # redirect one off-site block to another atom: its mirror no longer exists
    ops = small2d.operators
    b = int(np.flatnonzero(ops.row_atom != ops.col_atom)[0])
    keep["ca"][b] = (keep["ca"][b] + 7) % small2d.basis.n_atoms
    rc, res = _create(req)
    assert rc == 1 and res.status in (50, 51)


def test_status_invalid_settings(small2d):
    req, keep = _request(small2d)
    req.tile_size = 48
    rc, res = _create(req)
    assert rc == 1 and res.status == 52


def test_compute_calls_need_a_device(small2d):
    nb = NativeBundle(small2d, device=-1)
    with pytest.raises(SolverError, match="status=53"):
        nb.assemble(np.zeros((1, 3)))
    nb.close()


def test_shift_solve_batch_limits_checked_before_device(small2d):
    nb = NativeBundle(small2d, device=-1)
    lib = N.load()
    for batch, rhs in ((65, 1), (1, 257), (0, 4)):
        req = N.new(N.ShiftRequest)
        req.batch_count = batch
        req.rhs_count = rhs
        res = N.ShiftResult()
        rc = lib.aosx_shift_solve_v1(nb.handle, C.byref(req), C.byref(res))
        assert rc == 1 and res.status == 12
    req = N.new(N.ShiftRequest)
    req.batch_count, req.rhs_count = 1, 4
    res = N.ShiftResult()
    assert lib.aosx_shift_solve_v1(nb.handle, C.byref(req), C.byref(res)) == 1 and res.status == 13
    nb.close()


def test_band_solve_settings_validation(small2d):
    from aospectrum_amd.ciss import CissSettings

    with pytest.raises(InputError):
        CissSettings(contour_batch_size=5)
    with pytest.raises(InputError):
        CissSettings(precision="fp8")
    with pytest.raises(InputError):
        CissSettings(integration_points=31, conjugate_pairs=True)
    with pytest.raises(InputError):
        CissSettings(precision="per_shift_ir", far_shift_threshold=2.5)
    assert CissSettings(precision="per_shift_ir", far_shift_threshold=0.7).far_shift_threshold == 0.7

