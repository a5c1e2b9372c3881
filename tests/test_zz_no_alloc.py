"""The allocation-free assembly entry point against aosx_assemble_v1 (needs an
MI355X): the same H(k), S(k) bit for bit, on two calls in a row with the
second smaller, written into caller-owned device buffers.

The test allocates those buffers with hipMalloc from the HIP runtime the
library itself links (/opt/rocm), not through torch: the torch wheel carries
its own HIP runtime, which does not initialise once the library's runtime has
used the device in this process."""

import ctypes as C
import sys

import numpy as np
import pytest

from aospectrum_amd import _native as N
from aospectrum_amd import synth
from aospectrum_amd.sparse import NativeBundle

pytestmark = pytest.mark.gpu


def say(*parts):
    print("[gpu-test]", *parts, flush=True)
    sys.stdout.flush()


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libaosx_hip.so links."""

    def __init__(self) -> None:
        N.load()
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), nbytes) == 0
        return p.value

    def upload(self, dst: int, host: np.ndarray) -> None:
        assert self.lib.hipMemcpy(dst, host.ctypes.data, host.nbytes, 1) == 0  # host to device

    def download(self, host: np.ndarray, src: int) -> None:
        assert self.lib.hipMemcpy(host.ctypes.data, src, host.nbytes, 2) == 0  # device to host

    def free(self, p: int) -> None:
        assert self.lib.hipFree(p) == 0


def _noalloc(nb, kpoints, h_out, s_out):
    k = np.ascontiguousarray(kpoints, dtype=np.float64)
    req = N.new(N.AssembleNoallocRequest)
    req.n_kpoints = k.shape[0]
    req.k_fractional = N.ptr(k, C.c_double)
    req.h_out = h_out
    req.s_out = s_out
    res = N.StatusResult()
    rc = N.load().aosx_assemble_noalloc_v1(nb.handle, C.byref(req), C.byref(res))
    return rc, res


@pytest.mark.parametrize("name", ["small2d", "tiny3d"])
def test_noalloc_assembly_matches_assemble_v1(name, tiny3d):
    bundle = tiny3d if name == "tiny3d" else synth.generate(synth.SPECS[name])
    say(f"no-alloc assembly {name} n={bundle.n_orbitals}")
    nb = NativeBundle(bundle, device=0)
    hip = _Hip()
    cap = 5 * nb.nnz * 8  # bytes of the larger call's outputs
    bufs = [hip.alloc(cap), hip.alloc(cap)]
    try:
        rng = np.random.default_rng(9)
        for nk in (5, 2):  # the second call is smaller: stale scratch must not leak
            k = rng.uniform(-0.5, 0.5, size=(nk, 3))
            h_ref, s_ref = nb.assemble(k)  # aosx_assemble_v1, host outputs
            poison = np.full(5 * nb.nnz, np.nan, dtype=np.complex64)
            for b in bufs:
                hip.upload(b, poison)
            rc, res = _noalloc(nb, k, bufs[0], bufs[1])
            assert rc == 0 and res.status == 0, bytes(res.error_message)
            got_h = np.empty((nk, nb.nnz), dtype=np.complex64)
            got_s = np.empty((nk, nb.nnz), dtype=np.complex64)
            hip.download(got_h, bufs[0])
            hip.download(got_s, bufs[1])
            say(f"  nk={nk}: bitwise equal h={np.array_equal(got_h.view(np.uint32), h_ref.view(np.uint32))}")
            assert np.array_equal(got_h.view(np.uint32), h_ref.view(np.uint32))
            assert np.array_equal(got_s.view(np.uint32), s_ref.view(np.uint32))
        # more k points than the scratch array is sized for: a clear error
        # from the host-side check, before any launch
        rc, res = _noalloc(nb, np.zeros((65, 3)), bufs[0], bufs[1])
        assert rc == 1 and res.status == 12
        assert b"[1, 64]" in bytes(res.error_message)
    finally:
        for b in bufs:
            hip.free(b)
        nb.close()
