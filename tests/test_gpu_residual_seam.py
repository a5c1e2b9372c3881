"""Eigenpair residual, returned Ritz vectors, the interval-solver plugin seam
and filter-pass saturation, against host FP64 (needs an MI355X).

* The device residual (src/aospectrum/ciss.py:362-372) is compared with the
  host FP64 ||H v - e S v|| on exact FP64 eigenpairs (residual at rounding
  level) and on perturbed pairs (a known, large residual).
* The Ritz vectors a band solve returns reproduce the residual it reports.
* ``HipCissSolver.solve_interval`` (the seam of ciss.py:253-391, injected
  through ``BandCalculation(solver=...)``, band.py:195,212-213) runs the
  explicit-H/S path on a pattern analysed with ``NativeBundle.from_csr``.
* A window whose states outnumber the filter-pass block while the moment rank
  stays below L*M must come back incomplete.
"""

import sys

import numpy as np
import pytest
import scipy.linalg as sl
from scipy import sparse

from aospectrum_amd import synth
from aospectrum_amd.band import BandCalculation, KPath
from aospectrum_amd.ciss import CissSettings, HipCissSolver, band_solve, eig_residual
from aospectrum_amd.data import EnergyInterval
from aospectrum_amd.sparse import NativeBundle, SparsePair
from oracle import bloch, dense

pytestmark = pytest.mark.gpu


def say(*parts):
    print("[gpu-test]", *parts, flush=True)
    sys.stdout.flush()


def _dense_c64(pattern, data):
    """The complex64 CSR values as an exact complex128 dense matrix."""
    return pattern.csr(data).toarray().astype(np.complex128)


@pytest.fixture(scope="module")
def native_small(small2d):
    say("create bundle small2d n=832")
    nb = NativeBundle(small2d, device=0)
    yield nb
    nb.close()


def test_device_residual_matches_host_fp64(small2d, small2d_pattern, native_small):
    k = (0.13, 0.37, 0.0)
    h, s = small2d_pattern.assemble(small2d.operators, k)
    H, S = _dense_c64(small2d_pattern, h), _dense_c64(small2d_pattern, s)
    w, V = sl.eigh(H, S)
    sel = np.flatnonzero((w > -2.02) & (w < 2.02))[:48]
    e, X = w[sel], V[:, sel]
    say("residual: exact FP64 pairs of the complex64 pencil,", sel.size, "pairs")
    got = eig_residual(native_small, e[None], X[None], h_data=h[None], s_data=s[None])[0]
    host = dense.eigenpair_residuals(H, S, X, e)
    assert np.max(host) < 1e-11
    assert np.max(got) < 1e-11
    # perturbed pairs: a known residual of order 1e-2..1 eV
    rng = np.random.default_rng(5)
    Xp = X + 1e-3 * (rng.standard_normal(X.shape) + 1j * rng.standard_normal(X.shape))
    ep = e + rng.uniform(-1e-2, 1e-2, e.size)
    say("residual: perturbed pairs")
    got = eig_residual(native_small, ep[None], Xp[None], h_data=h[None], s_data=s[None])[0]
    host = dense.eigenpair_residuals(H, S, Xp, ep)
    assert np.min(host) > 1e-4
    np.testing.assert_allclose(got, host, rtol=1e-10, atol=1e-13)
    # the same through on-device assembly at k: the device H(k), S(k) agree
    # with the oracle's to float32 rounding
    got_k = eig_residual(native_small, ep[None], Xp[None], kpoints=np.asarray([k]))[0]
    np.testing.assert_allclose(got_k, host, rtol=1e-5)


@pytest.mark.parametrize("name,window", [("small2d", 2.02), ("stress2d", 1.52)])
def test_band_vectors_reproduce_reported_residual(name, window):
    bundle = synth.generate(synth.SPECS[name])
    pat = bloch.Pattern.build(bundle.operators, bundle.basis.atom_offsets, bundle.n_orbitals)
    k = np.asarray([[0.21, 0.08, 0.0]])
    say(f"band vectors {name} n={bundle.n_orbitals}")
    nb = NativeBundle(bundle, device=0)
    try:
        st = CissSettings(refinement_passes=2, precision="fp32_ir", refinement_steps=2)
        (res,), _ = band_solve(nb, st, EnergyInterval(-window, window), kpoints=k, return_vectors=True)
        hd, sd = nb.assemble(k)
    finally:
        nb.close()
    assert res.complete and res.pass_rank < res.pass_width
    V = res.eigenvectors
    assert V.shape == (bundle.n_orbitals, res.energies_ev.size)
    H, S = _dense_c64(pat, hd[0]), _dense_c64(pat, sd[0])
    host = dense.eigenpair_residuals(H, S, V, res.energies_ev)
    say(f"  count={res.energies_ev.size} reported={res.maximum_residual_ev:.3e} host={host.max():.3e}")
    # the vectors come back in complex64 and the energies in float32: their
    # rounding adds up to ~|H| * 1e-7 to the FP64 residual of the device
    assert abs(host.max() - res.maximum_residual_ev) < 2e-6
    assert host.max() <= 1e-4
    # S-orthonormal Ritz vectors
    G = V.conj().T @ (S @ V)
    assert np.max(np.abs(G - np.eye(G.shape[0]))) < 1e-4
    # and they are eigenvectors of the FP64 pencil: the Rayleigh quotients
    rq = np.real(np.einsum("ij,ij->j", V.conj(), H @ V)) / np.real(np.einsum("ij,ij->j", V.conj(), S @ V))
    np.testing.assert_allclose(rq, res.energies_ev, atol=1e-4)


@pytest.mark.parametrize("name,window", [("small2d", 2.02), ("stress2d", 1.52)])
def test_hip_ciss_solver_plugin_seam(name, window):
    """BandCalculation(solver=HipCissSolver(...)) and a direct solve_interval
    on an assembled SparsePair, checked against FP64 dense eigh."""
    bundle = synth.generate(synth.SPECS[name])
    k = np.array([[0.0, 0.5, 0.0], [0.31, 0.12, 0.0]])
    kpath = KPath(k, (0, 1), ("A", "B"))
    st = CissSettings(refinement_passes=2, precision="bf16x3_ir", refinement_steps=2, refinement_policy="on_demand")
    display = EnergyInterval(-window + 0.02, window - 0.02)
    say(f"plugin seam {name}: BandCalculation(solver=HipCissSolver)")
    calc = BandCalculation(bundle, kpath, display, 0.0, st, solver=HipCissSolver(st))
    try:
        points = calc.solve_points([0, 1])
        pair = calc.workspace.update(bundle, k[1])
    finally:
        calc.close()
    pat = bloch.Pattern.build(bundle.operators, bundle.basis.atom_offsets, bundle.n_orbitals)
    lo, hi = calc.absolute_interval_ev.lower_ev, calc.absolute_interval_ev.upper_ev
    for p, kp in zip(points, k):
        h, s = pat.assemble(bundle.operators, kp)
        ref = dense.window_spectrum(pat.csr(h).toarray(), pat.csr(s).toarray(), lo, hi)
        say(f"  k={kp} count={p.energies_ev.size} oracle={ref.size} res={p.maximum_residual_ev:.2e}")
        assert p.energies_ev.size == ref.size
        assert np.max(np.abs(np.sort(p.energies_ev) - ref)) < 1e-4
        assert p.maximum_residual_ev <= 1e-4
    # a direct call on a host SparsePair: the reference's shared-pattern checks
    assert isinstance(pair, SparsePair)
    solver = HipCissSolver(st)
    try:
        res = solver.solve_interval(pair, calc.absolute_interval_ev)
        with pytest.raises(Exception):
            bad = SparsePair(pair.hamiltonian.astype(np.complex128), pair.overlap, pair.k_fractional)
            solver.solve_interval(bad, calc.absolute_interval_ev)
    finally:
        solver.close()
    np.testing.assert_allclose(np.sort(res.energies_ev), np.sort(points[1].energies_ev), atol=1e-4)


def test_filter_pass_saturation_marks_incomplete():
    """130 states packed within 1e-3 eV at one end of the window: the moment
    block loses all but ~2L directions (rank < L*M), and the filter-pass block
    (rank + 2L columns) cannot hold the cluster, so the k point must come back
    incomplete (the reference's policy for a saturated subspace, ciss.py:342)."""
    n = 400
    rng = np.random.default_rng(11)
    levels = np.concatenate([
        -0.95 + 1e-3 * rng.uniform(size=130),
        rng.uniform(-12.0, -1.6, 135),
        rng.uniform(1.6, 12.0, 135),
    ])
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    H = (Q * levels) @ Q.conj().T
    H = 0.5 * (H + H.conj().T)
    Hc = sparse.csr_matrix(H.astype(np.complex64))
    assert Hc.nnz == n * n and Hc.has_sorted_indices
    # S = I on the same (full) pattern: the shared pattern the seam requires
    Sfull = Hc.copy()
    rows = np.repeat(np.arange(n), np.diff(Hc.indptr))
    Sfull.data = (rows == Hc.indices).astype(np.complex64)
    pair = SparsePair(Hc, Sfull, (0.0, 0.0, 0.0))
    st = CissSettings(refinement_passes=2, precision="fp32_ir", refinement_steps=2)
    say("saturation: 130 clustered states, block 32, moment 4")
    solver = HipCissSolver(st)
    try:
        res = solver.solve_interval(pair, EnergyInterval(-1.0, 1.0))
    finally:
        solver.close()
    say(f"  moment rank={res.subspace_rank} pass rank={res.pass_rank}/{res.pass_width} "
        f"count={res.energies_ev.size} complete={res.complete}")
    assert res.subspace_rank < st.search_dimension
    assert res.pass_width < 130
    assert not res.complete
