"""The CPU oracle pinned against closed forms, invariants and the FP64
dense generalized eigensolver (CPU only)."""

import numpy as np
import pytest

from oracle import bloch, ciss, dense


@pytest.mark.parametrize("k", [(0.0, 0.0, 0.0), (0.13, 0.37, 0.0), (0.5, 0.5, 0.0)])
def test_assembly_matches_dense_closed_form(small2d, small2d_pattern, k):
    h, s = small2d_pattern.assemble(small2d.operators, k)
    Hd, Sd = bloch.dense_bloch(small2d.operators, small2d.basis.atom_offsets, small2d.n_orbitals, k)
    assert np.max(np.abs(small2d_pattern.csr(h).toarray() - Hd)) < 2e-6 * np.max(np.abs(Hd))
    assert np.max(np.abs(small2d_pattern.csr(s).toarray() - Sd)) < 2e-6
    # every dense nonzero lies on the pattern
    mask = small2d_pattern.csr(np.ones(small2d_pattern.nnz, np.complex64)).toarray() != 0
    assert not np.any(np.abs(Hd[~mask]) > 0)


def test_assembly_sums_duplicate_shifts(tiny3d):
    pat = bloch.Pattern.build(tiny3d.operators, tiny3d.basis.atom_offsets, tiny3d.n_orbitals)
    # several blocks share an atom pair: fewer CSR entries than block entries
    assert pat.source.size > pat.nnz
    k = (0.21, -0.33, 0.4)
    h, s = pat.assemble(tiny3d.operators, k)
    Hd, Sd = bloch.dense_bloch(tiny3d.operators, tiny3d.basis.atom_offsets, tiny3d.n_orbitals, k)
    assert np.max(np.abs(pat.csr(h).toarray() - Hd)) < 5e-6 * np.max(np.abs(Hd))
    assert np.max(np.abs(pat.csr(s).toarray() - Sd)) < 5e-6


def test_assembly_invariants(small2d, small2d_pattern):
    ops = small2d.operators
    k = np.array([0.13, 0.37, 0.0])
    h, _ = small2d_pattern.assemble(ops, k)
    H = small2d_pattern.csr(h).toarray()
    assert np.array_equal(H, H.conj().T)                       # exact Hermitian projection
    hg, _ = small2d_pattern.assemble(ops, k + np.array([1.0, -1.0, 0.0]))
    assert np.max(np.abs(hg - h)) < 1e-5 * np.max(np.abs(h))   # H(k + G) = H(k)
    hm, _ = small2d_pattern.assemble(ops, -k)
    assert np.max(np.abs(hm - np.conj(h))) < 1e-6 * np.max(np.abs(h))  # real blocks
    h0, s0 = small2d_pattern.assemble(ops, np.zeros(3))
    assert np.all(h0.imag == 0) and np.all(s0.imag == 0)       # Gamma is real


def test_non_hermitian_topology_raises(small2d):
    from aospectrum_amd.data import LocalizedOperatorBlocks

    ops = small2d.operators
    # drop one off-site block: its mirror loses its partner
    off = np.flatnonzero(ops.row_atom != ops.col_atom)[0]
    keep = np.ones(ops.n_blocks, bool)
    keep[off] = False
    sizes = np.diff(ops.block_offsets)[keep]
    vals = np.concatenate([np.arange(ops.block_offsets[b], ops.block_offsets[b + 1]) for b in np.flatnonzero(keep)])
    broken = LocalizedOperatorBlocks(
        ops.row_atom[keep], ops.col_atom[keep], ops.lattice_shift[keep],
        np.concatenate(([0], np.cumsum(sizes))), ops.block_shapes[keep],
        np.asarray(ops.h_values)[vals], np.asarray(ops.s_values)[vals],
    )
    with pytest.raises(bloch.TopologyError):
        bloch.Pattern.build(broken, small2d.basis.atom_offsets, small2d.n_orbitals)


def test_contour_quadrature_filter():
    """f(l) = -sum_j w_j / (l - z_j) approximates the window indicator."""
    st = ciss.Settings()
    z, t, w = ciss.contour(-1.0, 1.0, st)
    lam = np.array([-0.9, -0.5, 0.0, 0.3, 0.95, -3.0, 2.5, 10.0])
    f = np.array([-(w / (l - z)).sum() for l in lam])
    # 32 nodes on a flat (0.1) ellipse: ~10% ripple inside, ~1e-8 leak outside
    assert np.all(np.abs(f[:5] - 1.0) < 0.15)
    assert np.all(np.abs(f[5:]) < 1e-6)
    # mirrored nodes are conjugate to float32 rounding
    assert np.max(np.abs(z[::-1] - np.conj(z))) < 1e-6
    assert np.max(np.abs(w[::-1] - np.conj(w))) < 1e-7


def test_probe_block_is_orthonormal():
    v = ciss.probes(300, ciss.Settings())
    assert v.shape == (300, 32) and v.dtype == np.complex64
    assert np.max(np.abs(v.conj().T @ v - np.eye(32))) < 1e-5


def test_oracle_ciss_matches_fp64_eigensolver(small2d, small2d_pattern):
    k = (0.13, 0.37, 0.0)
    lo, hi = -2.02, 2.02
    h, s = small2d_pattern.assemble(small2d.operators, k)
    H, S = small2d_pattern.csr(h), small2d_pattern.csr(s)
    res = ciss.solve_interval(H, S, lo, hi)
    ref = dense.window_spectrum(H.toarray(), S.toarray(), lo, hi)
    assert res.energies_ev.size == ref.size == dense.inertia_count(H.toarray(), S.toarray(), lo, hi)
    assert np.max(np.abs(res.energies_ev - ref)) < 1e-4
    assert not res.saturated


def test_moments_match_closed_form_filter_response():
    """On a diagonal pencil H = diag(l), S = I with unit probes, the shifted
    solutions are X_j = diag(1 / (l - z_j)), so moment block m is the diagonal
    f_m(l) = -sum_j w_j t_j^m / (l - z_j), the trapezoid rule of
    (1 / 2 pi i) contour-integral ((z - c) / rho)^m / (z - l) dz: it must be
    ((l - c) / rho)^m inside the ellipse and ~0 outside.  A wrong sign of the
    weights, a conjugated t_j or a dropped power fails it."""
    st = ciss.Settings(block_size=8, moment_size=4)
    lo, hi = -1.0, 1.0
    z, t, w = ciss.contour(lo, hi, st)
    c = 0.5 * (lo + hi)
    rho = 0.5 * (hi - lo) + 2.0 * st.endpoint_guard_ev
    lam = np.array([-0.8, -0.45, -0.1, 0.0, 0.2, 0.55, 0.85, -2.5, 1.9, 4.0], dtype=np.float64)
    n = st.block_size
    assert lam.size >= n
    lam = lam[:n]
    sols = [np.diag(1.0 / (lam - complex(zj))).astype(np.complex64) for zj in z]
    M = ciss.moments_from_solutions(sols, t, w, st, n)
    inside = np.abs(lam - c) < 0.9 * rho
    # 32 nodes on a flat (0.1) ellipse: ~8 % ripple of f_0 inside, less on
    # the higher moments, whose integrands vanish at the centre
    tol = [0.1, 0.06, 0.06, 0.06]
    for m in range(st.moment_size):
        f = np.diag(M[:, m * n:(m + 1) * n])
        want = np.where(inside, ((lam - c) / rho) ** m, 0.0)
        assert np.all(np.abs(f[inside] - want[inside]) < tol[m]), (m, f, want)
        assert np.all(np.abs(f[~inside]) < 1e-6), (m, f)
        assert np.max(np.abs(M[:, m * n:(m + 1) * n] - np.diag(f))) == 0.0
    # the same check fails for plausible slips: a conjugated t_j, a flipped
    # weight sign
    for bad_t, bad_w in ((np.conj(t), w), (t, -w)):
        Mbad = ciss.moments_from_solutions(sols, bad_t, bad_w, st, n)
        errs = [np.max(np.abs(np.diag(Mbad[:, m * n:(m + 1) * n])[inside] - ((lam[inside] - c) / rho) ** m))
                for m in range(st.moment_size)]
        assert max(errs) > 0.1


def test_inertia_count_matches_eigenvalues(small2d, small2d_pattern):
    h, s = small2d_pattern.assemble(small2d.operators, (0.0, 0.25, 0.0))
    H, S = small2d_pattern.csr(h).toarray(), small2d_pattern.csr(s).toarray()
    for lo, hi in ((-2.02, 2.02), (-5.0, 0.0), (1.0, 3.0)):
        assert dense.inertia_count(H, S, lo, hi) == dense.window_spectrum(H, S, lo, hi).size
