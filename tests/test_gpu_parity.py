"""HIP Band path vs the CPU oracle, through the C-ABI (needs an MI355X)."""

import os
import sys

import numpy as np
import pytest

from aospectrum_amd import synth
from aospectrum_amd.ciss import CissSettings, HipShiftSolver, band_solve, probe_block
from aospectrum_amd.data import EnergyInterval
from aospectrum_amd.sparse import NativeBundle, SparsePair
from oracle import bloch, ciss as ociss, dense

pytestmark = pytest.mark.gpu


def say(*parts):
    print("[gpu-test]", *parts, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def native_small(small2d):
    say("create bundle small2d n=832")
    nb = NativeBundle(small2d, device=0)
    yield nb
    nb.close()


def test_assembly_matches_oracle(small2d, small2d_pattern, native_small):
    kpts = np.array([[0.0, 0.0, 0.0], [0.13, 0.37, 0.0], [0.5, 0.5, 0.0]])
    say("assembly small2d 3 k")
    h, s = native_small.assemble(kpts)
    indptr, indices, perm = native_small.pattern()
    assert np.array_equal(indptr, small2d_pattern.indptr)
    assert np.array_equal(indices, small2d_pattern.indices)
    assert np.array_equal(np.sort(perm), np.arange(small2d.n_orbitals))
    for i, k in enumerate(kpts):
        ho, so = small2d_pattern.assemble(small2d.operators, k)
        # complex64 sums of at most a few terms: agreement to float32 rounding
        assert np.max(np.abs(h[i] - ho)) <= 2e-6 * max(1.0, np.max(np.abs(ho)))
        assert np.max(np.abs(s[i] - so)) <= 2e-6
        # exact Hermitian projection: conj pairs and a real diagonal
        H = small2d_pattern.csr(h[i]).toarray()
        assert np.array_equal(H, H.conj().T)


def test_assembly_duplicate_shifts(tiny3d):
    """Several lattice shifts on one atom pair are summed (sparse.py:107-113)."""
    pat = bloch.Pattern.build(tiny3d.operators, tiny3d.basis.atom_offsets, tiny3d.n_orbitals)
    say("assembly tiny3d")
    nb = NativeBundle(tiny3d, device=0)
    try:
        k = np.array([[0.21, -0.33, 0.4]])
        h, s = nb.assemble(k)
        ho, so = pat.assemble(tiny3d.operators, k[0])
        assert np.max(np.abs(h[0] - ho)) <= 5e-6 * np.max(np.abs(ho))
        assert np.max(np.abs(s[0] - so)) <= 5e-6
    finally:
        nb.close()


def _matrices(small2d, small2d_pattern, k):
    h, s = small2d_pattern.assemble(small2d.operators, k)
    return SparsePair(small2d_pattern.csr(h), small2d_pattern.csr(s), tuple(k))


@pytest.mark.parametrize("precision,steps", [("fp32", 0), ("fp32_ir", 2), ("bf16x3_ir", 3)])
def test_shift_solve_matches_dense(small2d, small2d_pattern, precision, steps):
    m = _matrices(small2d, small2d_pattern, [0.13, 0.37, 0.0])
    n = m.n_orbitals
    V = probe_block(n, CissSettings())
    rhs = np.asarray(m.overlap @ V, dtype=np.complex64)
    z, _, _ = ociss.contour(-2.02, 2.02, ociss.Settings())
    shifts = z[[0, 7, 15, 16, 31]]
    say("shift solve", precision, "batch", shifts.size)
    solver = HipShiftSolver(precision=precision, refinement_steps=steps)
    try:
        out = solver.solve_shifted(m, shifts, rhs)
    finally:
        solver.close()
    Hd = m.hamiltonian.toarray().astype(np.complex128)
    Sd = m.overlap.toarray().astype(np.complex128)
    for b, zb in enumerate(shifts):
        A = Hd - complex(zb) * Sd
        x = np.linalg.solve(A, rhs.astype(np.complex128))
        err = np.linalg.norm(out.solutions[b] - x) / np.linalg.norm(x)
        say("  shift", b, "relative error %.2e" % err)
        # FP32 LU without pivoting: error ~ kappa * eps32; refinement -> ~eps32
        assert err < (2e-3 if precision == "fp32" else 2e-4)


def _oracle_window(small2d, small2d_pattern, k, lo, hi):
    h, s = small2d_pattern.assemble(small2d.operators, k)
    H = small2d_pattern.csr(h).toarray()
    S = small2d_pattern.csr(s).toarray()
    return dense.window_spectrum(H, S, lo, hi)


@pytest.mark.parametrize(
    "label,settings",
    [
        ("plain", CissSettings()),
        ("plain_pass2", CissSettings(refinement_passes=2)),
        ("batched_conj", CissSettings(contour_batch_size=16, k_batch=2, conjugate_pairs=True, refinement_passes=2)),
        (
            "mixed",
            CissSettings(
                contour_batch_size=16,
                k_batch=2,
                conjugate_pairs=True,
                precision="bf16x3_ir",
                refinement_steps=2,
                refinement_passes=2,
            ),
        ),
    ],
)
def test_band_solve_matches_fp64_oracle(small2d, small2d_pattern, native_small, label, settings):
    kpts = np.array([[0.13, 0.37, 0.0], [0.0, 0.0, 0.0]])
    interval = EnergyInterval(-2.02, 2.02)
    say("band solve", label)
    results, info = band_solve(native_small, settings, interval, kpoints=kpts)
    for k, r in zip(kpts, results):
        ref = _oracle_window(small2d, small2d_pattern, k, interval.lower_ev, interval.upper_ev)
        say(f"  k={k} count={r.energies_ev.size} ref={ref.size} rank={r.subspace_rank} "
            f"res={r.maximum_residual_ev:.2e} pivots={info['replaced_pivots']}")
        assert r.energies_ev.size == ref.size
        assert np.max(np.abs(np.sort(r.energies_ev) - ref)) < 1e-4
        assert r.complete
        if settings.refinement_passes >= 2:
            assert r.maximum_residual_ev <= 1e-4


MIXED_ON_DEMAND = CissSettings(
    contour_batch_size=16,
    k_batch=2,
    conjugate_pairs=True,
    precision="bf16x3_ir",
    refinement_steps=2,
    refinement_passes=2,
    refinement_policy="on_demand",
)


def test_on_demand_refinement_matches_oracle(small2d, small2d_pattern, native_small):
    """Refinement on demand matches the FP64 oracle (whichever k points it
    had to re-solve)."""
    kpts = np.array([[0.13, 0.37, 0.0], [0.0, 0.0, 0.0]])
    interval = EnergyInterval(-2.02, 2.02)
    say("band solve on_demand")
    results, info = band_solve(native_small, MIXED_ON_DEMAND, interval, kpoints=kpts)
    say("  refined", info["refined"])
    for k, r in zip(kpts, results):
        ref = _oracle_window(small2d, small2d_pattern, k, interval.lower_ev, interval.upper_ev)
        say(f"  k={k} count={r.energies_ev.size} ref={ref.size} res={r.maximum_residual_ev:.2e}")
        assert r.energies_ev.size == ref.size
        assert np.max(np.abs(np.sort(r.energies_ev) - ref)) < 1e-4
        assert r.maximum_residual_ev <= 1e-4


def test_on_demand_refinement_resolves_missed_targets(native_small):
    """A residual target no solve can meet sends every k point through the
    refined re-solve, whose results are those of the fixed policy."""
    from dataclasses import replace

    kpts = np.array([[0.13, 0.37, 0.0], [0.25, 0.0, 0.0]])
    interval = EnergyInterval(-2.02, 2.02)
    strict = replace(MIXED_ON_DEMAND, residual_target_ev=1e-12)
    fixed = replace(MIXED_ON_DEMAND, refinement_policy="fixed")
    say("band solve on_demand, unreachable target")
    got, info = band_solve(native_small, strict, interval, kpoints=kpts)
    want, info_fixed = band_solve(native_small, fixed, interval, kpoints=kpts)
    assert info["refined"] == [0, 1]
    # the first (unrefined) pass plus the refined re-solve
    assert info["factorizations"] == 2 * info_fixed["factorizations"]
    for a, b in zip(got, want):
        assert a.energies_ev.size == b.energies_ev.size
        assert np.max(np.abs(a.energies_ev - b.energies_ev), initial=0.0) <= 1e-6
        assert a.maximum_residual_ev == pytest.approx(b.maximum_residual_ev, rel=1e-6)


def test_graph_replays_match_eager_launches(small2d):
    """Factorisations and solves replayed from captured HIP graphs give the
    eager launches' results, replay after replay, with eager work and
    synchronisation in between; kernel timing works inside the graphs."""
    kpts = [np.array([[0.13, 0.37, 0.0]]), np.array([[0.25, 0.1, 0.0]]), np.array([[0.0, 0.5, 0.0]])]
    interval = EnergyInterval(-2.02, 2.02)
    settings = CissSettings(contour_batch_size=16, conjugate_pairs=True, precision="bf16x3_ir",
                            refinement_steps=1, refinement_passes=2)
    os.environ["AOSX_GRAPHS"] = "0"
    try:
        eager = NativeBundle(small2d, device=0)
    finally:
        del os.environ["AOSX_GRAPHS"]
    graphed = NativeBundle(small2d, device=0)
    try:
        for rep in range(2):
            for k in kpts:
                say("graph replay", rep, k[0])
                (a,), _ = band_solve(graphed, settings, interval, kpoints=k)
                (b,), _ = band_solve(eager, settings, interval, kpoints=k)
                assert a.energies_ev.size == b.energies_ev.size
                assert np.max(np.abs(a.energies_ev - b.energies_ev), initial=0.0) <= 1e-6
                assert a.maximum_residual_ev == pytest.approx(b.maximum_residual_ev, rel=1e-5)
        graphed.kernel_stats(True)
        band_solve(graphed, settings, interval, kpoints=kpts[0])
        band_solve(graphed, settings, interval, kpoints=kpts[1])
        ks = graphed.kernel_stats(False)
        say("graph kernel stats", {k: v for k, v in ks.items() if k.endswith("launches")})
        # small2d has 13 tile columns: one step group, so no trailing launches
        assert ks["inner_launches"] > 0 and ks["inner_ms"] > 0.0
        assert ks["getrf_launches"] > 0 and ks["solve_launches"] > 0 and ks["solve_ms"] > 0.0
    finally:
        graphed.close()
        eager.close()


def _bundle_with_env(bundle, **env):
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        return NativeBundle(bundle, device=0)
    finally:
        for k, v in saved.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.mark.parametrize("precision,form", [("bf16x3_ir", 1), ("bf16_ir", 1)])
def test_pipelined_quad_updates_match_step_kernel(small2d, precision, form):
    """The trailing update on 2 x 2 quads, staged through LDS by K quarter and
    component, accumulates the same three-multiplication products in the same
    order as the one-target reference kernel (AOSX_QUAD_PIPELINE=0): with step
    groups of 4 (three trailing launches per factorisation) both give the same
    energies and residuals."""
    kpts = np.array([[0.13, 0.37, 0.0], [0.25, 0.1, 0.0]])
    interval = EnergyInterval(-2.02, 2.02)
    settings = CissSettings(contour_batch_size=16, conjugate_pairs=True, precision=precision,
                            refinement_steps=1, refinement_passes=2)
    step = _bundle_with_env(small2d, AOSX_SCHUR_GROUP=4, AOSX_QUAD_PIPELINE=0)
    piped = _bundle_with_env(small2d, AOSX_SCHUR_GROUP=4, AOSX_QUAD_PIPELINE=form)
    try:
        piped.kernel_stats(True)
        got, _ = band_solve(piped, settings, interval, kpoints=kpts)
        ks = piped.kernel_stats(False)
        ref, _ = band_solve(step, settings, interval, kpoints=kpts)
        say("pipelined quad", precision, ks["schur_launches"], [r.energies_ev.size for r in got])
        assert ks["schur_launches"] > 0
        for a, b in zip(got, ref):
            assert a.energies_ev.size == b.energies_ev.size
            assert np.array_equal(a.energies_ev, b.energies_ev)
            # the FP64 residual reduction order varies from run to run
            assert a.maximum_residual_ev == pytest.approx(b.maximum_residual_ev, rel=1e-9)
    finally:
        piped.close()
        step.close()


@pytest.mark.parametrize("precision", ["bf16x3_ir", "fp32_ir"])
def test_folded_group_solves_match_step_launches(small2d, precision):
    """One launch per step group for the in-group triangular steps applies
    the same products in the same order as a diagonal and an update launch per
    step: with step groups of 4 and 32 matrices per solve (grouped sweeps,
    plain and conjugate-transposed factors) both give identical energies."""
    kpts = np.array([[0.13, 0.37, 0.0], [0.25, 0.1, 0.0]])
    interval = EnergyInterval(-2.02, 2.02)
    settings = CissSettings(contour_batch_size=16, conjugate_pairs=True, precision=precision,
                            refinement_steps=1, refinement_passes=2)
    step = _bundle_with_env(small2d, AOSX_SCHUR_GROUP=4, AOSX_SOLVE_FOLD=0)
    folded = _bundle_with_env(small2d, AOSX_SCHUR_GROUP=4, AOSX_SOLVE_FOLD=1)
    try:
        got, _ = band_solve(folded, settings, interval, kpoints=kpts)
        ref, _ = band_solve(step, settings, interval, kpoints=kpts)
        say("folded group solves", precision, [r.energies_ev.size for r in got])
        for a, b in zip(got, ref):
            assert a.energies_ev.size == b.energies_ev.size
            assert np.array_equal(a.energies_ev, b.energies_ev)
            assert a.maximum_residual_ev == pytest.approx(b.maximum_residual_ev, rel=1e-9)
    finally:
        folded.close()
        step.close()


def test_per_shift_precision_dispatch(small2d, small2d_pattern):
    """Per-shift precision runs the BF16x3 and the single-BF16 matrices of a
    batch in launches of their own: threshold 2 (no far point) gives the
    BF16x3 results and threshold 0 (every point far) the single-BF16 results,
    bit for bit; a threshold in between matches the FP64 oracle."""
    kpts = np.array([[0.13, 0.37, 0.0], [0.25, 0.1, 0.0]])
    interval = EnergyInterval(-2.02, 2.02)
    common = dict(contour_batch_size=16, k_batch=2, conjugate_pairs=True, refinement_steps=2, refinement_passes=2)
    nb = NativeBundle(small2d, device=0)
    try:
        out = {}
        for label, st in [
            ("bf16x3", CissSettings(precision="bf16x3_ir", **common)),
            ("bf16", CissSettings(precision="bf16_ir", **common)),
            ("none_far", CissSettings(precision="per_shift_ir", far_shift_threshold=2.0, **common)),
            ("all_far", CissSettings(precision="per_shift_ir", far_shift_threshold=0.0, **common)),
            ("half_far", CissSettings(precision="per_shift_ir", far_shift_threshold=0.7, **common)),
        ]:
            out[label], _ = band_solve(nb, st, interval, kpoints=kpts)
            say("per-shift", label, [r.energies_ev.size for r in out[label]],
                [f"{r.maximum_residual_ev:.2e}" for r in out[label]])
    finally:
        nb.close()
    for a, b in [("none_far", "bf16x3"), ("all_far", "bf16")]:
        for x, y in zip(out[a], out[b]):
            assert np.array_equal(x.energies_ev, y.energies_ev)
            assert x.maximum_residual_ev == pytest.approx(y.maximum_residual_ev, rel=1e-9)
    for k, r in zip(kpts, out["half_far"]):
        ref = _oracle_window(small2d, small2d_pattern, k, interval.lower_ev, interval.upper_ev)
        assert r.energies_ev.size == ref.size
        assert np.max(np.abs(np.sort(r.energies_ev) - ref)) < 1e-4
        assert r.maximum_residual_ev <= 1e-4


def test_autotune_picks_a_timed_group_and_keeps_results(small2d, small2d_pattern):
    """The autotuner times every candidate step-group size and keeps the
    fastest; solves under each plan match the FP64 oracle."""
    nb = NativeBundle(small2d, device=0)
    try:
        k = np.array([[0.13, 0.37, 0.0]])
        interval = EnergyInterval(-2.02, 2.02)
        settings = CissSettings(contour_batch_size=16, conjugate_pairs=True, precision="bf16x3_ir",
                                refinement_steps=1, refinement_passes=2)
        ref = _oracle_window(small2d, small2d_pattern, k[0], interval.lower_ev, interval.upper_ev)
        for cands in ((1,), (4,), (1, 4, 16)):
            say("autotune", cands)
            tune = nb.autotune(settings.factorizations_per_launch, "bf16x3_ir", candidates=cands)
            assert tune["chosen_group"] in cands
            assert set(tune["factor_ms"]) == set(cands) and all(v > 0 for v in tune["factor_ms"].values())
            assert tune["chosen_group"] == min(tune["factor_ms"], key=tune["factor_ms"].get)
            (r,), _ = band_solve(nb, settings, interval, kpoints=k)
            assert r.energies_ev.size == ref.size
            assert np.max(np.abs(np.sort(r.energies_ev) - ref)) < 1e-4
            assert r.maximum_residual_ev <= 1e-4
    finally:
        nb.close()


def test_streamed_factorisations_match_resident(native_small):
    """With at most one factorisation allowed resident, every filter pass
    factors each shift group again; the result is the resident one's."""
    from dataclasses import replace

    kpts = np.array([[0.13, 0.37, 0.0], [0.0, 0.0, 0.0]])
    interval = EnergyInterval(-2.02, 2.02)
    resident = replace(MIXED_ON_DEMAND, refinement_policy="fixed", refinement_steps=1)
    streamed = replace(resident, resident_limit=1)
    say("band solve resident vs streamed")
    a, ia = band_solve(native_small, resident, interval, kpoints=kpts)
    b, ib = band_solve(native_small, streamed, interval, kpoints=kpts)
    # moment pass plus two filter passes, each factoring all 2 x 16 shifts
    assert ib["factorizations"] == 3 * ia["factorizations"]
    for x, y in zip(a, b):
        assert x.energies_ev.size == y.energies_ev.size
        assert np.max(np.abs(x.energies_ev - y.energies_ev), initial=0.0) <= 1e-6
        assert y.maximum_residual_ev <= 1e-4


def test_slab26k_precisions_agree():
    """26,000-orbital slab (too large for the dense oracle): the BF16x3 path
    with refinement on demand and the FP32 path with two fixed refinement
    steps find the same in-window states, each with an FP64 eigenpair
    residual below 1e-4 eV."""
    bundle = synth.generate(synth.SPECS["slab26k_a"])
    nb = NativeBundle(bundle, device=0)
    try:
        k = np.array([[0.25, 0.0, 0.0]])
        interval = EnergyInterval(-2.02, 2.02)  # the calibrated window (about 50 states)
        base = dict(contour_batch_size=16, conjugate_pairs=True, refinement_passes=2)
        runs = {}
        for label, extra in (("bf16x3", dict(precision="bf16x3_ir", refinement_steps=2,
                                              refinement_policy="on_demand")),
                             ("fp32", dict(precision="fp32_ir", refinement_steps=2))):
            say("slab26k", label)
            (r,), info = band_solve(nb, CissSettings(**base, **extra), interval, kpoints=k)
            say(f"  count={r.energies_ev.size} rank={r.subspace_rank} res={r.maximum_residual_ev:.2e} "
                f"peak={info['peak_device_bytes'] / 1e9:.1f} GB")
            assert r.complete and r.maximum_residual_ev <= 1e-4
            runs[label] = r
        a, b = runs["bf16x3"], runs["fp32"]
        assert a.energies_ev.size == b.energies_ev.size > 0
        assert np.max(np.abs(a.energies_ev - b.energies_ev)) < 1e-4
    finally:
        nb.close()


def test_band_solve_matches_oracle_ciss(small2d, small2d_pattern, native_small):
    """Single-pass plain CISS: same energies and rank as the CPU restatement
    of the reference algorithm (same probes, same settings)."""
    k = np.array([0.13, 0.37, 0.0])
    interval = EnergyInterval(-2.02, 2.02)
    m = _matrices(small2d, small2d_pattern, k)
    ref = ociss.solve_interval(m.hamiltonian, m.overlap, interval.lower_ev, interval.upper_ev)
    say("band solve vs oracle CISS")
    (r,), _ = band_solve(native_small, CissSettings(), interval, kpoints=k[None, :])
    assert r.energies_ev.size == ref.energies_ev.size
    assert np.max(np.abs(r.energies_ev - ref.energies_ev)) < 1e-4
    assert abs(r.subspace_rank - ref.rank) <= 1
    # same order of residual as the reference algorithm (both ~1e-3 eV here)
    assert r.maximum_residual_ev < 5 * ref.maximum_residual_ev + 1e-4


@pytest.fixture(scope="module")
def proxy3d():
    """216-atom periodic 3D bundle with 77 states in [-1.02, 1.02] eV -- more
    than two moment blocks of 32 probes can hold, so moments 2 and 3 matter."""
    return synth.generate(synth.scaled_spec("bulk3d", grid=(6, 6, 6), cell=(14.4, 14.4, 14.4), gap=1.8))


@pytest.mark.parametrize("label", ["plain", "mixed"])
def test_band_solve_needs_all_moments(proxy3d, label):
    pat = bloch.Pattern.build(proxy3d.operators, proxy3d.basis.atom_offsets, proxy3d.n_orbitals)
    k = np.array([0.0, 0.2, 0.0])
    interval = EnergyInterval(-1.02, 1.02)
    h, s = pat.assemble(proxy3d.operators, k)
    ref = dense.window_spectrum(pat.csr(h).toarray(), pat.csr(s).toarray(), interval.lower_ev, interval.upper_ev)
    settings = {
        "plain": CissSettings(precision="fp32_ir", refinement_steps=2, refinement_passes=2),
        "mixed": CissSettings(contour_batch_size=16, conjugate_pairs=True, precision="bf16x3_ir",
                              refinement_steps=2, refinement_passes=2),
    }[label]
    say("proxy3d", label, "oracle count", ref.size)
    nb = NativeBundle(proxy3d, device=0)
    try:
        (r,), _ = band_solve(nb, settings, interval, kpoints=k[None, :])
    finally:
        nb.close()
    say(f"  count={r.energies_ev.size} rank={r.subspace_rank} res={r.maximum_residual_ev:.2e}")
    assert r.subspace_rank >= ref.size
    assert r.energies_ev.size == ref.size
    assert np.max(np.abs(np.sort(r.energies_ev) - ref)) < 1e-4
    assert r.maximum_residual_ev <= 1e-4


# path points of the 121-point Y-G-X-M path on both sides of each change of
# the FP64 count (66 -> 67 at 29/30, 67 -> 68 at 66/67, 68 -> 67 at 69/70)
BULK3D_GOLDEN_K = (0, 29, 30, 66, 67, 69, 70, 120)


def test_bulk3d_matches_golden_inertia_counts():
    """13,000-orbital bulk: in-window counts equal the FP64 Sylvester-inertia
    counts committed in tests/golden/bulk3d_counts.json (all 121 path points;
    checked here at the points where the count changes, covering counts 66,
    67 and 68), residual <= 1e-4 eV."""
    import json
    from pathlib import Path

    golden = json.loads((Path(__file__).parent / "golden" / "bulk3d_counts.json").read_text())
    assert golden["k_index"] == list(range(121))
    bundle = synth.generate(synth.SPECS["bulk3d"])
    lo, hi = golden["interval_ev"]
    kp = np.asarray([golden["kpoints"][i] for i in BULK3D_GOLDEN_K])
    want = [golden["counts"][i] for i in BULK3D_GOLDEN_K]
    assert {66, 67, 68} <= set(want)
    settings = CissSettings(contour_batch_size=16, k_batch=4, conjugate_pairs=True, precision="bf16x3_ir",
                            refinement_steps=2, refinement_passes=2)
    say("bulk3d golden counts", dict(zip(BULK3D_GOLDEN_K, want)))
    nb = NativeBundle(bundle, device=0)
    try:
        results, _ = band_solve(nb, settings, EnergyInterval(lo, hi), kpoints=kp)
    finally:
        nb.close()
    for i, r, count in zip(BULK3D_GOLDEN_K, results, want):
        say(f"  k={i} count={r.energies_ev.size} golden={count} rank={r.subspace_rank} res={r.maximum_residual_ev:.2e}")
        assert r.energies_ev.size == count
        assert r.maximum_residual_ev <= 1e-4


def test_bulk3d_full_path_headline_settings_match_golden_counts():
    """The benchmark's configuration (4 k points per batch, 16 conjugate
    pairs, BF16x3 factors, refinement on demand, 2 filter passes) over all 121
    path points of the 13,000-orbital bulk: every in-window count equals the
    committed FP64 Sylvester-inertia count and every residual is <= 1e-4 eV."""
    import json
    from pathlib import Path

    golden = json.loads((Path(__file__).parent / "golden" / "bulk3d_counts.json").read_text())
    bundle = synth.generate(synth.SPECS["bulk3d"])
    lo, hi = golden["interval_ev"]
    kp = np.asarray(golden["kpoints"])
    settings = CissSettings(contour_batch_size=16, k_batch=4, conjugate_pairs=True, precision="bf16x3_ir",
                            refinement_steps=2, refinement_passes=2, refinement_policy="on_demand")
    say("bulk3d full path, 121 k points")
    nb = NativeBundle(bundle, device=0)
    try:
        got = []
        for b0 in range(0, kp.shape[0], 20):
            results, info = band_solve(nb, settings, EnergyInterval(lo, hi), kpoints=kp[b0:b0 + 20])
            got += [(r.energies_ev.size, r.maximum_residual_ev) for r in results]
            say(f"  k {b0}..{b0 + len(results) - 1}: counts {[c for c, _ in got[b0:]]} refined {info['refined']}")
    finally:
        nb.close()
    counts = [c for c, _ in got]
    mismatch = {i: (c, g) for i, (c, g) in enumerate(zip(counts, golden["counts"])) if c != g}
    assert not mismatch, mismatch
    assert max(r for _, r in got) <= 1e-4


@pytest.fixture(scope="module")
def stress2d():
    return synth.generate(synth.SPECS["stress2d"])


@pytest.mark.parametrize(
    "label,settings,refined",
    [
        # complex64 FP32 LU without refinement, reference launch shape, with
        # the filter passes (one pass alone misses states bunched at a window
        # end, as the reference algorithm does)
        ("plain_c64", CissSettings(refinement_passes=2), False),
        ("fp32_fp64ir", CissSettings(precision="fp32_ir", refinement_steps=2, refinement_passes=2), True),
        ("bf16x3_fp64ir", CissSettings(contour_batch_size=16, k_batch=2, conjugate_pairs=True,
                                       precision="bf16x3_ir", refinement_steps=2, refinement_passes=2), True),
        # one BF16 product for the half of the contour farthest from the real
        # axis (|sin theta| >= 0.7: 8 of the 16 conjugate pairs), BF16x3 for
        # the rest
        ("per_shift_fp64ir", CissSettings(contour_batch_size=16, k_batch=2, conjugate_pairs=True,
                                          precision="per_shift_ir", far_shift_threshold=0.7, refinement_steps=2,
                                          refinement_passes=2), True),
    ],
)
def test_precision_stress_against_fp64(stress2d, label, settings, refined):
    """3,328 orbitals, min eig S(k) = 1e-3, clustered onsite levels: every
    precision mode against the FP64 dense generalized eigensolver."""
    pat = bloch.Pattern.build(stress2d.operators, stress2d.basis.atom_offsets, stress2d.n_orbitals)
    kpts = np.array([[0.0, 0.0, 0.0], [0.5, 0.25, 0.0]])
    interval = EnergyInterval(-1.52, 1.52)
    say("stress2d", label)
    nb = NativeBundle(stress2d, device=0)
    try:
        results, info = band_solve(nb, settings, interval, kpoints=kpts)
    finally:
        nb.close()
    for k, r in zip(kpts, results):
        h, s = pat.assemble(stress2d.operators, k)
        ref = dense.window_spectrum(pat.csr(h).toarray(), pat.csr(s).toarray(), interval.lower_ev, interval.upper_ev)
        err = float(np.max(np.abs(np.sort(r.energies_ev) - ref))) if r.energies_ev.size == ref.size else np.inf
        say(f"  k={k} count={r.energies_ev.size} ref={ref.size} rank={r.subspace_rank} "
            f"max|dE|={err:.2e} res={r.maximum_residual_ev:.2e} pivots={info['replaced_pivots']}")
        assert r.energies_ev.size == ref.size
        assert err < 1e-4
        if refined:
            assert r.maximum_residual_ev <= 1e-4
