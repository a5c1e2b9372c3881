#!/usr/bin/env python3
"""Band-path benchmark on one MI355X.

Workload: the 13,000-orbital 3D bulk bundle (1,000 atoms on a jittered
10x10x10 grid, 13 AOs per atom, 6 Angstrom cutoff over R in {-1,0,1}^3),
the 121-point Y-G-X-M path and the [-1, 1] eV window (solve interval
[-1.02, 1.02] eV), CISS with 32 contour points, block 32, moment 4.

One "step" is one batch of ``--k-batch`` k points through the whole path
(on-device assembly, all shifted factorisations and solves, moments,
compression, Rayleigh-Ritz, filter passes and FP64 residuals) in the batched
mixed-precision configuration: 4 k points x 16 conjugate-pair shifts per
factorisation launch, BF16x3 Schur updates, refinement on demand (a k point
whose FP64 eigenpair residual misses 1e-4 eV is solved again with FP64-
residual refinement, inside the timed region; ``refined_kpoints`` lists them)
and two filter passes.  ``value`` is its k-points/s.

The same k points are also solved in the plain configuration -- one shift per
factorisation launch, one k point at a time, FP32 LU with the reference's
two iterative-refinement steps (cuDSS IR_N_STEPS = 2) on the last filter
application and one step on the moment pass and the first filter pass -- for
the speedup figure.  Both must meet the 1e-4 eV residual bar, and every timed
k point's in-window count is checked against the FP64 Sylvester-inertia
counts committed in tests/golden/bulk3d_counts.json.

The roofline entry reports the dominant kernel (the trailing Schur update),
timed with HIP events on the library's stream over the timed steps; its
algorithmic flops are 8 real flops per complex multiply-add of the tile
updates, and its HBM traffic per launch is read from the rocprofv3 counter
record committed under profiles/.  A bounded serving leg (Poisson arrivals
of jobs over several bundles through the continuous-batching server) reports
job latency percentiles beside throughput.  The CPU baseline runs the
oracle's CISS for a bounded sample of one k point on the host cores.
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_TFLOPS = {"fp32": 157.3, "fp32_ir": 157.3, "bf16x3_ir": 2500.0, "bf16_ir": 2500.0, "per_shift_ir": 2500.0}
# MFMA flops the trailing update issues per algorithmic flop (8 per complex
# multiply-add): the BF16 forms take three real products per complex product
# (re*re, im*im, (re+im)*(re+im)), each as hi*hi + hi*lo + lo*hi (BF16x3:
# 9 x 2 / 8) or hi*hi (3 x 2 / 8); FP32 takes the four textbook products
MFMA_PRODUCTS = {"fp32": 1, "fp32_ir": 1, "bf16x3_ir": 2.25, "bf16_ir": 0.75, "per_shift_ir": None}


def log(*parts):
    print("[bench]", *parts, file=sys.stderr, flush=True)


def schur_kernel(precision: str) -> str:
    """The trailing-update kernel the library dispatches for `precision`
    (pipeline.hip, Bundle::factor)."""
    if not (precision.startswith("bf16") or precision == "per_shift_ir"):
        return "k_schur_group"
    three = {"bf16x3_ir": "true", "bf16_ir": "false", "per_shift_ir": "true|false"}[precision]
    form = "k_schur_pairs_3m" if os.environ.get("AOSX_QUAD_PIPELINE", "1") == "0" else "k_schur_quad_3m"
    return f"{form}<{three}>"


def traffic_record(workload: str, precision: str, kernel: str) -> dict | None:
    """HBM bytes per dispatch of `kernel` from the counter records committed
    under profiles/ (lines 'record: {json}'), if one matches."""
    for path in sorted((ROOT / "profiles").glob("*.md")):
        for line in path.read_text(encoding="utf-8").splitlines():
            m = re.match(r"^record:\s*(\{.*\})\s*$", line.strip())
            if not m:
                continue
            rec = json.loads(m.group(1))
            if (rec.get("workload"), rec.get("precision"), rec.get("kernel")) == (workload, precision, kernel):
                return {**rec, "file": str(path.relative_to(ROOT))}
    return None


def golden_counts(workload: str) -> dict[int, int]:
    path = ROOT / "tests" / "golden" / f"{workload}_counts.json"
    if not path.is_file():
        return {}
    doc = json.loads(path.read_text())
    idx = doc.get("k_index", list(range(len(doc["counts"]))))
    return {int(i): int(c) for i, c in zip(idx, doc["counts"])}


def cpu_baseline(bundle, kpoint, interval, shifts_sampled: int) -> dict:
    """The oracle's CISS of one k point on the host: assembly, the shifted
    solves (dense complex64 LAPACK on all host threads, `shifts_sampled` of
    the 32 timed and scaled), moments, compression, Rayleigh-Ritz and the
    residual."""
    from oracle import bloch, ciss as ociss

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    st = ociss.Settings()
    t0 = time.perf_counter()
    pat = bloch.Pattern.build(bundle.operators, bundle.basis.atom_offsets, bundle.n_orbitals)
    h, s = pat.assemble(bundle.operators, kpoint)
    H, S = pat.csr(h), pat.csr(s)
    z, t, w = ociss.contour(interval.lower_ev, interval.upper_ev, st)
    V = ociss.probes(bundle.n_orbitals, st)
    rhs = np.asarray(S @ V, dtype=np.complex64)
    t1 = time.perf_counter()
    sols = [ociss.lapack_solver(H, S, z[j], rhs) for j in range(shifts_sampled)]
    t2 = time.perf_counter()
    # the remaining shifts' solutions are stood in by the sampled ones: the
    # subspace stage's cost depends on the shapes only
    full = [sols[j % shifts_sampled] for j in range(st.integration_points)]
    M = ociss.moments_from_solutions(full, t, w, st, bundle.n_orbitals)
    Q, sigma = ociss.compress(S, M, st.rank_relative_tolerance)
    E, X, C = ociss.rayleigh_ritz(H, S, Q)
    keep = ociss.nonspurious(C, sigma, st.spurious_relative_threshold)
    if keep.any():
        R = H @ X[:, keep] - (S @ X[:, keep]) * E[keep][None, :]
        float(np.max(np.linalg.norm(R, axis=0)))
    t3 = time.perf_counter()
    per_k = (t1 - t0) + (t2 - t1) * st.integration_points / shifts_sampled + (t3 - t2)
    return {
        "value": 1.0 / per_k,
        "unit": "k-points/s",
        "cores": threads,
        "kind": "port",
        "sample": f"one k point of the workload: host assembly {t1 - t0:.1f} s, {shifts_sampled} of the 32 shifted "
        f"factorisations+solves {t2 - t1:.1f} s (scaled x{st.integration_points / shifts_sampled:.2f}), moments, "
        f"compression, Rayleigh-Ritz and residual {t3 - t2:.1f} s; dense complex64 LAPACK, {threads} threads",
        "method_note": "the planned SciPy SuperLU (COLAMD) baseline with one process per core needs about 2 minutes "
        "per shifted factorisation of this matrix on one core (about 66 min per k point), beyond a bounded bench; "
        "dense LAPACK on all host threads is the faster CPU method at this fill and is timed instead",
    }


# ------------------------------------------------------------------ serving
def serving_leg(args, *, jobs: int, kscale: float, names: list[str], plain_loop: bool,
                plain_per_k: dict[str, float] | None = None) -> dict:
    """Mixed request load through the continuous-batching server.

    Jobs arrive as a Poisson stream (mean ``args.rate`` jobs/s, seed 4242).
    Each draws a bundle uniformly from ``names``, a k-path length from
    {20, 61, 121} scaled by ``kscale`` (points spread evenly over the
    Y-G-X-M path), a window half-width U[0.5, 2] eV about 0 and a block size
    from {32, 64}.  The sequential plain-CISS comparison runs the same
    completed jobs one after another (``plain_loop``), or is estimated from
    each bundle's measured plain seconds per k point (``plain_per_k``)."""
    from aospectrum_amd import synth
    from aospectrum_amd.band import ygxm_path
    from aospectrum_amd.ciss import CissSettings, band_solve
    from aospectrum_amd.data import EnergyInterval
    from aospectrum_amd.serving import BandJob, BandServer, latency_summary
    from aospectrum_amd.sparse import NativeBundle

    bundles = {name: synth.generate(synth.SPECS[name]) for name in names}
    rng = np.random.default_rng(4242)
    path = ygxm_path(40)
    plan = []
    t = 0.0
    for j in range(jobs):
        t += rng.exponential(1.0 / args.rate)
        name = names[rng.integers(len(names))]
        length = max(1, int(round(int(rng.choice([20, 61, 121])) * kscale)))
        idx = np.unique(np.linspace(0, path.n_points - 1, length).round().astype(int))
        half = float(rng.uniform(0.5, 2.0))
        block = int(rng.choice([32, 64]))
        st = CissSettings(block_size=block, contour_batch_size=16, conjugate_pairs=True, precision=args.precision,
                          refinement_steps=args.ir_steps, refinement_passes=args.passes,
                          refinement_policy=args.refinement_policy)
        plan.append((t, BandJob(name, path.fractional_points[idx], EnergyInterval(-half - 0.02, half + 0.02), st,
                                f"job{j}-{name}-{idx.size}k-h{half:.2f}-L{block}")))
    log(f"serving: {len(plan)} jobs over {names}, {sum(j.kpoints.shape[0] for _, j in plan)} k points")
    results, failures = [], []
    kb = args.k_batch or 4
    with BandServer(k_batch=kb, on_batch=lambda e: log(
            f"batch {e['bundle']}: {e['k_points']} k from {e['jobs']} job(s), {e['end'] - e['start']:.2f} s"
            + (f", error {e['error'][:80]}" if e["error"] else ""))) as server:
        for name, b in bundles.items():
            server.register(name, b)
        t0 = time.perf_counter()
        futures = []
        for arrival, job in plan:
            delay = t0 + arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            futures.append((job, server.submit(job)))
        for job, fut in futures:
            try:
                results.append(fut.result())
                log(f"done {job.job_id}: latency {results[-1].latency_s:.2f} s")
            except Exception as exc:  # noqa: BLE001 -- failed jobs are reported
                failures.append({"job": job.job_id, "error": str(exc)[:200]})
                log(f"failed {job.job_id}: {str(exc)[:120]}")
        batches = list(server.batch_log)
    summary = latency_summary(results) if results else {}
    done = {r.job_id for r in results}
    seq_rate, seq_kind = None, None
    if results and plain_loop:
        natives = {name: NativeBundle(b, device=0) for name, b in bundles.items()}
        os.environ["AOSX_PROGRESS"] = "1"
        t1 = time.perf_counter()
        nk = 0
        for _, job in plan:
            if job.job_id in done:
                plain = CissSettings(block_size=job.settings.block_size, precision="fp32_ir", refinement_steps=2,
                                     refinement_passes=args.passes)
                band_solve(natives[job.bundle_key], plain, job.interval, kpoints=job.kpoints)
                nk += job.kpoints.shape[0]
        seq_rate, seq_kind = nk / (time.perf_counter() - t1), "measured: the completed jobs one after another"
        os.environ.pop("AOSX_PROGRESS", None)
        for nb in natives.values():
            nb.close()
    elif results and plain_per_k:
        secs = sum(job.kpoints.shape[0] * plain_per_k[job.bundle_key] for _, job in plan if job.job_id in done)
        nk = sum(job.kpoints.shape[0] for _, job in plan if job.job_id in done)
        seq_rate = nk / secs if secs > 0 else None
        seq_kind = "estimated: sum over completed jobs of k points x the bundle's measured plain seconds per k point"
    return {
        "kpoints_per_s": summary.get("kpoints_per_s"),
        "p50_job_latency_s": summary.get("p50_latency_s"),
        "p99_job_latency_s": summary.get("p99_latency_s"),
        "jobs": len(plan),
        "jobs_completed": len(results),
        "jobs_failed": failures,
        "batches": len(batches),
        "batches_shared_by_jobs": sum(1 for b in batches if b["jobs"] > 1),
        "sequential_plain_kpoints_per_s": seq_rate,
        "sequential_plain_kind": seq_kind,
        "speedup_vs_sequential_plain": (summary["kpoints_per_s"] / seq_rate) if seq_rate and summary else None,
        "workload": f"Poisson {args.rate} jobs/s, {len(plan)} jobs, bundles {names}, k-path lengths "
        f"{{20,61,121}} x {kscale}, half-width U[0.5,2] eV, block {{32,64}}, k_batch {kb}",
    }


def plain_seconds_per_k(names: list[str], args, known: dict[str, float]) -> dict[str, float]:
    """Plain-configuration seconds per k point of each bundle (one k point at
    the middle of the path, window +/-1.25 eV, block 32), for the serving
    leg's sequential estimate."""
    from aospectrum_amd import synth
    from aospectrum_amd.band import ygxm_path
    from aospectrum_amd.ciss import CissSettings, band_solve
    from aospectrum_amd.data import EnergyInterval
    from aospectrum_amd.sparse import NativeBundle

    out = dict(known)
    k = ygxm_path(40).fractional_points[60:61]
    for name in names:
        if name in out:
            continue
        nb = NativeBundle(synth.generate(synth.SPECS[name]), device=0)
        try:
            plain = CissSettings(precision="fp32_ir", refinement_steps=2, refinement_passes=args.passes)
            band_solve(nb, plain, EnergyInterval(-1.27, 1.27), kpoints=k)  # captures the graphs
            t = time.perf_counter()
            band_solve(nb, plain, EnergyInterval(-1.27, 1.27), kpoints=k)
            out[name] = time.perf_counter() - t
        finally:
            nb.close()
        log(f"plain {name}: {out[name]:.2f} s per k point")
    return out


# --------------------------------------------------------------------- main
def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="bulk3d")
    ap.add_argument("--k-batch", type=int, default=None, help="k points per batch (default: per workload)")
    ap.add_argument("--plain-kpoints", type=int, default=2)
    ap.add_argument("--precision", default="bf16x3_ir")
    ap.add_argument("--far-threshold", type=float, default=0.9,
                    help="per_shift_ir: |sin theta| at and above which a contour point takes one BF16 product")
    ap.add_argument("--ir-steps", type=int, default=2)
    ap.add_argument("--refinement-policy", default="on_demand", choices=["fixed", "on_demand"])
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--cpu-shifts", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the bounded serving leg")
    ap.add_argument("--serving", action="store_true", help="only the serving load, with a measured plain loop")
    ap.add_argument("--serving-bundles", default="small2d,stress2d,bulk3d,slab26k_a,slab26k_b")
    ap.add_argument("--jobs", type=int, default=12)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--kscale", type=float, default=0.1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: the Band path batches k points on a single MI355X "
            "and is not distributed across GPUs (run with --gpus 1)",
            file=sys.stderr,
        )
        return 2
    if args.serving:
        names = args.serving_bundles.split(",")
        leg = serving_leg(args, jobs=args.jobs, kscale=args.kscale, names=names, plain_loop=True)
        line = {
            "metric": "serving: p50/p99 job latency and aggregate k-points/s under mixed request load",
            "value": leg["kpoints_per_s"], "unit": "k-points/s", "n_gpus": 1, "higher_is_better": True,
            **leg, "config": {"workload": leg["workload"], "precision": args.precision,
                              "refinement_policy": args.refinement_policy},
            "data": "synthetic (seeded generator)",
        }
        print(json.dumps(line), flush=True)
        return 0

    from aospectrum_amd import synth
    from aospectrum_amd.band import ygxm_path
    from aospectrum_amd.ciss import CissSettings, band_solve
    from aospectrum_amd.data import EnergyInterval
    from aospectrum_amd.sparse import NativeBundle

    # window (eV), block size, k points per batch, points per path segment
    workloads = {
        "bulk3d": ((-1.0, 1.0), 32, 4, 40),
        "small2d": ((-2.0, 2.0), 32, 4, 40),
        "stress2d": ((-1.5, 1.5), 32, 4, 20),
        "slab52k": ((-0.8, 0.8), 64, 1, 20),
        "slab26k_a": ((-2.0, 2.0), 32, 2, 20),
    }
    (lo, hi), block, k_default, seg = workloads[args.workload]
    spec = synth.SPECS[args.workload]
    interval = EnergyInterval(lo - 0.02, hi + 0.02)
    t0 = time.perf_counter()
    bundle = synth.generate(spec)
    gen_s = time.perf_counter() - t0
    log(f"generated {spec.name}: n={bundle.n_orbitals} blocks={bundle.operators.n_blocks} ({gen_s:.1f} s)")
    t0 = time.perf_counter()
    native = NativeBundle(bundle, device=0)
    setup_s = time.perf_counter() - t0
    info = native.info
    log(f"symbolic+upload {setup_s:.1f} s: nnz={info['nnz']} tiles={info['n_factor_tiles']} "
        f"flops/factor={info['factor_flops']:.3e}")
    path = ygxm_path(seg)
    K = args.k_batch or k_default
    mixed = CissSettings(block_size=block, contour_batch_size=16, k_batch=K, conjugate_pairs=True,
                         precision=args.precision, refinement_steps=args.ir_steps, refinement_passes=args.passes,
                         refinement_policy=args.refinement_policy, far_shift_threshold=args.far_threshold)
    # plain: the reference's configuration -- one shift per factorisation
    # launch, one k point at a time, FP32 LU with refinement
    plain = CissSettings(block_size=block, precision="fp32_ir", refinement_steps=2, refinement_passes=args.passes)

    def kslice(step):
        # timed steps walk the path from its start; warmup steps (negative)
        # come from its end
        start = (step * K) % path.n_points
        idx = [(start + i) % path.n_points for i in range(K)]
        return idx, path.fractional_points[idx]

    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tune = native.autotune(mixed.factorizations_per_launch, args.precision)
        tune["seconds"] = time.perf_counter() - t0
        log(f"autotune: group {tune['chosen_group']} "
            f"({', '.join(f'{g}: {ms:.0f} ms' for g, ms in tune['factor_ms'].items())})")
    # kernel timing is on from the warmup on, so the timed steps replay the
    # same captured launch sequences (with their event nodes) as the warmup
    native.kernel_stats(True)
    for w in range(args.warmup):
        band_solve(native, mixed, interval, kpoints=kslice(-(w + 1))[1])
        log(f"warmup step {w} done")
    native.kernel_stats(True)  # resets the counters
    lat, results, refined, peak_bytes = [], {}, [], 0
    t_start = time.perf_counter()
    for s in range(args.steps):
        idx, kp = kslice(s)
        t = time.perf_counter()
        res, binfo = band_solve(native, mixed, interval, kpoints=kp)
        lat.append(time.perf_counter() - t)
        peak_bytes = max(peak_bytes, binfo["peak_device_bytes"])
        refined += [idx[i] for i in binfo["refined"]]
        results.update(zip(idx, res))
        log(f"step {s}: {lat[-1]:.2f} s, counts {[r.energies_ev.size for r in res]}, "
            f"ranks {[r.subspace_rank for r in res]}, max res {max(r.maximum_residual_ev for r in res):.2e}, "
            f"stages {{{', '.join(f'{k}: {v:.2f}' for k, v in binfo['stage_seconds'].items() if v > 0.01)}}}")
    total = time.perf_counter() - t_start
    ks = native.kernel_stats(False)
    value = args.steps * K / total

    plain_info = {}
    if not args.no_plain and args.plain_kpoints > 0:
        pidx = sorted(results)[: args.plain_kpoints]
        if bundle.n_orbitals > 20000:
            os.environ["AOSX_PROGRESS"] = "1"  # a plain slab k point factors every shift in every pass
        native.kernel_stats(True)
        # plain warmup on a k point outside the timed set (captures its graphs)
        band_solve(native, plain, interval, kpoints=path.fractional_points[path.n_points - 1:])
        native.kernel_stats(True)
        t = time.perf_counter()
        pres, _ = band_solve(native, plain, interval, kpoints=path.fractional_points[pidx])
        pdt = time.perf_counter() - t
        pks = native.kernel_stats(False)
        os.environ.pop("AOSX_PROGRESS", None)
        same = [(i, r) for i, r in zip(pidx, pres) if r.energies_ev.size == results[i].energies_ev.size]
        plain_info = {
            "plain_kpoints_per_s": len(pidx) / pdt,
            "plain_seconds_per_kpoint": pdt / len(pidx),
            "plain_max_residual_ev": max(r.maximum_residual_ev for r in pres),
            "plain_vs_mixed_count_mismatch": [i for i, r in zip(pidx, pres)
                                              if r.energies_ev.size != results[i].energies_ev.size],
            "plain_vs_mixed_max_dE_ev": max((float(np.max(np.abs(r.energies_ev - results[i].energies_ev)))
                                             for i, r in same if r.energies_ev.size), default=0.0),
            "plain_schur_tflops": pks["schur_flops"] / max(pks["schur_ms"], 1e-9) / 1e9,
        }
        log(f"plain: {pdt / len(pidx):.2f} s per k point; mixed/plain = {value * pdt / len(pidx):.2f}x")

    gold = golden_counts(args.workload)
    checked = {i: (gold[i], int(r.energies_ev.size)) for i, r in results.items() if i in gold}
    count_check = {
        "fixture": f"tests/golden/{args.workload}_counts.json",
        "checked": len(checked),
        "timed": len(results),
        "mismatches": {str(i): {"oracle": c, "gpu": g} for i, (c, g) in checked.items() if c != g},
    }
    log(f"in-window counts: {len(checked)} of {len(results)} timed k points checked against FP64 inertia, "
        f"{len(count_check['mismatches'])} mismatches")

    native.close()  # frees the batch's resident factorisations before the serving leg
    serving = None
    if not args.no_serving:
        log("serving leg ...")
        try:
            names = ["small2d", "stress2d", "bulk3d", "slab26k_a"]
            known = {args.workload: plain_info["plain_seconds_per_kpoint"]} if plain_info else {}
            per_k = plain_seconds_per_k(names, args, known)
            serving = serving_leg(args, jobs=10, kscale=0.1, names=names, plain_loop=False, plain_per_k=per_k)
        except Exception as exc:  # noqa: BLE001 -- the main figure stands on its own
            serving = {"error": f"{type(exc).__name__}: {exc}"[:300]}
        log(f"serving: {serving}")

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(bundle, path.fractional_points[0], interval, args.cpu_shifts)
        log(f"cpu baseline {cpu['value']:.2e} k-points/s")

    kernel = schur_kernel(args.precision)
    achieved = ks["schur_flops"] / max(ks["schur_ms"], 1e-9) / 1e9  # TFLOP/s
    peak = PEAK_TFLOPS[args.precision]
    traffic = traffic_record(args.workload, args.precision, kernel)
    launches = max(ks["schur_launches"], 1)
    lat_ms = np.asarray(lat) * 1e3
    line = {
        "metric": "k-points/s at residual <=1e-4 eV, batched-mixed vs plain CISS; p50/p99 latency",
        "value": value,
        "unit": "k-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(total / args.steps * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x3_ir": "complex64+bf16x3", "bf16_ir": "complex64+bf16",
                  "per_shift_ir": "complex64+bf16/bf16x3 per shift"}.get(args.precision, "complex64"),
        "data": "synthetic (seeded generator, no datasets)",
        "config": {
            "workload": f"{spec.name}: n={bundle.n_orbitals}, {bundle.structure.n_atoms} atoms, "
            f"nnz={info['nnz']}, {path.n_points}-pt YGXM path, window [{lo},{hi}] eV",
            "integration_points": 32,
            "block_size": block,
            "moment_size": 4,
            "k_batch": K,
            "precision": args.precision,
            **({"far_shift_threshold": args.far_threshold} if args.precision == "per_shift_ir" else {}),
            "ir_steps": args.ir_steps,
            "refinement_policy": args.refinement_policy,
            "filter_passes": args.passes,
            "conjugate_pairs": True,
            "parallelism": "single GPU, k points batched on device",
        },
        "speedup_vs_plain": value * plain_info["plain_seconds_per_kpoint"] if plain_info else None,
        **plain_info,
        "p50_step_ms": float(np.percentile(lat_ms, 50)),
        "p99_step_ms": float(np.percentile(lat_ms, 99)),
        "max_residual_ev": max(r.maximum_residual_ev for r in results.values()),
        "count_check": count_check,
        "in_window_counts": {str(i): int(r.energies_ev.size) for i, r in sorted(results.items())},
        "subspace_ranks": {str(i): int(r.subspace_rank) for i, r in sorted(results.items())},
        "pass_ranks": {str(i): [int(r.pass_rank), int(r.pass_width)] for i, r in sorted(results.items())},
        "saturated": [i for i, r in sorted(results.items()) if not r.complete],
        "refined_kpoints": refined,
        "factor_tiles": info["n_factor_tiles"],
        "factor_fill_entries": info["n_factor_tiles"] * 64 * 64,
        "atom_fill_entries": info["atom_fill_entries"],
        "factor_bytes_per_matrix": info["factor_bytes"],
        "factor_flops_per_matrix": info["factor_flops"],
        "peak_device_bytes": peak_bytes,
        "roofline": {
            "bound": "mfma",
            "kernel": f"{kernel} ({args.precision}, trailing update of each step group)",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": None if traffic is None else traffic["bytes_per_dispatch"],
            "traffic_unit": "HBM bytes per dispatch (rocprofv3 FETCH_SIZE x scale + WRITE_SIZE)",
            "traffic_source": None if traffic is None else
            {k: traffic[k] for k in ("file", "measured_avg_ms", "source_commit") if k in traffic},
            "launches": ks["schur_launches"],
            "avg_launch_ms": ks["schur_ms"] / launches,
            "algorithmic_bytes_per_launch": ks["schur_bytes"] / launches,
            "mfma_issued_tflops": None if MFMA_PRODUCTS[args.precision] is None
            else achieved * MFMA_PRODUCTS[args.precision],
        },
        "kernel_ms": {k: v for k, v in ks.items() if k.endswith("_ms")},
        # streaming stages in achieved algorithmic GB/s (HBM peak 8,000)
        "assembly_gbps": ks["assemble_bytes"] / max(ks["assemble_ms"], 1e-9) / 1e6,
        "solve_gbps": ks["solve_bytes"] / max(ks["solve_ms"], 1e-9) / 1e6,
        "serving": serving,
        "cpu_baseline": cpu,
        "setup_seconds": {"generate": gen_s, "symbolic_and_upload": setup_s},
        "autotune": tune,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
