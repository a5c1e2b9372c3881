"""Band-job serving on one MI355X with continuous batching.

The reference has no request loop (its ``serve`` is a static HTTP file server,
src/aospectrum/cli.py:561-571); this module is a new caller of the Band path.

* **Resident bundles.**  ``register`` runs the symbolic analysis once and keeps
  the bundle's block data and analysis in HBM for every later job.
* **Continuous batching.**  Every bundle has a worker thread with a queue of
  pending k points.  Whenever the worker is free it takes up to ``k_batch``
  waiting k points that share solver settings -- from *any* jobs, each with its
  own energy window -- and hands them to one ``aosx_band_solve_v1`` call, whose
  factorisation launches cover all their contour shifts.  Jobs that arrive
  while others run join the next batch.
* **Overlap on the device.**  Workers of different bundles run concurrently;
  each bundle owns its HIP stream, and the ctypes call releases the GIL, so
  their kernels overlap on the GPU.
* **Admission by fill memory.**  Before a batch runs, its device memory is
  estimated from the bundle's factor size (resident factorisations of every k
  and shift, plus vectors) and reserved against ``memory_budget_bytes``.  When
  the batch ends its reservation is released; the bundle's workspaces stay
  allocated as a cache (counted against the budget as ``cached``) so that its
  next batch reuses them, and the cache of an idle bundle is freed
  (``aosx_bundle_trim_v1``) as soon as another batch needs the room.  A batch
  that does not fit waits for memory to free up; a single k point that can
  never fit fails its job at submission.
* **Policy for saturated subspaces.**  As in the reference
  (src/aospectrum/band.py:239-243) a k point whose CISS search subspace
  saturates fails its job with ``SolverError``.

Each job reports per-k energies and residuals, its latency (arrival to last k
point) and how many batches it shared.
"""

from __future__ import annotations

import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from .ciss import CissSettings, band_solve
from .data import EnergyInterval, LocalizedOperatorBundle
from .errors import InputError, SolverError
from .sparse import NativeBundle


@dataclass
class BandJob:
    """One request: k points of one bundle in one absolute energy window."""

    bundle_key: str
    kpoints: np.ndarray
    interval: EnergyInterval
    settings: CissSettings = field(default_factory=CissSettings)
    job_id: str = ""


@dataclass
class JobResult:
    job_id: str
    bundle_key: str
    energies_ev: list[np.ndarray]
    maximum_residual_ev: list[float]
    arrival: float
    finished: float
    batches: int
    shared_batches: int
    error: str | None = None

    @property
    def latency_s(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Item:
    job: "_JobState"
    index: int


@dataclass
class _JobState:
    job: BandJob
    future: Future
    arrival: float
    energies: list = field(default_factory=list)
    residuals: list = field(default_factory=list)
    remaining: int = 0
    batches: int = 0
    shared: int = 0
    error: str | None = None


def _settings_key(s: CissSettings) -> tuple:
    return (
        s.integration_points,
        s.block_size,
        s.moment_size,
        s.ellipse_vertical_scale,
        s.endpoint_guard_ev,
        s.rank_relative_tolerance,
        s.spurious_relative_threshold,
        s.random_seed,
        s.precision,
        s.refinement_steps,
        s.refinement_passes,
        s.conjugate_pairs,
        s.refinement_policy,
        s.residual_target_ev,
        s.resident_limit,
    )


class _Resident:
    def __init__(self, key: str, native: NativeBundle) -> None:
        self.key = key
        self.native = native
        self.pending: deque[_Item] = deque()
        self.reserved = 0   # bytes held by the batch running now
        self.cached = 0     # bytes of workspaces kept after the last batch
        self.busy = False
        self.thread: threading.Thread | None = None


class BandServer:
    """Continuous-batching Band server for one GPU."""

    def __init__(
        self,
        *,
        device: int = 0,
        memory_budget_bytes: int = 200 * 2**30,
        k_batch: int = 4,
        on_batch=None,
    ) -> None:
        self.device = device
        self.budget = int(memory_budget_bytes)
        self.k_batch = int(k_batch)
        self.on_batch = on_batch
        self._lock = threading.Condition()
        self._bundles: dict[str, _Resident] = {}
        self._closed = False
        self._next_id = 0
        self.batch_log: list[dict] = []
        # bundles whose next batch waits for memory: (arrival of its oldest
        # k point, bytes it needs)
        self._waiting: dict[str, tuple[float, int]] = {}

    # ----------------------------------------------------------- bundles
    def register(self, key: str, bundle: LocalizedOperatorBundle, *, autotune: CissSettings | None = None) -> dict:
        """Analyse and upload a bundle once; it stays resident.  With
        ``autotune`` the bundle's step-group size is tuned for batches of
        those settings (``k_batch`` k points) before it takes jobs."""
        native = NativeBundle(bundle, device=self.device)
        res = _Resident(key, native)
        if autotune is not None:
            # tune for the largest batch of these settings the budget admits,
            # then give the tuning workspace back
            per_k = autotune.factorizations_per_launch // autotune.k_batch
            k = autotune.k_batch
            while k > 1 and self._estimate(native, autotune, k) > self.budget:
                k -= 1
            native.autotune(k * per_k, autotune.precision)
            native.trim()
        with self._lock:
            if key in self._bundles:
                native.close()
                raise InputError(f"bundle {key!r} is already registered")
            self._bundles[key] = res
            res.thread = threading.Thread(target=self._worker, args=(res,), name=f"band-{key}", daemon=True)
            res.thread.start()
        return native.info

    def estimate_bytes(self, key: str, settings: CissSettings, n_k: int) -> int:
        """Device bytes one batch of ``n_k`` k points needs beyond residency."""
        return self._estimate(self._bundles[key].native, settings, n_k)

    @staticmethod
    def _estimate(native: NativeBundle, settings: CissSettings, n_k: int) -> int:
        info = native.info
        n = int(info["n"])
        nfac = settings.integration_points // (2 if settings.conjugate_pairs else 1)
        slots = n_k * (nfac if settings.refinement_passes else min(settings.contour_batch_size, nfac))
        width = settings.block_size * settings.moment_size + 2 * settings.block_size
        vectors = n_k * n * width * 8 * 12 + slots * n * width * 8 * 3
        matrices = n_k * int(info["nnz"]) * 16
        # BF16 precisions: split copies of one step group's panel tiles per
        # factorisation (at most 2 x 32 steps x n_diag tiles of 48 KiB)
        split_mode = settings.precision in ("bf16x3_ir", "bf16_ir", "per_shift_ir")
        split = 2 * 32 * int(info["n_diagonal_tiles"]) * 49152 if split_mode else 0
        return int(slots * (int(info["factor_bytes"]) + split) + vectors + matrices)

    # -------------------------------------------------------------- jobs
    def submit(self, job: BandJob) -> Future:
        if job.bundle_key not in self._bundles:
            raise InputError(f"unknown bundle {job.bundle_key!r}")
        k = np.asarray(job.kpoints, dtype=np.float64).reshape(-1, 3)
        if k.shape[0] == 0:
            raise InputError("a Band job needs at least one k point")
        fut: Future = Future()
        with self._lock:
            if self._closed:
                raise SolverError("server is closed")
            if not job.job_id:
                job.job_id = f"job-{self._next_id}"
                self._next_id += 1
            job.kpoints = k
            state = _JobState(job, fut, time.perf_counter(), [None] * k.shape[0], [None] * k.shape[0], k.shape[0])
            if self.estimate_bytes(job.bundle_key, job.settings, 1) > self.budget:
                fut.set_exception(SolverError("one k point of this job exceeds the device memory budget"))
                return fut
            res = self._bundles[job.bundle_key]
            for i in range(k.shape[0]):
                res.pending.append(_Item(state, i))
            self._lock.notify_all()
        return fut

    def _take_batch(self, res: _Resident) -> list[_Item]:
        """Up to k_batch pending items sharing the first item's settings."""
        first = res.pending[0]
        key = _settings_key(first.job.job.settings)
        take, keep = [], deque()
        while res.pending:
            item = res.pending.popleft()
            if item.job.error is None and len(take) < self.k_batch and _settings_key(item.job.job.settings) == key:
                take.append(item)
            elif item.job.error is None:
                keep.append(item)
        res.pending = keep
        return take

    def _admit(self, res: _Resident, need: int, arrival: float) -> bool:
        """Hold the lock.  Admission in order of arrival across bundles: a
        batch whose oldest k point arrived after that of another bundle's
        waiting batch starts only if both fit beside the running batches
        (idle caches can be freed), so that a bundle which finishes a batch
        and takes the next one at once cannot starve the others."""
        older = sum(n for k, (a, n) in self._waiting.items() if k != res.key and a < arrival)
        if older:
            running = sum(r.reserved for r in self._bundles.values() if r is not res)
            if running + need + older > self.budget:
                return False
        return self._reserve(res, need)

    def _reserve(self, res: _Resident, need: int) -> bool:
        """Hold the lock; make room for `need` bytes or return False.

        A bundle's own cache is taken over by its batch (the workspaces only
        grow, so the batch holds the larger of the two); caches of idle bundles
        are freed one at a time until the batch fits."""
        def used() -> int:
            return sum(r.reserved + r.cached for r in self._bundles.values() if r is not res)

        def take() -> bool:
            res.reserved = max(need, res.cached)
            res.cached = 0
            return True

        if used() + max(need, res.cached) <= self.budget:
            return take()
        for other in self._bundles.values():
            if other is not res and not other.busy and other.cached:
                other.native.trim()
                other.cached = 0
                if used() + max(need, res.cached) <= self.budget:
                    return take()
        if res.cached and used() + need <= self.budget:
            # the bundle's own cache is larger than this batch needs
            res.native.trim()
            res.cached = 0
            return take()
        return False

    def memory_in_use(self) -> dict[str, tuple[int, int]]:
        """Per bundle: (bytes reserved by a running batch, bytes cached)."""
        with self._lock:
            return {k: (r.reserved, r.cached) for k, r in self._bundles.items()}

    def _worker(self, res: _Resident) -> None:
        while True:
            with self._lock:
                while not self._closed and not res.pending:
                    self._lock.wait()
                if self._closed and not res.pending:
                    return
                batch = self._take_batch(res)
                if not batch:
                    continue
                settings = batch[0].job.job.settings
                # a batch that could never fit the budget is cut down (one k
                # point always fits: submit() rejects jobs where it does not)
                while len(batch) > 1 and self.estimate_bytes(res.key, settings, len(batch)) > self.budget:
                    res.pending.appendleft(batch.pop())
                need = self.estimate_bytes(res.key, settings, len(batch))
                arrival = min(it.job.arrival for it in batch)
                self._waiting[res.key] = (arrival, need)
                while not self._admit(res, need, arrival):
                    self._lock.wait(timeout=0.05)
                del self._waiting[res.key]
                res.busy = True
            self._run(res, batch, settings)
            with self._lock:
                res.busy = False
                res.cached, res.reserved = res.reserved, 0
                self._lock.notify_all()

    def _run(self, res: _Resident, batch: list[_Item], settings: CissSettings) -> None:
        kp = np.stack([it.job.job.kpoints[it.index] for it in batch])
        iv = np.array([[it.job.job.interval.lower_ev, it.job.job.interval.upper_ev] for it in batch])
        jobs = {id(it.job) for it in batch}
        started = time.perf_counter()
        try:
            run_settings = CissSettings(**{**_asdict(settings), "k_batch": len(batch)})
            results, info = band_solve(res.native, run_settings, EnergyInterval(iv[:, 0].min(), iv[:, 1].max()),
                                       kpoints=kp, intervals=iv)
            failure = None
        except Exception as exc:  # noqa: BLE001 -- reported per job
            results, info, failure = None, {}, f"{type(exc).__name__}: {exc}"
        ended = time.perf_counter()
        entry = {"bundle": res.key, "k_points": len(batch), "jobs": len(jobs), "start": started, "end": ended,
                 "error": failure}
        self.batch_log.append(entry)
        if self.on_batch is not None:
            self.on_batch(entry)
        with self._lock:
            for pos, it in enumerate(batch):
                st = it.job
                st.batches += 1
                st.shared += len(jobs) > 1
                if failure is not None:
                    st.error = st.error or failure
                else:
                    r = results[pos]
                    if not r.complete:
                        st.error = st.error or "SolverError: CISS search subspace is saturated; increase block_size or moment_size"
                    st.energies[it.index] = r.energies_ev
                    st.residuals[it.index] = r.maximum_residual_ev
                st.remaining -= 1
                if st.error is not None or st.remaining == 0:
                    self._finish(st, ended)

    def _finish(self, st: _JobState, ended: float) -> None:
        if st.future.done():
            return
        if st.error is not None:
            # drop the job's remaining queued k points
            for res in self._bundles.values():
                res.pending = deque(it for it in res.pending if it.job is not st)
            st.future.set_exception(SolverError(f"{st.job.job_id}: {st.error}"))
            return
        st.future.set_result(
            JobResult(st.job.job_id, st.job.bundle_key, st.energies, st.residuals, st.arrival, ended,
                      st.batches, st.shared)
        )

    def close(self) -> None:
        with self._lock:
            self._closed = True
            self._lock.notify_all()
        for res in list(self._bundles.values()):
            if res.thread is not None:
                res.thread.join()
            res.native.close()
        self._bundles.clear()

    def __enter__(self) -> "BandServer":
        return self

    def __exit__(self, *exc) -> None:
        self.close()


def _asdict(settings: CissSettings) -> dict:
    return {name: getattr(settings, name) for name in CissSettings.__dataclass_fields__}


def latency_summary(results: list[JobResult]) -> dict:
    lat = np.array([r.latency_s for r in results])
    nk = sum(len(r.energies_ev) for r in results)
    wall = max(r.finished for r in results) - min(r.arrival for r in results)
    return {
        "jobs": len(results),
        "k_points": nk,
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "kpoints_per_s": nk / wall if wall > 0 else float("nan"),
        "wall_s": wall,
    }
