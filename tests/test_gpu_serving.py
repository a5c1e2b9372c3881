"""Serving loop on the GPU: a job gives the same result alone and batched
with other jobs (different windows and bundles), saturated jobs fail without
disturbing the rest, and memory admission trims idle bundles."""

import sys

import numpy as np
import pytest

from aospectrum_amd import synth
from aospectrum_amd.band import ygxm_path
from aospectrum_amd.ciss import CissSettings
from aospectrum_amd.data import EnergyInterval
from aospectrum_amd.errors import SolverError
from aospectrum_amd.serving import BandJob, BandServer, latency_summary

pytestmark = pytest.mark.gpu

SETTINGS = CissSettings(contour_batch_size=16, conjugate_pairs=True, precision="bf16x3_ir",
                        refinement_steps=2, refinement_passes=2)
# Same job alone and in a mix: identical arithmetic per k point, so energies
# agree to float32 rounding of the reductions; stated tolerance 1e-6 eV.
TOL_EV = 1e-6


def say(*parts):
    print("[gpu-serving]", *parts, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def bundles(small2d):
    proxy = synth.generate(synth.scaled_spec("bulk3d", grid=(6, 6, 6), cell=(14.4, 14.4, 14.4), gap=1.8))
    return {"small2d": small2d, "proxy3d": proxy}


def _job_a():
    path = ygxm_path(20)
    return BandJob("small2d", path.fractional_points[:6], EnergyInterval(-2.02, 2.02), SETTINGS, "A")


def test_job_same_alone_and_in_mix(bundles):
    say("alone")
    with BandServer(k_batch=4) as server:
        server.register("small2d", bundles["small2d"])
        alone = server.submit(_job_a()).result(timeout=300)
    path = ygxm_path(20)
    say("mix")
    with BandServer(k_batch=4) as server:
        server.register("small2d", bundles["small2d"])
        server.register("proxy3d", bundles["proxy3d"])
        futures = [
            server.submit(BandJob("small2d", path.fractional_points[30:33], EnergyInterval(-1.52, 1.52), SETTINGS, "B")),
            server.submit(_job_a()),
            server.submit(BandJob("proxy3d", path.fractional_points[10:13], EnergyInterval(-1.02, 1.02), SETTINGS, "C")),
            server.submit(BandJob("small2d", path.fractional_points[50:55], EnergyInterval(-2.02, 2.02), SETTINGS, "D")),
        ]
        results = [f.result(timeout=300) for f in futures]
        log = list(server.batch_log)
    mixed = results[1]
    say("batches", [(b["bundle"], b["k_points"], b["jobs"]) for b in log])
    assert any(b["jobs"] > 1 for b in log), "no batch combined several jobs"
    assert mixed.shared_batches >= 1
    for e0, e1, r in zip(alone.energies_ev, mixed.energies_ev, mixed.maximum_residual_ev):
        assert e0.size == e1.size
        assert np.max(np.abs(e0 - e1), initial=0.0) <= TOL_EV
        assert r <= 1e-4
    summary = latency_summary(results)
    say(summary)
    assert summary["k_points"] == 3 + 6 + 3 + 5


def test_saturated_job_fails_alone(bundles):
    path = ygxm_path(20)
    with BandServer(k_batch=4) as server:
        server.register("small2d", bundles["small2d"])
        bad = server.submit(BandJob("small2d", path.fractional_points[:2], EnergyInterval(-9.0, 9.0), SETTINGS, "wide"))
        good = server.submit(BandJob("small2d", path.fractional_points[2:4], EnergyInterval(-2.02, 2.02), SETTINGS, "ok"))
        with pytest.raises(SolverError, match="saturated"):
            bad.result(timeout=300)
        assert len(good.result(timeout=300).energies_ev) == 2


def test_admission_trims_idle_bundles(bundles):
    path = ygxm_path(20)
    with BandServer(k_batch=2) as server:
        server.register("small2d", bundles["small2d"])
        server.register("proxy3d", bundles["proxy3d"])
        need = max(server.estimate_bytes("small2d", SETTINGS, 2), server.estimate_bytes("proxy3d", SETTINGS, 2))
        server.budget = int(1.2 * need)  # room for one batch at a time
        futs = [
            server.submit(BandJob("small2d", path.fractional_points[:2], EnergyInterval(-2.02, 2.02), SETTINGS, "s")),
            server.submit(BandJob("proxy3d", path.fractional_points[:2], EnergyInterval(-1.02, 1.02), SETTINGS, "p")),
        ]
        for f in futs:
            assert all(r <= 1e-4 for r in f.result(timeout=300).maximum_residual_ev)
