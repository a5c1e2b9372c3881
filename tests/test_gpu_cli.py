"""``aospectrum band band.toml`` end to end on the GPU: bundle directory,
k-path JSON and TOML in, ``bands.csv`` / ``summary.json`` / checkpoints out,
with the MI355X solver keys (batched BF16x3, refinement on demand,
autotuned step groups), checked against the FP64 dense oracle; then the
output and resume contract of the reference (cli.py:270-303)."""

import csv
import json
import sys

import numpy as np
import pytest

from aospectrum_amd import write_bundle
from aospectrum_amd.band import ygxm_path
from aospectrum_amd.cli import main
from oracle import dense

pytestmark = pytest.mark.gpu


def say(*parts):
    print("[gpu-cli]", *parts, flush=True)
    sys.stdout.flush()


def test_band_cli_end_to_end(small2d, small2d_pattern, tmp_path, capsys):
    write_bundle(small2d, tmp_path / "b.aobundle")
    path = ygxm_path(3)  # 10 points
    (tmp_path / "k.json").write_text(json.dumps({
        "fractional_points": path.fractional_points.tolist(),
        "node_indices": [int(i) for i in path.node_indices],
        "labels": list(path.node_labels),
    }))
    cfg = tmp_path / "band.toml"
    cfg.write_text("""
[input]
bundle = "b.aobundle"
[band]
kpath = "k.json"
energy_interval_ev = [-2.0, 2.0]
energy_reference = "input_zero"
[compute]
devices = [0]
[solver]
contour_batch_size = 16
conjugate_pairs = true
precision = "bf16x3_ir"
refinement_steps = 2
refinement_passes = 2
refinement_policy = "on_demand"
k_batch = 4
autotune = true
[output]
directory = "out"
""")
    say("aospectrum band")
    assert main(["band", str(cfg)]) == 0
    out = tmp_path / "out"
    summary = json.loads((out / "summary.json").read_text())
    assert summary["backend"] == "ciss-hip" and summary["kpoint_count"] == path.n_points
    assert summary["maximum_eigenpair_residual_ev"] <= 1e-4
    rows = list(csv.DictReader((out / "bands.csv").open()))
    by_k = {}
    for r in rows:
        by_k.setdefault(int(r["k_index"]), []).append(float(r["energy_ev"]))
    for i, k in enumerate(path.fractional_points):
        h, s = small2d_pattern.assemble(small2d.operators, k)
        # bands.csv lists the solve interval (the display window widened by
        # the 0.02 eV continuity guard), relative to the reference energy 0
        ref = dense.window_spectrum(small2d_pattern.csr(h).toarray(), small2d_pattern.csr(s).toarray(), -2.02, 2.02)
        got = np.sort(np.asarray(by_k.get(i, []), dtype=np.float64))
        say(f"k {i}: {got.size} states (oracle {ref.size})")
        assert got.size == ref.size
        assert np.max(np.abs(got - ref), initial=0.0) < 1e-4
    # a published output is never overwritten (cli.py:270-272)
    assert main(["band", str(cfg)]) == 2
    assert "output directory is not empty" in capsys.readouterr().err

    # resume: an interrupted run (work directory with the first five k points
    # checkpointed) completes the rest and publishes the same bands
    from aospectrum_amd import load_bundle
    from aospectrum_amd.band import BandCalculation, write_point
    from aospectrum_amd.cli import _band_specification, _prepare_work, read_band_config

    cfg2 = tmp_path / "band2.toml"
    cfg2.write_text(cfg.read_text().replace('directory = "out"', 'directory = "out2"'))
    config = read_band_config(cfg2)
    bundle = load_bundle(config.bundle)
    _prepare_work(config.work, _band_specification(config, bundle), resume=False)
    calc = BandCalculation(bundle, config.kpath, config.interval, 0.0, config.settings)
    try:
        for point in calc.solve_points([0, 1, 2, 3, 4]):
            write_point(config.work / f"k-{point.index:06d}.npz", point)
    finally:
        calc.close()
    say("aospectrum band --resume")
    assert main(["band", str(cfg2), "--resume"]) == 0
    first = list(csv.DictReader((out / "bands.csv").open()))
    second = list(csv.DictReader((tmp_path / "out2" / "bands.csv").open()))
    assert [(r["k_index"], r["track_id"]) for r in first] == [(r["k_index"], r["track_id"]) for r in second]
    assert max(abs(float(a["energy_ev"]) - float(b["energy_ev"])) for a, b in zip(first, second)) <= 1e-6
    capsys.readouterr()
