"""Host-side formats and control of the Band path (CPU only): bundle I/O,
k paths, checkpoints, track assignment, outputs, TOML configuration and the
resume contract -- the user-visible contract of ``aospectrum band``."""

import csv
import json
from pathlib import Path

import numpy as np
import pytest

from aospectrum_amd import load_bundle, write_bundle
from aospectrum_amd.band import (
    BandPoint,
    BandResult,
    KPath,
    assign_track_ids,
    read_kpath,
    read_point,
    resolve_energy_reference,
    write_band_output,
    write_point,
    ygxm_path,
)
from aospectrum_amd.cli import _prepare_work, main, read_band_config
from aospectrum_amd.data import EnergyInterval
from aospectrum_amd.errors import InputError

GOLDEN = Path(__file__).parent / "golden"


def test_bundle_round_trip(small2d, tmp_path):
    root = write_bundle(small2d, tmp_path / "b.aobundle")
    back = load_bundle(root)
    assert json.loads((root / "manifest.json").read_text())["schema"] == "aospectrum.operator-bundle/v1"
    for name in ("row_atom", "col_atom", "lattice_shift", "block_offsets", "block_shapes", "h_values", "s_values"):
        assert np.array_equal(getattr(back.operators, name), getattr(small2d.operators, name))
    assert back.n_orbitals == small2d.n_orbitals
    assert back.provenance["spec"]["name"] == "small2d"
    with pytest.raises(InputError):
        write_bundle(small2d, root)


def test_kpath_fixture():
    path = read_kpath(GOLDEN / "YGXM-121.json")
    assert path.n_points == 121
    assert list(path.node_indices) == [0, 40, 80, 120]
    assert path.node_labels == ("Y", "G", "X", "M")
    assert np.allclose(path.fractional_points[40], 0.0)
    assert np.array_equal(path.fractional_points, ygxm_path(40).fractional_points)
    d = path.distances(np.diag([20.0, 20.0, 30.0]))
    assert d[0] == 0.0 and np.all(np.diff(d) > 0)
    # Y -> G is |b2|/2 = pi / 20 per inverse Angstrom
    assert abs(d[40] - np.pi / 20.0) < 1e-12


def test_kpath_explicit_and_invalid(tmp_path):
    p = tmp_path / "k.json"
    p.write_text(json.dumps({"fractional_points": [[0, 0, 0], [0.5, 0, 0]], "node_indices": [0, 1],
                             "labels": ["G", "X"]}))
    assert read_kpath(p).n_points == 2
    p.write_text(json.dumps({"labels": ["G"], "fractional_points": [[0, 0, 0]], "node_indices": [0]}))
    with pytest.raises(InputError):
        read_kpath(p)
    with pytest.raises(InputError):
        KPath(np.zeros((3, 3)), np.array([0, 2, 1]), ("a", "b", "c"))


def test_energy_reference(small2d):
    assert resolve_energy_reference(small2d, "input_zero") == 0.0
    assert resolve_energy_reference(small2d, "fixed", 1.5) == 1.5
    with pytest.raises(InputError):
        resolve_energy_reference(small2d, "fixed")
    with pytest.raises(InputError):
        resolve_energy_reference(small2d, "bundle_fermi")  # synthetic filling has no Fermi level
    with pytest.raises(InputError):
        resolve_energy_reference(small2d, "vacuum")


def _point(i, energies):
    return BandPoint(i, np.array([0.1 * i, 0.0, 0.0]), np.asarray(energies, np.float32), 1e-6,
                     {"native_numeric_factorization": 0.5}, 0.01, 123, ("w",))


def test_checkpoint_round_trip(tmp_path):
    p = _point(3, [-0.5, 0.25])
    write_point(tmp_path / "k-000003.npz", p)
    back = read_point(tmp_path / "k-000003.npz", index=3, k_fractional=p.k_fractional)
    assert np.array_equal(back.energies_ev, p.energies_ev)
    assert back.stage_seconds == p.stage_seconds and back.peak_gpu_memory_bytes == 123
    assert back.warnings == ("w",)
    assert read_point(tmp_path / "missing.npz", index=0, k_fractional=p.k_fractional) is None
    with pytest.raises(InputError):
        read_point(tmp_path / "k-000003.npz", index=4, k_fractional=p.k_fractional)


def test_track_ids_follow_minimum_cost_matching():
    tracks = assign_track_ids([np.array([0.0, 1.0]), np.array([0.4, 0.6]), np.array([1.0, 0.0]),
                               np.array([0.0, 1.0, 5.0])])
    assert [t.tolist() for t in tracks] == [[0, 1], [0, 1], [1, 0], [0, 1, 2]]


def test_band_outputs(tmp_path):
    path = KPath(np.array([[0, 0, 0], [0.25, 0, 0], [0.5, 0, 0]], float), np.array([0, 2]), ("G", "X"))
    points = tuple(_point(i, e) for i, e in enumerate([[-0.5, 0.3], [-0.45, 0.2], [-0.4]]))
    result = BandResult(path, EnergyInterval(-1, 1), EnergyInterval(-1.02, 1.02), 0.1,
                        np.eye(3) * 10.0, points, 2.5)
    root = write_band_output(result, tmp_path / "out")
    rows = list(csv.reader((root / "bands.csv").open()))
    assert rows[0] == ["k_index", "k_distance_inv_angstrom", "track_id", "energy_ev"]
    assert len(rows) == 1 + 5
    assert float(rows[1][3]) == pytest.approx(-0.5 - 0.1, abs=1e-7)  # reference subtracted
    summary = json.loads((root / "summary.json").read_text())
    for key in ("backend", "scalar_dtype", "kpoint_count", "energy_interval_ev", "reference_energy_ev",
                "maximum_eigenpair_residual_ev", "elapsed_seconds", "assembly_seconds",
                "solver_stage_seconds", "peak_gpu_memory_bytes", "warnings"):
        assert key in summary
    assert summary["kpoint_count"] == 3 and summary["warnings"] == ["w"]
    assert summary["solver_stage_seconds"]["native_numeric_factorization"] == pytest.approx(1.5)
    assert (root / "band.png").is_file()


def _toml(tmp_path, extra=""):
    (tmp_path / "k.json").write_text((GOLDEN / "YGXM-121.json").read_text())
    cfg = tmp_path / "band.toml"
    cfg.write_text(f"""
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
{extra}
[output]
directory = "out"
""")
    return cfg


def test_read_band_config(tmp_path):
    cfg = read_band_config(_toml(tmp_path, 'precision = "bf16x3_ir"\nk_batch = 4\nconjugate_pairs = true\n'
                                           "refinement_passes = 2\nrefinement_steps = 2\n"
                                           'refinement_policy = "on_demand"\nresidual_target_ev = 5e-5\n'
                                           "autotune = true"))
    assert cfg.settings.autotune and cfg.settings.factorizations_per_launch == 4 * 16
    assert cfg.settings.refinement_policy == "on_demand" and cfg.settings.residual_target_ev == 5e-5
    assert cfg.settings.contour_batch_size == 16 and cfg.settings.integration_points == 32
    assert cfg.settings.precision == "bf16x3_ir" and cfg.settings.k_batch == 4 and cfg.settings.conjugate_pairs
    assert cfg.kpath.n_points == 121 and cfg.output == (tmp_path / "out").resolve()
    assert cfg.work.name == ".out.work"
    with pytest.raises(InputError):
        read_band_config(_toml(tmp_path, "block_size = \"x\""))
    with pytest.raises(InputError):
        read_band_config(_toml(tmp_path, "conjugate_pairs = 1"))
    with pytest.raises(InputError):
        read_band_config(_toml(tmp_path, 'refinement_policy = "sometimes"'))


def test_resume_contract(tmp_path):
    work = tmp_path / ".out.work"
    _prepare_work(work, {"kind": "band", "a": 1}, resume=False)
    with pytest.raises(InputError, match="use --resume"):
        _prepare_work(work, {"kind": "band", "a": 1}, resume=False)
    _prepare_work(work, {"kind": "band", "a": 1}, resume=True)
    with pytest.raises(InputError, match="another calculation"):
        _prepare_work(work, {"kind": "band", "a": 2}, resume=True)


def test_cli_exit_status_on_input_error(tmp_path, capsys):
    assert main(["band", str(tmp_path / "missing.toml")]) == 2
    assert "aospectrum:" in capsys.readouterr().err
