"""Shared fixtures; the ``gpu`` marker selects tests that need an MI355X."""

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def small2d():
    from aospectrum_amd import synth

    return synth.generate(synth.SPECS["small2d"])


@pytest.fixture(scope="session")
def small2d_pattern(small2d):
    from oracle import bloch

    return bloch.Pattern.build(small2d.operators, small2d.basis.atom_offsets, small2d.n_orbitals)


@pytest.fixture(scope="session")
def tiny3d():
    """A 2x2x3-atom periodic 3D system (156 orbitals) whose cell is smaller
    than twice the cutoff, so several lattice shifts fall on one atom pair."""
    from aospectrum_amd import synth

    spec = synth.scaled_spec("bulk3d", grid=(2, 2, 3), cell=(4.8, 4.8, 7.2), seed=77, gap=1.0)
    return synth.generate(spec)


def rng(seed=0):
    return np.random.default_rng(seed)
