"""pytest configuration: the `gpu` marker and shared helpers."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


def rel_max(ref, cmp) -> float:
    """max|ref - cmp| / max|ref| -- the metric of corrct's tests/test_projectors.py:135."""
    ref = np.asarray(ref, dtype=np.float64)
    cmp = np.asarray(cmp, dtype=np.float64)
    return float(np.max(np.abs(ref - cmp)) / max(np.max(np.abs(ref)), 1e-300))


def rel_l2(ref, cmp) -> float:
    ref = np.asarray(ref, dtype=np.float64)
    cmp = np.asarray(cmp, dtype=np.float64)
    return float(np.linalg.norm(ref - cmp) / max(np.linalg.norm(ref), 1e-300))


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "skimage_golden.npz"))
