import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
sys.path.insert(0, os.path.join(ROOT, "amr-demo_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle.Oracle()


@pytest.fixture(scope="session")
def hip():
    import amrhip
    amrhip.lib()
    return amrhip


def golden_meshes():
    return sorted(glob.glob(os.path.join(GOLDEN, "mesh_*.npz")))


def parity_err(got, ref):
    """max over cells and fields of |got - ref| / max(|ref|, ||ref_field||_inf)
    (the floor keeps free-stream cells, whose residual is round-off, from
    dominating)"""
    r4 = np.asarray(ref).reshape(-1, 4)
    g4 = np.asarray(got).reshape(-1, 4)
    floor = np.maximum(np.abs(r4), np.abs(r4).max(axis=0, keepdims=True))
    floor = np.where(floor == 0, 1.0, floor)
    return float(np.max(np.abs(g4 - r4) / floor))
