import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP path)")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz")) as d:
        return {k: d[k] for k in d.files}


@pytest.fixture(scope="session")
def golden_border():
    return load_golden("tiny_border")


@pytest.fixture(scope="session")
def golden_noborder():
    return load_golden("tiny_noborder")


@pytest.fixture(scope="session", params=["tiny_border", "tiny_noborder"])
def golden(request):
    return load_golden(request.param)


def value_range(g):
    """Range of the reference volume, estimated from the stored LolaBunny
    slice and samples (the normaliser of every tolerance)."""
    v = np.concatenate([g["bunny_slice"].ravel(), g["bunny_samples"]])
    return float(v.max() - v.min())


def compare_summary(vol, g, alg):
    """Max |diff| / range of a full (L,L,L) volume against the stored slice,
    samples and per-z sums of reference output ``alg``."""
    L = vol.shape[0]
    rng = value_range(g)
    d_slice = np.abs(vol[L // 2 - 1] - g[f"{alg}_slice"]).max()
    d_samp = np.abs(vol.reshape(-1)[g["sample_idx"]] - g[f"{alg}_samples"]).max()
    zsum = vol.reshape(L, -1).astype(np.float64).sum(1)
    d_z = np.abs(zsum - g[f"{alg}_zsum"]).max() / (L * L)
    return max(d_slice, d_samp) / rng, d_z / rng
