"""Shared test setup.

Markers: `gpu` tests run the HIP product library on an MI355X; everything
else runs on the CPU (oracle against the golden fixtures, host logic, and the
C-ABI exports).  The oracle under oracle/ is only ever the checker.
"""
import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE_LIB = os.path.join(ROOT, "oracle", "libsps_oracle.so")
PRODUCT_LIB = os.path.join(ROOT, "stress-particle-sph_amd", "libsps_hip.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP product library)")


def load_pkg():
    return importlib.import_module("stress-particle-sph_amd")


@pytest.fixture(scope="session")
def sps():
    return load_pkg()


@pytest.fixture(scope="session")
def oracle_lib():
    if not os.path.exists(ORACLE_LIB):
        pytest.fail("oracle/libsps_oracle.so is not built: run __graft_entry__.build()")
    return ORACLE_LIB


def deck_dir(name):
    return os.path.join(ROOT, "stress-particle-sph_amd", "decks", name)


def golden(name):
    path = os.path.join(GOLDEN, name + ".npz")
    return np.load(path, allow_pickle=False)


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def describe_diff(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    d = np.abs(a - b)
    scale = max(np.abs(a).max(), 1e-300)
    return f"max|diff| {d.max():.3e}, rel {d.max() / scale:.3e}, n differing {(a != b).sum()}"


def open_varied_h(sps, deck, lib_path, factor=1.05):
    """A deck's initial state with every third particle's smoothing length
    scaled by `factor`: a case where pairs have different h, which the
    reference handles through mhsml = (h_i + h_j)/2."""
    import importlib
    import numpy as np
    setup = importlib.import_module("stress-particle-sph_amd.setup")
    parts = setup.build_particles(deck)
    st = sps.initial_state(parts)
    st["hsml"] = st["hsml"] * np.where(np.arange(len(st["hsml"])) % 3 == 0, factor, 1.0)
    sim = sps.Simulation(sps.make_params(deck, parts), lib_path)
    sim.set_state(st)
    sim.deck, sim.parts = deck, parts
    return sim
