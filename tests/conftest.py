import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


TRAP_KEYS = ["trap", "kappa", "Vdc", "Vrf", "r0", "z0", "freq_rf", "tof", "E_kick", "coulomb",
             "coulomb_scale"]


def golden_fields(arr, make):
    """Field struct (oracle.Fields or simbuca_amd.Fields via `make`) from a
    golden fixture's `fields` vector."""
    kw = {}
    for k, v in zip(TRAP_KEYS, arr.tolist()):
        kw[k] = int(v) if k in ("trap", "tof", "coulomb") else float(v)
    return make(**kw)


@pytest.fixture(scope="session")
def rng():
    return np.random.Generator(np.random.PCG64(12345))
