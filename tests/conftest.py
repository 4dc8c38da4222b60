import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "ref_ogrid_33x17.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    """Arrays produced by the compiled reference on a 33x17 O-grid (see make_ref_golden.py)."""
    with np.load(GOLDEN, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    m = d["meta"]
    d["n0"], d["n1"] = int(m[0]), int(m[1])
    d["dxi"], d["deta"] = float(m[2]), float(m[3])
    d["Re"], d["Pr"], d["dt"], d["Ri"] = float(m[4]), float(m[5]), float(m[6]), float(m[7])
    d["uinf"], d["vinf"], d["ar"] = float(m[8]), float(m[9]), float(m[10])
    return d


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
