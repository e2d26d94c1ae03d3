import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


def golden(name):
    """Load a fixture recorded from the compiled reference (tests/golden/make_golden.py)."""
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def decided_steps(stats, scene, band=0.1):
    """Leading DFSPH steps of a reference run whose loops all stopped clear of
    their thresholds: the reference's final average density error below
    (1 - band) eta and its divergence error below (1 - band) etaV / dt
    (DFSPHSolver.cpp:446, :746). Past the first step within the band, an
    iteration count is decided by the last bits of the sums: on the 40-step
    golden run, the oracle's per-step divergence error in the reference's own
    summation order already differs from the compiled reference's by up to
    4.6% (FP32) and 8.4% (FP64) at step 20. stats columns: t, dt, iterations,
    divergence iterations, density error, divergence error."""
    c = scene["Configuration"]
    n = 0
    for t, dt, its, itsv, e_d, e_v in stats:
        if e_d > (1 - band) * c["eta"] or e_v > (1 - band) * c["etaV"] / dt:
            break
        n += 1
    return n


@pytest.fixture
def tmpdir_path(tmp_path):
    return str(tmp_path)
