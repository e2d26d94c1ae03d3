"""The decided-steps rule of the DFSPH trajectory checks (tests/conftest.py
decided_steps): leading steps whose loops the reference stopped with its
error at least 10% under the threshold (DFSPHSolver.cpp:446, :746)."""
import json

import numpy as np

from conftest import decided_steps, golden

SCENE = {"Configuration": {"eta": 0.001, "etaV": 0.001}}


def _row(dt, e_d, e_v):
    return [0.0, dt, 2, 1, e_d, e_v]


def test_stops_at_first_step_within_the_band():
    dt = 1e-3   # divergence threshold etaV / dt = 1.0
    stats = np.array([_row(dt, 1e-4, 0.5), _row(dt, 8.9e-4, 0.89), _row(dt, 1e-4, 0.91), _row(dt, 1e-4, 0.1)])
    assert decided_steps(stats, SCENE) == 2
    stats[1, 4] = 9.1e-4   # the density error within 10% of eta
    assert decided_steps(stats, SCENE) == 1
    assert decided_steps(stats, SCENE, band=0.05) == 4


def test_divergence_threshold_scales_with_dt():
    # the same divergence error is decided at dt = 1e-3 (threshold 1.0) but
    # not at dt = 2e-3 (threshold 0.5)
    assert decided_steps(np.array([_row(1e-3, 0.0, 0.6)]), SCENE) == 1
    assert decided_steps(np.array([_row(2e-3, 0.0, 0.6)]), SCENE) == 0


def test_golden_runs():
    for name, k in (("ref_traj_dfsph40.npz", 24), ("ref_emit_dfsph.npz", 18), ("ref_traj_dfsph20_art.npz", 16)):
        g = golden(name)
        assert decided_steps(g["stats"], json.loads(str(g["scene"]))) == k
