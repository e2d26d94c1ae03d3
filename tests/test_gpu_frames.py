"""Frame output (SceneLoader::writeFluid through SPHSimulation::step,
SPHSimulation.cpp:157-191, TimeHelper.h:100-110) of the headless CLI
(sph2o_amd/sph2o_cli, the reference's `-g false` path) and of the Python
SPHSimulation mirror, against the compiled reference's own frame files
(tests/golden/ref_frames_wcsph.npz: the WCSPH emitter scene at 1000 frames
per second for 120 steps).

Checked, per frame: the file name (<scene>_NNNN.json), the steps that wrote
a frame (the cadence follows the reference's FP32 lastTimeSave), the frame
count, numParticles (all particles, inactive emitter ones included) and
numActiveParticles exact; time and timeStep equal in FP32; positions within
1e-5 m and rms(dv)/rms(v) <= 1e-3 (the bars of the 120-step emitter run).
"""
import glob
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden
from sph2o_amd import scenes as SC
from sph2o_amd.simulation import SPHSimulation

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "sph2o_amd", "sph2o_cli")


def _rms_rel(a, b):
    d = a - b
    return float(np.sqrt((d ** 2).mean() / max(1e-30, (b ** 2).mean())))


def _check_frames(files, g):
    names = [os.path.basename(f) for f in files]
    assert names == [str(n) for n in g["names"]]
    for k, path in enumerate(files):
        with open(path) as f:
            fr = json.load(f)
        assert fr["numParticles"] == int(g["num_particles"][k])
        assert fr["numActiveParticles"] == int(g["num_active"][k])
        assert np.float32(fr["time"]) == np.float32(g["time"][k])
        assert np.float32(fr["timeStep"]) == np.float32(g["time_step"][k])
        x = np.array(fr["positions"], np.float32)
        v = np.array(fr["velocities"], np.float32)
        assert x.shape == g["positions"][k].shape
        assert np.abs(x - g["positions"][k]).max() <= 1e-5, k
        assert _rms_rel(v, g["velocities"][k]) <= 1e-3, k


def test_cli_frames_match_reference(tmp_path):
    g = golden("ref_frames_wcsph.npz")
    sc = json.loads(str(g["scene"]))
    path = os.path.join(str(tmp_path), "scene.json")
    SC.write_json(sc, path)
    r = subprocess.run([CLI, "-c", path, "-n", str(int(g["nsteps"])), "-o", "true"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("steps %d " % int(g["nsteps"]))
    _check_frames(sorted(glob.glob(os.path.join(str(tmp_path), "scene_*.json"))), g)


def test_simulation_step_frames_match_reference(tmp_path):
    g = golden("ref_frames_wcsph.npz")
    sc = json.loads(str(g["scene"]))
    path = os.path.join(str(tmp_path), "scene.json")
    SC.write_json(sc, path)
    sim = SPHSimulation(0)
    try:
        assert sim.importScene(path)
        sim.activateSave(True)
        sim.init()
        saved = []
        for _ in range(int(g["nsteps"])):
            s = sim.step()
            saved.append((1.0 if s else 0.0, sim.getFrame()))
        assert np.array_equal(np.array(saved), g["saved"])
        _check_frames(sorted(glob.glob(os.path.join(str(tmp_path), "scene_*.json"))), g)
    finally:
        sim.ctx.close()
