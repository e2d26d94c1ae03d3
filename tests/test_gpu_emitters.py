"""Emitters (FluidObject::addEmitter, Emitter.cpp:5-123) on the HIP path,
against the compiled reference's runs of the same scene JSON
(tests/golden/ref_emit_*.npz, written by tests/golden/make_golden.py).

Tolerances, stated per check:
  active particle count after every step: exact
  state right after the first emission (1 DFSPH step, 3 WCSPH / PCISPH steps):
      |dx| <= 1e-6 m for every particle (the emitted ones are placed on the
      host in the reference's FP32 arithmetic), rms(dv)/rms(v) <= 1e-3
  longer runs (WCSPH 120 steps, PCISPH 100, DFSPH 26 = just past its third
      emission): per-step time and time step equal to the reference's (|dt|
      relative 1e-6); WCSPH and PCISPH iteration counts equal at every step,
      |dx| <= 1e-5 m; DFSPH iteration counts equal over the leading steps
      whose loops the reference stopped clear of the thresholds (tests/
      conftest.py decided_steps: 18 here; at step 19 its divergence error
      sits 3.5% under the threshold, inside the spread of FP32 evaluations);
      kinetic energy within 5%, centre of mass within 1e-4 m, mean density
      error within 20%.
"""
import json
import os

import numpy as np
import pytest

from conftest import decided_steps, golden
from sph2o_amd import scenes as SC
from sph2o_amd.simulation import Context

pytestmark = pytest.mark.gpu

TAGS = ["dfsph", "wcsph", "pcisph"]


def _load(tmp, sc):
    path = os.path.join(tmp, "scene.json")
    SC.write_json(sc, path)
    ctx = Context(0)
    sid = ctx.load_scene_json(path)
    ctx.init()
    return ctx, sid


def _rms_rel(a, b):
    d = a - b
    return float(np.sqrt((d ** 2).mean() / max(1e-30, (b ** 2).mean())))


@pytest.mark.parametrize("tag", TAGS)
def test_gpu_first_emission_matches_reference(tmp_path, tag):
    g = golden("ref_emit_%s.npz" % tag)
    sc = json.loads(str(g["scene"]))
    ctx, sid = _load(str(tmp_path), sc)
    n_total = len(g["x0"])
    assert ctx.num_fluid(sid) == n_total
    na0 = ctx.num_active(sid)
    assert na0 < n_total
    st0 = ctx.fluid_state(sid)
    assert np.array_equal(st0["x"][:na0], g["x0"][:na0])
    assert not st0["x"][na0:].any()   # inactive particles wait at the origin, as in the reference
    first = int(g["first"])
    act = []
    for _ in range(first):
        ctx.step(1)
        act.append(ctx.num_active(sid))
    assert act == [int(a) for a in g["nact1"]]
    st = ctx.fluid_state(sid)
    na = act[-1]
    assert na > na0
    assert np.abs(st["x"] - g["x1"]).max() <= 1e-6
    assert np.abs(st["x"][na0:na] - g["x1"][na0:na]).max() <= 1e-6
    assert _rms_rel(st["v"][:na], g["v1"][:na]) <= 1e-3
    assert not st["x"][na:].any()


@pytest.mark.parametrize("tag", TAGS)
def test_gpu_emitter_run_matches_reference(tmp_path, tag):
    g = golden("ref_emit_%s.npz" % tag)
    sc = json.loads(str(g["scene"]))
    ctx, sid = _load(str(tmp_path), sc)
    n = int(g["nsteps"])
    act, stats = [], []
    for _ in range(n):
        stats.append(ctx.step(1)[0][sid])
        act.append(ctx.num_active(sid))
    assert act == [int(a) for a in g["nact"]]
    ref = g["stats"]
    assert np.allclose([s.t for s in stats], ref[:, 0], rtol=1e-6, atol=0)
    assert np.allclose([s.dt for s in stats], ref[:, 1], rtol=1e-6, atol=0)
    if tag == "dfsph":
        k = decided_steps(ref, sc)
        assert k >= 10
        its = np.array([[s.iterations, s.iterations_v] for s in stats])
        assert np.array_equal(its[:k], ref[:k, 2:4].astype(int))
        assert abs(its[:, 0].sum() - ref[:, 2].sum()) <= 2
        assert abs(its[:, 1].sum() - ref[:, 3].sum()) <= 2
    else:
        assert [s.iterations for s in stats] == [int(q) for q in ref[:, 2]]
    st = ctx.fluid_state(sid)
    na = act[-1]
    x, v = st["x"][:na], st["v"][:na]
    gx, gv = g["x"][:na], g["v"][:na]
    if tag != "dfsph":
        assert np.abs(x - gx).max() <= 1e-5
    assert abs((v ** 2).sum() - (gv ** 2).sum()) / (gv ** 2).sum() < 0.05
    assert np.abs(x.mean(0) - gx.mean(0)).max() < 1e-4
    e_ref = np.mean(np.maximum(g["rho"][:na] / 1000 - 1, 0))
    e_me = np.mean(np.maximum(st["rho"][:na] / 1000 - 1, 0))
    assert abs(e_me - e_ref) < 0.2 * e_ref + 1e-5


def test_gpu_add_emitter_abi_matches_json(tmp_path):
    # the same DFSPH emitter added through sph_add_emitter instead of the JSON
    sc = SC.emitter_scenes()["dfsph"]
    e = sc["Fluid"]["FluidModel"][0].pop("emitters")[0]
    path = os.path.join(str(tmp_path), "plain.json")
    SC.write_json(sc, path)
    ctx = Context(0)
    sid = ctx.load_scene_json(path)
    ctx.add_emitter(sid, e["type"], e["numParticles"], e["position"], e["rotation"], e["velocity"],
                    e["startTime"], e["width"], e["height"], e["spacing"])
    ctx.init()
    ctx.step(30)
    sc2 = SC.emitter_scenes()["dfsph"]
    ctx2, sid2 = _load(str(tmp_path), sc2)
    ctx2.step(30)
    a, b = ctx.fluid_state(sid), ctx2.fluid_state(sid2)
    assert ctx.num_active(sid) == ctx2.num_active(sid2)
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["v"], b["v"])
