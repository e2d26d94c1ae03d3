"""Lanes per particle of the search's list build (sph_set_search_lanes).

Every count (1, 2, 4, 8) builds the same lists in the same row order, so the
count a context picks by timing its first searches cannot change a result:
  lists after two steps: identical rows (not only sets) for every count
  a tuned context against one pinned to 1 lane, 16 DFSPH steps: bit-identical
  positions, velocities and densities
"""
import os

import numpy as np
import pytest

from sph2o_amd import _lib, scenes as SC
from sph2o_amd.simulation import Context

pytestmark = pytest.mark.gpu


def _scene(tmp, method):
    sc = SC.small(method, n=12, viscosity_method=2, viscosity=0.01, surface_tension_method=3, surface_tension=0.1)
    path = os.path.join(tmp, "scene.json")
    SC.write_json(sc, path)
    p, x, xb = _lib.scene_from_json(path)
    return p, SC.jitter(x, 0.1 * 0.02, 5), xb


@pytest.mark.parametrize("method", [0, 2], ids=["wcsph", "dfsph"])
def test_search_lanes_build_the_same_lists(tmp_path, method):
    p, x, xb = _scene(str(tmp_path), method)
    ref = None
    for lanes in (1, 2, 4, 8):
        print("search lanes", lanes, flush=True)
        ctx = Context(0)
        try:
            ctx.set_search_lanes(lanes)
            sid = ctx.add_scene(p, x, None, xb)
            ctx.init()
            ctx.step(2)
            got = ctx.neighbors(sid)
            assert ctx.search_lanes()[:2] == (lanes, False)
        finally:
            ctx.close()
        if ref is None:
            ref = got
            assert got[0][-1] > 20 * len(x)   # a populated list
            continue
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), lanes


def test_tuned_search_steps_like_pinned(tmp_path):
    p, x, xb = _scene(str(tmp_path), 2)
    states = []
    for lanes in (0, 1):
        ctx = Context(0)
        try:
            ctx.set_search_lanes(lanes)
            sid = ctx.add_scene(p, x, None, xb)
            ctx.init()
            ctx.step(16)   # the init search and one per step: the 12 timed searches are done
            used, tuned, ms = ctx.search_lanes()
            if lanes == 0:
                assert tuned and used in (1, 2, 4, 8), (used, tuned)
                assert all(np.isfinite(t) and t > 0 for t in ms.values()), ms
                assert ms[used] == min(ms.values())
            else:
                assert (used, tuned) == (1, False)
            states.append(ctx.fluid_state(sid))
        finally:
            ctx.close()
    for k in ("x", "v", "rho"):
        assert np.array_equal(states[0][k], states[1][k]), k


def test_search_lanes_rejects_other_counts():
    ctx = Context(0)
    try:
        with pytest.raises(_lib.SPHError) as e:
            ctx.set_search_lanes(3)
        assert e.value.code == _lib.SPH_E_ARG
    finally:
        ctx.close()


def _growing_scene():
    """WCSPH: a 27-particle block and an emitter of 10 x 10 sheets (every 50
    steps at 4 m/s), so the active particles more than double while it runs."""
    return {
        "Configuration": {"startTime": 0.0, "endTime": 1e6, "timeStep": 1e-4, "fps": 60.0, "minTimeStep": 0.0,
                          "maxTimeStep": 0.005, "gravity": [0.0, 0.0, -9.81], "particleRadius": 0.01,
                          "simulationMethod": 0, "boundaryMethod": 2, "cflFactor": 1.0, "stiffness": 1000.0,
                          "gamma": 7.0},
        "Fluid": {"viscosityMethod": 2, "surfaceTensionMethod": -1, "adhesionMethod": -1,
                  "FluidModel": [{"density0": 1000.0,
                                  "fluidBlocks": [{"min": [0.02, 0.02, 0.02], "max": [0.08, 0.08, 0.08]}],
                                  "viscosity": 0.01, "boundaryViscosity": 0.0,
                                  "emitters": [{"position": [0.3, 0.15, 0.25],
                                                "rotation": [0.70710678, -0.70710678, 0.0, 0.0], "velocity": 4.0,
                                                "numParticles": 600, "startTime": 0.0002, "type": 0, "width": 0.2,
                                                "height": 0.2, "spacing": 1.0}]}]},
        "Boundary": [{"box": [{"min": [0, 0, 0], "max": [0.6, 0.3, 0.4]}], "sphere": [], "geometry": []}],
    }


def test_search_lanes_retuned_as_emitters_grow(tmp_path):
    """The tuned count is timed again once the active particles have more
    than doubled since it was picked; the trajectory stays the pinned one's."""
    path = os.path.join(str(tmp_path), "grow.json")
    SC.write_json(_growing_scene(), path)
    finals = []
    for lanes in (0, 1):
        ctx = Context(0)
        try:
            sid = ctx.load_scene_json(path)
            if lanes:
                ctx.set_search_lanes(lanes)
            ctx.init()
            tuned, active = [], []
            for _ in range(160):
                ctx.step(1)
                tuned.append(ctx.search_lanes()[1])
                active.append(ctx.num_active(sid))
            finals.append(ctx.fluid_state(sid))
        finally:
            ctx.close()
        if lanes:
            assert not any(tuned)
            continue
        first = tuned.index(True)                  # the first pick
        assert active[-1] > 2 * active[first], active
        again = tuned.index(False, first)          # timed again after the growth
        assert active[again - 1] > 2 * active[first], (again, active[first], active[again - 1])
        assert tuned[-1], "the second tuning finishes"
    for k in ("x", "v", "rho"):
        assert np.array_equal(finals[0][k], finals[1][k]), k
