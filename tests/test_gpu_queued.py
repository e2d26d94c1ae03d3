"""WCSPH steps queued in chunks (sph_step without a host check per step).

A chunk's steps run without host waits; an even chunk replays one captured
HIP graph of two steps. Each is checked against a path that takes the same
steps another way, bit for bit:
  graph replays over several chunks, with host reads and uneven step counts
    between them, against the recording path (sph_forward_record: queued
    steps without graphs) and against profiled stepping (no graphs either)
  a chunk whose search check fails (the block leaves the grid box) is undone
    and retaken with per-step host checks (test_gpu_adjoint.py covers it)
"""
import os

import numpy as np
import pytest

from sph2o_amd import scenes as SC
from sph2o_amd.simulation import Context

pytestmark = pytest.mark.gpu


def _ctx(tmp, name):
    sc = SC.small(0, n=12, viscosity_method=2, viscosity=0.01, surface_tension_method=3, surface_tension=0.1)
    path = os.path.join(tmp, name + ".json")
    SC.write_json(sc, path)
    ctx = Context(0)
    sid = ctx.load_scene_json(path)
    ctx.init()
    x0 = SC.jitter(ctx.fluid_state(sid)["x"], 0.1 * 0.02, 5)
    ctx.set_fluid_state(sid, x0, None)
    return ctx, sid


def test_graph_replays_match_unqueued_paths(tmp_path):
    tmp = str(tmp_path)
    runs = {}
    for mode in ("graph", "record", "profiled"):
        print("mode", mode, flush=True)
        ctx, sid = _ctx(tmp, mode)
        try:
            if mode == "graph":
                seen = []
                for n in (64, 2, 31, 1, 32):   # chunks of 32 and shorter, odd counts between
                    ctx.step(n)
                    seen.append(ctx.fluid_state(sid)["x"][:4].copy())
                assert all(np.isfinite(s).all() for s in seen)
            elif mode == "record":
                ctx.forward_record(130, 130)
            else:
                ctx.set_profiling(True)
                ctx.step(130)
            runs[mode] = ctx.fluid_state(sid)
        finally:
            ctx.close()
    for other in ("record", "profiled"):
        for k in ("x", "v", "rho"):
            assert np.array_equal(runs["graph"][k], runs[other][k]), (other, k)
