"""The viewer feed (sph_viewer_feed): the vec4 positions the reference's viewer
instances (SimulationObject::getPositions, SimulationObject.h:16, w = 1) and
its scalar field |v| (SPHSimulation::updateScalarField, SPHSimulation.cpp:
638-657), for the active particles in original order, written on the GPU.

Checks: positions bit-identical to sph_get_fluid_state and w = 1; |v| equal
to glm::length in the reference's FMA rounding with a correctly rounded
square root, sqrt(fma(x, x, fma(z, z, y*y))), emulated here in FP64 (the
emulation rounds the inner sums twice, so at most 1 ulp apart, and equal for
at least 99.9% of the particles); emitter
scenes give sph_num_active rows; pageable host buffers are refused with
SPH_E_ARG and an uninitialised context with SPH_E_STATE.
"""
import ctypes
import os

import numpy as np
import pytest

from sph2o_amd import _lib, scenes as SC
from sph2o_amd.simulation import Context

pytestmark = pytest.mark.gpu


def _len_ref(v):
    v = v.astype(np.float64)
    yy = (v[:, 1] * v[:, 1]).astype(np.float32).astype(np.float64)
    t = (v[:, 2] * v[:, 2] + yy).astype(np.float32).astype(np.float64)
    d = (v[:, 0] * v[:, 0] + t).astype(np.float32).astype(np.float64)
    return np.sqrt(d).astype(np.float32)


def _ctx(tmp, sc, init=True):
    path = os.path.join(tmp, "scene.json")
    SC.write_json(sc, path)
    ctx = Context(0)
    sid = ctx.load_scene_json(path)
    if init:
        ctx.init()
    return ctx, sid


def test_gpu_viewer_feed_matches_state(tmp_path):
    ctx, sid = _ctx(str(tmp_path), SC.small(2, viscosity_method=2, viscosity=0.01))
    ctx.step(5)
    p, s = ctx.viewer_feed(sid)
    st = ctx.fluid_state(sid)
    assert p.shape == (ctx.num_fluid(sid), 4)
    assert np.array_equal(p[:, :3], st["x"]) and (p[:, 3] == 1.0).all()
    ref = _len_ref(st["v"])
    ulp = np.abs(s.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1 and (ulp == 0).mean() >= 0.999
    assert (s > 0).any()


def test_gpu_viewer_feed_active_particles_only(tmp_path):
    ctx, sid = _ctx(str(tmp_path), SC.emitter_scenes()["dfsph"])
    ctx.step(20)
    na = ctx.num_active(sid)
    assert na < ctx.num_fluid(sid)
    pos4, speed = ctx.viewer_feed(sid)
    assert pos4.shape == (na, 4) and speed.shape == (na,)
    st = ctx.fluid_state(sid)
    assert np.array_equal(pos4[:, :3], st["x"][:na])


def test_gpu_viewer_feed_refuses_host_buffers_and_uninitialised(tmp_path):
    ctx, sid = _ctx(str(tmp_path), SC.small(2))
    host = np.zeros((ctx.num_fluid(sid), 4), np.float32)
    rc = ctx.L.sph_viewer_feed(ctx.h, sid, ctypes.c_void_p(host.ctypes.data), None)
    assert rc == _lib.SPH_E_ARG
    ctx2, sid2 = _ctx(str(tmp_path), SC.small(2), init=False)
    assert ctx2.L.sph_viewer_feed(ctx2.h, sid2, None, None) == _lib.SPH_E_STATE
