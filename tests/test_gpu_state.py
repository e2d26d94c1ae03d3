"""Exact resume through sph_save_state / sph_load_state: a run saved after N1
steps and resumed in a fresh context continues bit-identically (positions,
velocities, densities, time, iteration counts, active particles) for the
three solvers, a batch of scenes, and an emitter scene mid-emission."""
import os

import numpy as np
import pytest

from sph2o_amd import _lib, scenes as SC
from sph2o_amd.simulation import Context

pytestmark = pytest.mark.gpu


def _make(tmp, scenes):
    ctx = Context(0)
    ids = []
    for k, sc in enumerate(scenes):
        path = os.path.join(tmp, "s%d.json" % k)
        SC.write_json(sc, path)
        ids.append(ctx.load_scene_json(path))
    ctx.init()
    return ctx, ids


def _snap(ctx, ids):
    out = []
    for sid in ids:
        st = ctx.fluid_state(sid)
        out.append((st["x"], st["v"], st["rho"], ctx.num_active(sid)))
    return out


CASES = {
    "dfsph": [SC.small(2, viscosity_method=2, viscosity=0.01, surface_tension_method=3, surface_tension=0.1)],
    "wcsph": [SC.small(0, viscosity_method=2, viscosity=0.01)],
    "pcisph": [SC.small(1, viscosity_method=2, viscosity=0.01)],
    "dfsph_batch": [SC.small(2, n=6, viscosity_method=2, viscosity=0.01),
                    SC.small(2, n=8, surface_tension_method=3, surface_tension=0.1, box=(0.5, 0.4, 0.5))],
    "dfsph_emitter": [SC.emitter_scenes()["dfsph"]],
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_resume_is_bit_identical(tmp_path, case):
    tmp = str(tmp_path)
    scenes = CASES[case]
    n1, n2 = (10, 12) if case != "dfsph_emitter" else (10, 10)   # the emitter fires at steps 1 and 16
    a, ids = _make(tmp, scenes)
    a.step(n1)
    path = os.path.join(tmp, "state.bin")
    a.save_state(path)
    sa = a.step(n2)
    snap_a = _snap(a, ids)
    b, ids_b = _make(tmp, scenes)
    b.load_state(path)
    sb = b.step(n2)
    snap_b = _snap(b, ids_b)
    for (xa, va, ra, na), (xb, vb, rb, nb) in zip(snap_a, snap_b):
        assert na == nb
        assert np.array_equal(xa, xb) and np.array_equal(va, vb) and np.array_equal(ra, rb)
    for s_a, s_b in zip(sa, sb):
        for sid in ids:
            assert (s_a[sid].t, s_a[sid].dt, s_a[sid].iterations, s_a[sid].iterations_v) == \
                   (s_b[sid].t, s_b[sid].dt, s_b[sid].iterations, s_b[sid].iterations_v)


def test_gpu_state_file_must_match_the_scenes(tmp_path):
    tmp = str(tmp_path)
    a, _ = _make(tmp, CASES["dfsph"])
    path = os.path.join(tmp, "state.bin")
    a.save_state(path)
    b, _ = _make(tmp, [SC.small(2, n=6)])
    with pytest.raises(_lib.SPHError) as e:
        b.load_state(path)
    assert e.value.code == _lib.SPH_E_ARG
    with open(path, "r+b") as f:
        f.write(b"NOTSTATE")
    with pytest.raises(_lib.SPHError):
        a.load_state(path)


def test_gpu_rejected_state_file_leaves_the_context_unchanged(tmp_path):
    """A truncated file, a duplicated particle id and an impossible grid are
    all refused before anything is applied: the context then steps exactly
    as an untouched twin does."""
    tmp = str(tmp_path)
    scenes = CASES["dfsph"]
    a, _ = _make(tmp, scenes)
    a.step(5)
    good = os.path.join(tmp, "state.bin")
    a.save_state(good)
    with open(good, "rb") as f:
        data = bytearray(f.read())
    NP = 1024   # 1,000 particles padded to whole 256-slot blocks
    header, scene_rec = 24, 16 + 28 + 12 + 8 + 8 + 80   # counts, grid, cells, t/dt, its, SceneState
    ids_at = header + scene_rec + 2 * NP * 16
    dup = bytearray(data)
    dup[ids_at + 4:ids_at + 8] = dup[ids_at:ids_at + 4]
    grid = bytearray(data)
    grid[header + 16 + 12:header + 16 + 16] = (0).to_bytes(4, "little")   # gn.x = 0
    variants = {"truncated": data[:len(data) // 2], "duplicate_id": dup, "grid": grid}
    b, ids_b = _make(tmp, scenes)
    c, ids_c = _make(tmp, scenes)
    b.step(3)
    c.step(3)
    for name, blob in variants.items():
        path = os.path.join(tmp, name + ".bin")
        with open(path, "wb") as f:
            f.write(bytes(blob))
        with pytest.raises(_lib.SPHError) as e:
            b.load_state(path)
        assert e.value.code == _lib.SPH_E_ARG, name
    sb, sc_ = b.step(6), c.step(6)
    for (xb, vb, rb, nb), (xc, vc, rc, nc) in zip(_snap(b, ids_b), _snap(c, ids_c)):
        assert nb == nc and np.array_equal(xb, xc) and np.array_equal(vb, vc) and np.array_equal(rb, rc)
    assert [(s[0].t, s[0].iterations, s[0].iterations_v) for s in sb] == \
           [(s[0].t, s[0].iterations, s[0].iterations_v) for s in sc_]
    b.load_state(good)   # the good file still loads afterwards
