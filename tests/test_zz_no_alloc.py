"""sph_dump_cells_noalloc (the allocation-free variant of sph_dump_cells)
through the C ABI: its keys equal sph_dump_cells' bit for bit on two calls
in a row, the second on a smaller scene, and again after the particles have
moved; a scene above its maximum is refused with SPH_E_CAPACITY. Inactive
emitter particles read zero, as from sph_dump_cells."""
import os

import numpy as np
import pytest

from sph2o_amd import _lib, scenes as SC
from sph2o_amd.simulation import Context

pytestmark = pytest.mark.gpu


def _add(ctx, sc, tmp, name, jit, seed):
    path = os.path.join(tmp, name)
    SC.write_json(sc, path)
    p, x, xb = _lib.scene_from_json(path)
    return ctx.add_scene(p, SC.jitter(x, jit, seed), None, xb)


def test_noalloc_cells_match_dump_cells(tmp_path):
    tmp = str(tmp_path)
    ctx = Context(0)
    try:
        big = _add(ctx, SC.small(2, n=14, viscosity_method=2, viscosity=0.01), tmp, "a.json", 0.004, 1)
        small = _add(ctx, SC.small(2, n=6, box=(0.3, 0.3, 0.3)), tmp, "b.json", 0.004, 2)
        ctx.init()
        assert ctx.num_fluid(big) > ctx.num_fluid(small)
        for _ in range(2):
            for sid in (big, small):   # a big call, then a smaller one on the same scratch
                ref = ctx.cells(sid)
                got = ctx.cells_noalloc(sid)
                assert got.shape == ref.shape
                assert np.array_equal(got, ref), sid
            ctx.step(3)
    finally:
        ctx.close()


def test_noalloc_cells_inactive_particles_and_limit(tmp_path):
    tmp = str(tmp_path)
    ctx = Context(0)
    try:
        path = os.path.join(tmp, "e.json")
        SC.write_json(SC.emitter_scenes()["wcsph"], path)
        sid = ctx.load_scene_json(path)
        ctx.init()
        ctx.step(2)
        assert ctx.num_active(sid) < ctx.num_fluid(sid)
        assert np.array_equal(ctx.cells_noalloc(sid), ctx.cells(sid))
    finally:
        ctx.close()
    # a 102^3 block: 1,061,208 active particles, above the 2^20 limit
    sc = SC.small(0, n=102, radius=0.005, box=(1.2, 1.2, 1.2))
    ctx = Context(0)
    try:
        path = os.path.join(tmp, "big.json")
        SC.write_json(sc, path)
        p, x, xb = _lib.scene_from_json(path)
        assert len(x) > 1 << 20
        sid = ctx.add_scene(p, x, None, xb)
        ctx.init()
        with pytest.raises(_lib.SPHError) as e:
            ctx.cells_noalloc(sid)
        assert e.value.code == _lib.SPH_E_CAPACITY
    finally:
        ctx.close()
