"""The five benchmark workloads at their own sizes, on the HIP path through the
C ABI, against the compiled reference's golden vectors and the oracle.

  W8k   8,000-particle WCSPH dam break (2,000 steps). 20 steps against the
        reference per particle: |dx| <= 1e-5 m, rms(dv)/rms(v) <= 1e-4,
        |drho|/rho0 <= 5e-4. The block sits flush in the box corner, so the
        first steps throw particles out at up to 550 m/s; two CPU builds of
        the same formulas (the FP32 and FP64 oracle) already differ from the
        reference by 1.4e-6 / 2.9e-6 m, rms(dv) 8.7e-6 / 1.5e-5 and
        4.2e-5 / 9.3e-5 of rho0 after 20 steps. After all 2,000 steps the
        box aggregates (particles inside, sum |v|^2, their centre of mass and
        mean density) within the spread of those builds: count, |v|^2 and
        density 1.5%, centre of mass 0.025 m (one particle spacing is 0.02).
  D1M   1,000,000 fluid + 400,002 boundary particles, DFSPH. Cell keys of
        every particle equal the FP32 oracle's; fluid and boundary
        within-support sets of 3,000 random particles exact after the first
        search and after 3 steps; iteration counts and dt of those steps
        equal the reference's; 4,000 sampled particles after one step within
        1e-6 m and rms(dv)/rms(v) of twice the FP32 oracle's own distance to
        the reference (2.8e-3; surface tension on a perfect lattice sits at
        the |k| gates); mean density error <= 0.1%.
        Initial densities within 2e-6 of rho0 of the FP64 oracle's (the
        reference's own FP32 sums are up to 6.3e-6 away; bar 1e-5 to them).
  B64   all 64 scenes of the sweep in one context for 10 steps: every scene
        bit-identical to the same scene stepped alone (state and per-step
        stats); the smallest scene against the reference (one step:
        iterations equal, rms(dv)/rms(v) within twice the FP32 oracle's own
        distance to the reference; 10 steps: iterations equal over the
        reference's decided steps, centre of mass within 5e-4 m).
  A100k 100,000-particle rollouts at A100k's geometry: dL/dv0 along a
        random direction, dL/d(viscosity) and dL/d(surface tension) against
        FP64 differences of the oracle evaluated on the GPU's branches, at
        the first eps where the loss is smooth (forward and backward
        differences agree within 1e-3, or one side's quotient is stable
        across two eps). WCSPH over 40 steps within 1e-3
        (measured 2.5e-5, 3.2e-4, 5.7e-5); past that the FP64 loss is not
        smooth at any eps down to 3e-8 (100 and 200 steps), so no difference
        quotient pins it. DFSPH with 8 + 8 iterations over 2 steps within
        5e-3 (measured 1.5e-3 to 2.8e-3 for dL/dv0): on blocks both wide and deep
        the FP32 trajectory's k, formed from rho* - rho0 ~ 1e-4 rho0, carries
        ~1e-3 relative rounding, and the gradient of that trajectory departs
        from the FP64 one by that much (800 to 20,000-particle blocks of
        depth 8 or width 10 agree within 2e-5). The 200-step gradients:
        WCSPH finite; the DFSPH adjoint reports non-finite values
        (SPH_E_NONFINITE): the discrete map's derivative grows ~10x per 5
        steps.
"""
import json
import os

import numpy as np
import pytest

from conftest import decided_steps, golden
from oracle import sphref as S
from sph2o_amd import _lib, scenes as SC
from sph2o_amd.simulation import Context
from test_gpu_adjoint import Case, _rel
from test_gpu_parity import _ctx, _rms_rel, _subset_rows

pytestmark = pytest.mark.gpu


def _box_aggregates(x, v, rho, box):
    inside = np.all((x >= 0) & (x <= np.asarray(box, np.float32)), axis=1)
    return np.array([inside.sum(), (v.astype(np.float64) ** 2).sum(), *x[inside].astype(np.float64).mean(0),
                     rho[inside].astype(np.float64).mean()])


# ---------------------------------------------------------------- W8k
def test_w8k_matches_reference(tmp_path):
    g = golden("ref_w8k.npz")
    sc = json.loads(str(g["scene"]))
    ctx, sid = _ctx(sc, str(tmp_path), x=g["x0"], xb=g["bx"])
    try:
        assert ctx.num_fluid(sid) == 8000 and ctx.num_boundary(sid) == 4002
        _, vol = ctx.boundary(sid)
        assert np.abs(vol - g["bvol"]).max() / np.abs(g["bvol"]).max() < 2e-6
        n = int(g["nsteps"])
        ctx.step(n)
        st = ctx.fluid_state(sid)
        o = S.Oracle(S.params_from_scene(sc), g["x0"], None, g["bx"], bits=64)
        o.init()
        o.step(n)
        so = o.state()
        o.close()
        for ref_x, ref_v, ref_rho in ((g["x"], g["v"], g["rho"]), (so["x"], so["v"], so["rho"])):
            assert np.abs(st["x"] - ref_x).max() <= 1e-5
            assert _rms_rel(st["v"], ref_v) <= 1e-4
            assert np.abs(st["rho"] - ref_rho).max() / 1000.0 <= 5e-4
        ctx.step(int(g["long_steps"]) - n)
        st = ctx.fluid_state(sid)
        assert np.isfinite(st["x"]).all() and np.isfinite(st["v"]).all()
        t, _ = ctx.time(sid)
        assert abs(t - 0.2) < 1e-4
        me = _box_aggregates(st["x"], st["v"], st["rho"], (0.8, 0.4, 0.4))
        ref = g["long_aggregates"]
        assert abs(me[0] - ref[0]) <= 0.015 * ref[0]
        assert abs(me[1] - ref[1]) <= 0.015 * ref[1]
        assert np.abs(me[2:5] - ref[2:5]).max() <= 0.025
        assert abs(me[5] - ref[5]) <= 0.015 * ref[5]
    finally:
        ctx.close()


# ---------------------------------------------------------------- D1M
def test_d1m_full_scale(tmp_path):
    g = golden("ref_d1m.npz")
    sc = json.loads(str(g["scene"]))
    p, x, xb = _lib.scene_from_json(_write(sc, str(tmp_path)))
    ctx = Context(0)
    try:
        sid = ctx.add_scene(p, x, None, xb)
        ctx.init()
        assert ctx.num_fluid(sid) == 1_000_000 and ctx.num_boundary(sid) == 400_002
        idx = g["idx"]
        st0 = ctx.fluid_state(sid)
        # initial densities: the reference's FP32 sums sit up to 6.3e-6 of rho0
        # from the FP64 oracle's on this scene (the FP32 oracle's too)
        o = S.Oracle(S.params_from_scene(sc), st0["x"], None, xb, bits=64)
        o.init()
        rho64 = o.state()["rho"]
        o.close()
        assert np.abs(st0["rho"] - rho64).max() / 1000.0 < 2e-6
        assert np.abs(st0["rho"][idx] - g["rho0"]).max() / 1000.0 < 1e-5
        # every cell key at 1M against the FP32 oracle
        o = S.Oracle(S.params_from_scene(sc), st0["x"], None, xb, bits=32)
        assert np.array_equal(ctx.cells(sid), o.cell_keys())
        o.close()
        pairs = _subset_rows(ctx, sid, sc, xb, 3000, 21)
        assert pairs > 3000 * 25
        stats = ctx.step(int(g["nsteps"]))
        ref = g["stats"]
        its = np.array([[s[sid].iterations, s[sid].iterations_v] for s in stats])
        k = decided_steps(ref, sc)
        assert k >= 1
        assert np.array_equal(its[:k], ref[:k, 2:4].astype(int))
        dts = np.array([s[sid].dt for s in stats])
        assert np.allclose(dts[:k], ref[:k, 1], rtol=1e-6, atol=0)
        st = ctx.fluid_state(sid)
        assert np.isfinite(st["x"]).all()
        e = float(np.mean(np.maximum(st["rho"].astype(np.float64) / 1000.0 - 1.0, 0.0)))
        assert e <= 1e-3
        assert abs(e - float(g["mean_density_error"])) <= 0.2 * float(g["mean_density_error"]) + 1e-5
        pairs = _subset_rows(ctx, sid, sc, xb, 3000, 22)
        assert pairs > 3000 * 25
    finally:
        ctx.close()
    # one step from the same start, sampled particles against the reference
    ctx = Context(0)
    try:
        sid = ctx.add_scene(p, x, None, xb)
        ctx.init()
        s = ctx.step(1)[0][sid]
        assert (s.iterations, s.iterations_v) == tuple(g["stats1"][0, 2:4].astype(int))
        st = ctx.fluid_state(sid)
        # surface tension on a perfect lattice puts |k_i + k_j| at the 1e-5
        # gates: the FP32 oracle lands rms 2.8e-3 from the reference here (the
        # FP64 one 8.9e-3), so the bar is twice the FP32 oracle's distance
        o = S.Oracle(S.params_from_scene(sc), x, None, xb, bits=32)
        o.init()
        o.step(1)
        so = o.state()
        o.close()
        spread = _rms_rel(so["v"][g["idx"]], g["v1"])
        assert _rms_rel(st["v"][g["idx"]], g["v1"]) <= max(1e-3, 2.0 * spread)
        assert np.abs(st["x"][g["idx"]] - g["x1"]).max() <= 1e-6
    finally:
        ctx.close()


def _write(sc, tmp, name="scene.json"):
    path = os.path.join(tmp, name)
    SC.write_json(sc, path)
    return path


# ---------------------------------------------------------------- B64
def test_b64_batched_equals_alone(tmp_path):
    tmp = str(tmp_path)
    scenes = SC.b64()
    assert len(scenes) == 64
    nsteps = 10
    paths = [_write(sc, tmp, "s%02d.json" % k) for k, sc in enumerate(scenes)]
    alone = []
    for path in paths:
        c = Context(0)
        try:
            sid = c.load_scene_json(path)
            c.init()
            stats = c.step(nsteps)
            st = c.fluid_state(sid)
            alone.append((st["x"], st["v"], [(s[sid].iterations, s[sid].iterations_v, s[sid].dt) for s in stats]))
        finally:
            c.close()
    total = sum(len(a[0]) for a in alone)
    assert 2_000_000 < total < 4_000_000
    batch = Context(0)
    try:
        for path in paths:
            batch.load_scene_json(path)
        batch.init()
        stats = batch.step(nsteps)
        for k in range(64):
            st = batch.fluid_state(k)
            assert np.array_equal(st["x"], alone[k][0]) and np.array_equal(st["v"], alone[k][1]), k
            assert [(s[k].iterations, s[k].iterations_v, s[k].dt) for s in stats] == alone[k][2], k
            e = float(np.mean(np.maximum(st["rho"].astype(np.float64) / 1000.0 - 1.0, 0.0)))
            assert e <= 1e-3, k
    finally:
        batch.close()


def test_b64_smallest_scene_matches_reference(tmp_path):
    g = golden("ref_b64_smallest.npz")
    sc = json.loads(str(g["scene"]))
    assert sc == json.loads(json.dumps(SC.b64()[int(g["index"])]))
    ctx, sid = _ctx(sc, str(tmp_path), x=g["x0"], xb=g["bx"])
    try:
        s = ctx.step(1)[0][sid]
        assert (s.iterations, s.iterations_v) == tuple(g["stats1"][0, 2:4].astype(int))
        st = ctx.fluid_state(sid)
        # as on the reference's example scene, surface tension puts many
        # |k_i + k_j| at the 1e-5 gates: the FP32 oracle already differs from
        # the reference by rms 2.5e-3 (the FP64 one by 2.9e-2)
        o = S.Oracle(S.params_from_scene(sc), g["x0"], None, g["bx"], bits=32)
        o.init()
        o.step(1)
        spread = _rms_rel(o.state()["v"], g["v1"])
        o.close()
        assert _rms_rel(st["v"], g["v1"]) <= max(1e-3, 2.0 * spread)
        n = int(g["nsteps"])
        stats = [[s]] + ctx.step(n - 1)
        its = np.array([[q[sid].iterations, q[sid].iterations_v] for q in stats])
        ref = g["stats"]
        k = decided_steps(ref, sc)
        assert k >= 1
        assert np.array_equal(its[:k], ref[:k, 2:4].astype(int))
        st = ctx.fluid_state(sid)
        assert np.abs(st["x"].mean(0) - g["x"].mean(0)).max() < 5e-4
    finally:
        ctx.close()


# ---------------------------------------------------------------- A100k
# method, dt, steps, tolerance
A100K_CASES = {"wcsph": (0, 2e-4, 40, 1e-3), "dfsph": (2, 5e-4, 2, 5e-3)}


@pytest.mark.parametrize("name", list(A100K_CASES))
def test_a100k_gradients_match_fp64_finite_differences(tmp_path, name):
    method, dt, nsteps, tol = A100K_CASES[name]
    sc, _, _ = SC.a100k(method, dt=dt)
    case = Case(str(tmp_path), sc, xseed=1, vseed=3)
    try:
        assert len(case.x0) == 100_000
        case.gpu_branch(nsteps)
        xT, res = case.gpu_grad(nsteps)
        lg = float(np.mean(np.sum((xT - case.y) ** 2, axis=1)))
        lo, _ = case.oracle(case.x0.astype(np.float64), case.v0.astype(np.float64), case.nu, case.gam, nsteps)
        assert abs(lg - lo) <= 1e-4 * lo
        checks = case.fd_checks(res, nsteps, ndir=1, kinds=("v0", "viscosity", "surface_tension"),
                                smooth=(1e-6, 1e-7, 3e-8, 1e-8))
        assert len(checks) == 3
        for kind, g_, fd in checks:
            assert _rel(g_, fd) <= tol, (kind, g_, fd)
    finally:
        case.ctx.close()


@pytest.mark.parametrize("name", ["wcsph", "dfsph"])
def test_a100k_200_step_gradients(tmp_path, name):
    """The config's 200 steps with checkpoints every 20: WCSPH gradients are
    finite; the DFSPH adjoint reports non-finite values with SPH_E_NONFINITE."""
    method, dt = (0, 2e-4) if name == "wcsph" else (2, 5e-4)
    sc, _, _ = SC.a100k(method, dt=dt)
    case = Case(str(tmp_path), sc, xseed=1, vseed=3)
    try:
        if name == "wcsph":
            _, res = case.gpu_grad(200, 20)
            assert np.isfinite(res["v0"]).all() and np.isfinite(res["x0"]).all()
            assert np.isfinite(res["viscosity"]) and np.isfinite(res["surface_tension"])
            assert np.abs(res["v0"]).max() > 0
        else:
            with pytest.raises(_lib.SPHError) as e:
                case.gpu_grad(200, 20)
            assert e.value.code == _lib.SPH_E_NONFINITE
    finally:
        case.ctx.close()
