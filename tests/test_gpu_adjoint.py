"""Gradients of WCSPH rollouts (sph_forward_record / sph_adjoint_*) on the GPU,
checked against FP64 central finite differences of the oracle.

WCSPH and DFSPH (8 + 8 Jacobi iterations, fixed dt). The loss is
L = mean_i |x_i(T) - y_i|^2 with y = x(0) + N(0, (0.5 s)^2); the
checked gradients are dL/dv0 and dL/dx0 (random unit directions), and
dL/d(viscosity), dL/d(surface tension). Central differences in FP64 on the
oracle with eps = 1e-6 (v0, coefficients) and 1e-8 (x0).

The GPU differentiates its FP32 trajectory, which stays within ~1e-8 m of
the FP64 one but resolves the reference's two non-smooth points on its own:
  * the pressure max(0, p) (WCSPHSolver.cpp:93) has a kink at rho = rho0;
    FP32 density sums carry ~3e-6 relative error, so a particle that close
    to rho0 can sit on the other side than in FP64, where the one-sided
    derivatives differ (one such particle-step moved a 20-step gradient by
    3e-3);
  * the cohesion kernel jumps 2x at |r| = h/2 (Cohesion.h:37), which is the
    lattice spacing, so pairs sit there all the time; a pair whose FP32 and
    FP64 lengths fall on different sides changes the surface-tension
    derivative by a fixed amount from that step on (1e-2 of dL/dgamma).
  * DFSPH adds the rho* >= rho0 and d >= 0 clamps, the warm-start conditions
    and the |k_i + k_j| > 1e-5, |k_i| > 1e-5 gates (DFSPHSolver.cpp:285, :608,
    :479, :483), all decided by FP32 values near their thresholds.
The oracle therefore evaluates its FP64 rollouts on the GPU's branches:
every GPU step's FP32 positions and p > 0 pattern (WCSPH) or the k and
rho* / divergence of every Jacobi pass (DFSPH, read from the recorded tape
through sph_recorded_field) are handed to it (Oracle.set_pressure_branch /
set_dfsph_branch / set_cohesion_positions; the branch of a pair is decided by
the reference's FP32 length of the GPU positions). Every
difference is still taken between runs whose inner-branch cohesion
fingerprints agree, which the checker asserts.

Tolerance: relative error <= 1e-3 per direction and coefficient.
"""
import os

import numpy as np
import pytest
import torch

from oracle import sphref as S
from sph2o_amd import _lib, scenes as SC
from sph2o_amd.autograd import rollout
from sph2o_amd.simulation import Context

pytestmark = pytest.mark.gpu

R = 0.005
SP = 2 * R


NIT = 8   # DFSPH: iterations of both solves (min = max)


def _scene(visc=True, st=True, n=(10, 10, 8), dt=5e-4, method=0):
    hi = tuple(SP + SP * c for c in n)
    return SC.scene_dict(method, R, [((SP, SP, SP), hi)], [((0, 0, 0), (0.2, 0.2, 0.15))], dt=dt, max_dt=dt,
                         min_dt=dt, min_it=NIT, max_it=NIT, min_it_v=NIT, max_it_v=NIT,
                         viscosity_method=2 if visc else -1, viscosity=0.01 if visc else 0.0,
                         surface_tension_method=3 if st else -1, surface_tension=0.1 if st else 0.0)


class Case:
    """One scene on the GPU plus the FP64 oracle of the same scene."""

    def __init__(self, tmp, sc, xseed=1, vseed=3):
        path = os.path.join(tmp, "scene.json")
        SC.write_json(sc, path)
        self.sc = sc
        self.ctx = Context(0)
        self.sid = self.ctx.load_scene_json(path)
        self.ctx.init()
        self.xb, _ = self.ctx.boundary(self.sid, volumes=False)
        self.x0 = SC.jitter(self.ctx.fluid_state(self.sid)["x"], 0.01 * SP, xseed)
        self.v0 = (0.05 * SC.normal(vseed, self.x0.size).reshape(self.x0.shape)).astype(np.float32)
        self.y = SC.adjoint_target(self.x0)
        self.P = S.params_from_scene(sc)
        self.nu = float(self.P.viscosity)
        self.gam = float(self.P.surface_tension)
        self.dfsph = sc["Configuration"]["simulationMethod"] == 2
        self.branch = None

    def gpu_branch(self, nsteps):
        """The branches the GPU rollout took, in original order: per step the
        FP32 positions at its start (cohesion) and, WCSPH, the p > 0 pattern;
        DFSPH, the k and rho* / divergence of every Jacobi pass (from the tape
        of a rollout recorded in one segment)."""
        c, sid = self.ctx, self.sid
        c.set_fluid_state(sid, self.x0, self.v0)
        if self.dfsph:
            c.forward_record(nsteps, nsteps)
            rec = lambda name, k: np.array([c.recorded(name, k, sid, q) for q in range(NIT + 1)])
            self.branch = [np.array([rec(name, k) for k in range(nsteps)])
                           for name in ("k_density", "rho_star", "k_divergence", "divergence")]
            self.branch_x = np.array([c.recorded("x", k, sid) for k in range(nsteps)], np.float32)
            return self.branch
        masks, xs = [], []
        for _ in range(nsteps):
            xs.append(c.fluid_state(sid)["x"])
            c.step(1)
            masks.append(c.fluid_state(sid)["p"] > 0)
        self.branch = np.array(masks)
        self.branch_x = np.array(xs, np.float32)
        return self.branch

    def gpu_grad(self, nsteps, every=20):
        c, sid = self.ctx, self.sid
        c.set_fluid_state(sid, self.x0, self.v0)
        c.forward_record(nsteps, every)
        xT = c.fluid_state(sid)["x"].astype(np.float64)
        seed = (2.0 * (xT - self.y) / len(xT)).astype(np.float32)
        return xT, c.adjoint({sid: (seed, None)})[sid]

    def oracle(self, x, v, nu, gam, nsteps):
        q = S.Params.from_buffer_copy(self.P)
        q.viscosity = nu
        q.surface_tension = gam
        o = S.Oracle(q, x.astype(np.float32), v.astype(np.float32), self.xb, bits=64)
        try:
            o.init()
            o.set_state64(x, v)
            if self.branch is not None:
                if self.dfsph:
                    o.set_dfsph_branch(*(b[:nsteps] for b in self.branch))
                else:
                    o.set_pressure_branch(self.branch[:nsteps])
                o.set_cohesion_positions(self.branch_x[:nsteps])
            o.step(nsteps)
            xt = o.state()["x"]
            return float(np.mean(np.sum((xt - self.y) ** 2, axis=1))), o.cohesion_print()
        finally:
            o.close()

    def fd_checks(self, res, nsteps, seed=8, ndir=2, kinds=None, smooth=None):
        """[(kind, gpu, fd)] over random directions and the coefficients
        (`kinds`: which of v0, x0 (ndir directions each), viscosity,
        surface_tension; default all).

        smooth: None takes central differences at one eps per kind and asserts
        that no run crossed a cohesion jump (fingerprints). A decreasing
        sequence of eps instead tests the loss for jumps, whatever their cause
        (on large scenes the FP64 rollouts also meet non-smooth points that
        the branches handed over do not fix): at each eps the forward and
        backward differences f+ = (L(+eps) - L(0)) / eps and
        f- = (L(0) - L(-eps)) / eps are formed; a jump J inside (0, eps] on
        one side moves that side's quotient by J / eps. The first eps at
        which f+ and f- agree within 1e-3 of their mean gives the central
        difference; failing that, a side whose quotient agrees within 1e-4 at
        two successive eps is smooth over the larger interval and gives its
        one-sided difference at the smaller eps (error eps L'' / 2)."""
        X0, V0 = self.x0.astype(np.float64), self.v0.astype(np.float64)
        L0, base = self.oracle(X0, V0, self.nu, self.gam, nsteps)
        if smooth is not None:
            return self._fd_smooth(res, nsteps, seed, ndir, kinds, smooth, X0, V0, L0)
        rng = np.random.default_rng(seed)
        out = []
        want = set(kinds or ("v0", "x0", "viscosity", "surface_tension"))
        kinds = [k for k in ["v0"] * ndir + ["x0"] * ndir if k in want]
        if self.nu and "viscosity" in want:
            kinds.append("viscosity")
        if self.gam and "surface_tension" in want:
            kinds.append("surface_tension")
        for kind in kinds:
            eps = 1e-8 if kind == "x0" else 1e-6
            if kind in ("v0", "x0"):
                d = rng.normal(size=self.x0.shape)
                d /= np.linalg.norm(d)
                g = float(np.sum(res[kind].astype(np.float64) * d))
                dx = eps * d if kind == "x0" else 0.0
                dv = eps * d if kind == "v0" else 0.0
                lp, pp = self.oracle(X0 + dx, V0 + dv, self.nu, self.gam, nsteps)
                lm, pm = self.oracle(X0 - dx, V0 - dv, self.nu, self.gam, nsteps)
            else:
                g = res[kind]
                dn = eps if kind == "viscosity" else 0.0
                dg = eps if kind == "surface_tension" else 0.0
                lp, pp = self.oracle(X0, V0, self.nu + dn, self.gam + dg, nsteps)
                lm, pm = self.oracle(X0, V0, self.nu - dn, self.gam - dg, nsteps)
            assert pp == base and pm == base, "finite difference crossed the cohesion jump at h/2 (%s)" % kind
            out.append((kind, g, (lp - lm) / (2 * eps)))
        return out


    def _fd_smooth(self, res, nsteps, seed, ndir, kinds, epss, X0, V0, L0):
        rng = np.random.default_rng(seed)
        want = set(kinds or ("v0", "x0", "viscosity", "surface_tension"))
        todo = [k for k in ["v0"] * ndir + ["x0"] * ndir if k in want]
        todo += [k for k, on in (("viscosity", self.nu), ("surface_tension", self.gam)) if on and k in want]
        out = []
        for kind in todo:
            d = None
            if kind in ("v0", "x0"):
                d = rng.normal(size=self.x0.shape)
                d /= np.linalg.norm(d)
                g = float(np.sum(res[kind].astype(np.float64) * d))
            else:
                g = res[kind]
            tried = []
            found = None
            for eps in epss:
                def loss(sgn):
                    e = sgn * eps
                    if kind == "v0":
                        return self.oracle(X0, V0 + e * d, self.nu, self.gam, nsteps)[0]
                    if kind == "x0":
                        return self.oracle(X0 + e * d, V0, self.nu, self.gam, nsteps)[0]
                    if kind == "viscosity":
                        return self.oracle(X0, V0, self.nu + e, self.gam, nsteps)[0]
                    return self.oracle(X0, V0, self.nu, self.gam + e, nsteps)[0]
                fp, fm = (loss(1) - L0) / eps, (L0 - loss(-1)) / eps
                c = 0.5 * (fp + fm)
                tried.append((eps, fp, fm))
                if abs(fp - fm) <= 1e-3 * abs(c):
                    found = c
                    break
                if len(tried) > 1:
                    (_, fp0, fm0), (_, fp1, fm1) = tried[-2], tried[-1]
                    if abs(fp1 - fp0) <= 1e-4 * abs(fp1):
                        found = fp1
                        break
                    if abs(fm1 - fm0) <= 1e-4 * abs(fm1):
                        found = fm1
                        break
            if found is None:
                raise AssertionError("no eps with a smooth loss for %s: %s" % (kind, tried))
            out.append((kind, g, found))
        return out


def _rel(g, fd):
    return abs(g - fd) / max(abs(fd), 1e-30)


@pytest.mark.parametrize("visc,st", [(True, True), (True, False), (False, True), (False, False)])
@pytest.mark.parametrize("nsteps", [1, 5])
@pytest.mark.parametrize("method", [0, 2], ids=["wcsph", "dfsph"])
def test_adjoint_matches_fp64_finite_differences(tmp_path, visc, st, nsteps, method):
    case = Case(str(tmp_path), _scene(visc, st, method=method))
    try:
        case.gpu_branch(nsteps)
        xT, res = case.gpu_grad(nsteps)
        lg = float(np.mean(np.sum((xT - case.y) ** 2, axis=1)))
        lo, _ = case.oracle(case.x0.astype(np.float64), case.v0.astype(np.float64), case.nu, case.gam, nsteps)
        assert abs(lg - lo) <= 1e-4 * lo
        for kind, g, fd in case.fd_checks(res, nsteps):
            assert _rel(g, fd) <= 1e-3, (kind, g, fd)
    finally:
        case.ctx.close()


def test_adjoint_long_rollout(tmp_path):
    """20 steps, XSPH + surface tension, on the GPU's branches. Also counts
    the particle-steps where the FP64 oracle on its own would take the other
    side of the pressure kink (reported, not asserted)."""
    nsteps = 20
    case = Case(str(tmp_path), _scene(True, True), xseed=2, vseed=102)
    try:
        branch = case.gpu_branch(nsteps)
        o = S.Oracle(case.P, case.x0, case.v0, case.xb, bits=64)
        o.init()
        flips = 0
        for k in range(nsteps):
            o.step(1)
            flips += int(np.sum(branch[k] != (o.state()["p"] > 0)))
        o.close()
        _, res = case.gpu_grad(nsteps)
        for kind, g, fd in case.fd_checks(res, nsteps, ndir=3):
            assert _rel(g, fd) <= 1e-3, (kind, g, fd, "fp32/fp64 kink flips: %d" % flips)
    finally:
        case.ctx.close()


def test_adjoint_dfsph_long_rollout(tmp_path):
    """12 DFSPH steps (8 + 8 Jacobi iterations each), XSPH + surface tension,
    on the GPU's branches."""
    nsteps = 12
    case = Case(str(tmp_path), _scene(True, True, method=2), xseed=2, vseed=102)
    try:
        case.gpu_branch(nsteps)
        _, res = case.gpu_grad(nsteps)
        for kind, g, fd in case.fd_checks(res, nsteps, ndir=3):
            assert _rel(g, fd) <= 1e-3, (kind, g, fd)
    finally:
        case.ctx.close()


@pytest.mark.parametrize("method", [0, 2], ids=["wcsph", "dfsph"])
def test_checkpoint_interval_does_not_change_the_gradient(tmp_path, method):
    """Segment recomputation from checkpoints is bit-exact: the gradients and
    the final state are identical for every checkpoint interval, and the
    recorded forward equals plain stepping. One segment (every = 12) runs the
    sweep on one stream; several segments run each segment's adjoint passes
    beside the recompute of the segment before it, on two streams, and a
    second sweep over the same recording gives the same gradients again."""
    case = Case(str(tmp_path), _scene(True, True, method=method))
    try:
        c, sid = case.ctx, case.sid
        c.set_fluid_state(sid, case.x0, case.v0)
        c.step(12)
        plain = c.fluid_state(sid)
        ref = None
        for every in (1, 5, 12):
            xT, res = case.gpu_grad(12, every)
            st = c.fluid_state(sid)
            assert np.array_equal(st["x"], plain["x"]) and np.array_equal(st["v"], plain["v"])
            if every == 5:
                seed = (2.0 * (xT - case.y) / len(xT)).astype(np.float32)
                again = c.adjoint({sid: (seed, None)})[sid]
                for k in ("x0", "v0"):
                    assert np.array_equal(again[k], res[k]), ("second sweep", k)
                assert again["viscosity"] == res["viscosity"]
            if ref is None:
                ref = res
                continue
            for k in ("x0", "v0"):
                assert np.array_equal(res[k], ref[k]), (every, k)
            assert res["viscosity"] == ref["viscosity"] and res["surface_tension"] == ref["surface_tension"]
    finally:
        case.ctx.close()


@pytest.mark.parametrize("method", [0, 2], ids=["wcsph", "dfsph"])
def test_replayed_segments_follow_a_growing_grid(tmp_path, method):
    """An 11^3 block thrown along +x at 40 m/s (16 steps of 0.5 ms, 0.32 m)
    away from a small wall box behind it: the context's dense grid box spans
    both with a margin of 7 cells of 0.02 m, so after about nine steps more
    than half the overflow buckets' worth of particles (512) sit outside it,
    and the recording's search grows the grid and runs again. A recomputed
    segment repeats such a step's search with a host check and queues the
    others without one; the gradients and the final state equal those of
    one segment (nothing recomputed) bit for bit. Plain stepping from the
    same state ends there too (WCSPH queues its 16 steps as one chunk, whose
    failed search check undoes it and retakes the steps with host checks)."""
    sc = SC.scene_dict(method, R, [((SP, SP, SP), (12 * SP, 12 * SP, 12 * SP))], [((-0.1, -0.1, -0.1), (-0.08, -0.08, -0.08))],
                       dt=5e-4, max_dt=5e-4, min_dt=5e-4, min_it=NIT, max_it=NIT, min_it_v=NIT, max_it_v=NIT,
                       viscosity_method=2, viscosity=0.01, surface_tension_method=3, surface_tension=0.1)
    case = Case(str(tmp_path), sc)
    try:
        case.v0 = (case.v0 + np.array([40.0, 0.0, 0.0], np.float32)).astype(np.float32)
        ref = None
        for every in (16, 5, 3):
            print("checkpoint every", every, flush=True)
            xT, res = case.gpu_grad(16, every)
            assert xT[:, 0].min() > 0.28   # the whole block has left the first dense box (x < 0.26)
            if ref is None:
                ref = (xT, res)
                continue
            assert np.array_equal(xT, ref[0]), every
            for k in ("x0", "v0"):
                assert np.array_equal(res[k], ref[1][k]), (every, k)
            assert res["viscosity"] == ref[1]["viscosity"] and res["surface_tension"] == ref[1]["surface_tension"]
        assert np.isfinite(ref[1]["v0"]).all() and np.abs(ref[1]["v0"]).max() > 0
        case.ctx.set_fluid_state(case.sid, case.x0, case.v0)
        stats = case.ctx.step(16)
        assert np.array_equal(case.ctx.fluid_state(case.sid)["x"].astype(np.float64), ref[0])
        ts = np.array([row[case.sid].t for row in stats])
        assert np.allclose(np.diff(ts), 5e-4, rtol=1e-2) and all(row[case.sid].dt == np.float32(5e-4) for row in stats)
    finally:
        case.ctx.close()


@pytest.mark.parametrize("method", [0, 2], ids=["wcsph", "dfsph"])
def test_batched_adjoint_equals_single_scene(tmp_path, method):
    """Two scenes in one context get the gradients each gets alone, bit for bit."""
    tmp = str(tmp_path)
    a = _scene(True, True, method=method)
    b = _scene(True, True, n=(8, 9, 7), method=method)
    b["Fluid"]["FluidModel"][0]["viscosity"] = 0.03
    b["Fluid"]["FluidModel"][0]["surfaceTension"] = 0.05
    singles = []
    for k, sc in enumerate((a, b)):
        os.makedirs(os.path.join(tmp, str(k)))
        case = Case(os.path.join(tmp, str(k)), sc, xseed=1 + k, vseed=3 + k)
        _, res = case.gpu_grad(6, 4)
        singles.append((case.x0, case.v0, case.y, res))
        case.ctx.close()
    ctx = Context(0)
    try:
        for k, sc in enumerate((a, b)):
            path = os.path.join(tmp, "b%d.json" % k)
            SC.write_json(sc, path)
            ctx.load_scene_json(path)
        ctx.init()
        for k in range(2):
            ctx.set_fluid_state(k, singles[k][0], singles[k][1])
        ctx.forward_record(6, 4)
        seeds = {}
        for k in range(2):
            xT = ctx.fluid_state(k)["x"].astype(np.float64)
            seeds[k] = ((2.0 * (xT - singles[k][2]) / len(xT)).astype(np.float32), None)
        res = ctx.adjoint(seeds)
        for k in range(2):
            ref = singles[k][3]
            for f in ("x0", "v0"):
                assert np.array_equal(res[k][f], ref[f]), (k, f)
            assert res[k]["viscosity"] == ref["viscosity"]
            assert res[k]["surface_tension"] == ref["surface_tension"]
    finally:
        ctx.close()


@pytest.mark.parametrize("method", [0, 2], ids=["wcsph", "dfsph"])
def test_torch_autograd_function(tmp_path, method):
    """loss.backward() through sph2o_amd.autograd.rollout returns the C-ABI gradients."""
    case = Case(str(tmp_path), _scene(True, True, method=method))
    try:
        _, ref = case.gpu_grad(5)
        x0 = torch.tensor(case.x0, requires_grad=True)
        v0 = torch.tensor(case.v0, requires_grad=True)
        nu = torch.tensor(case.nu, dtype=torch.float64, requires_grad=True)
        gam = torch.tensor(case.gam, dtype=torch.float64, requires_grad=True)
        xT, vT = rollout(case.ctx, case.sid, x0, v0, nu, gam, 5)
        y = torch.tensor(case.y)
        loss = ((xT.double() - y) ** 2).sum(dim=1).mean()
        loss.backward()
        assert np.array_equal(v0.grad.numpy(), ref["v0"])
        assert np.array_equal(x0.grad.numpy(), ref["x0"])
        assert nu.grad.item() == pytest.approx(ref["viscosity"], rel=1e-6)
        assert gam.grad.item() == pytest.approx(ref["surface_tension"], rel=1e-6)
    finally:
        case.ctx.close()


def test_adjoint_errors(tmp_path):
    tmp = str(tmp_path)
    case = Case(tmp, _scene(True, True))
    try:
        c, sid = case.ctx, case.sid
        with pytest.raises(_lib.SPHError) as e:     # nothing recorded yet
            c.adjoint({sid: (np.zeros_like(case.x0), None)})
        assert e.value.code == _lib.SPH_E_STATE
        c.forward_record(3, 2)
        c.set_fluid_state(sid, case.x0, case.v0)    # state changed after the recording
        with pytest.raises(_lib.SPHError) as e:
            c.adjoint({sid: (np.zeros_like(case.x0), None)})
        assert e.value.code == _lib.SPH_E_STATE
        with pytest.raises(_lib.SPHError) as e:
            c.forward_record(0, 1)
        assert e.value.code == _lib.SPH_E_ARG
    finally:
        case.ctx.close()
    # PCISPH, DFSPH with a free iteration count or an adaptive dt: not differentiable here
    for k, sc in enumerate((SC.small(1), SC.small(2), _scene(method=2))):
        if k == 2:
            sc["Configuration"]["maxTimeStep"] = 1e-3
        path = os.path.join(tmp, "u%d.json" % k)
        SC.write_json(sc, path)
        ctx = Context(0)
        try:
            ctx.load_scene_json(path)
            ctx.init()
            with pytest.raises(_lib.SPHError) as e:
                ctx.forward_record(2, 1)
            assert e.value.code == _lib.SPH_E_UNSUPPORTED
        finally:
            ctx.close()
