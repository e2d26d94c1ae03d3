"""Record golden vectors from the compiled reference (oracle/_ref/sph2o_ref).

Runs only where the reference tree and its headless build exist (this
container): ``make -C oracle ref && python tests/golden/make_golden.py``.
Writes small .npz fixtures next to this script; the tests read only those
fixtures, never the reference itself.

Fixtures:
  ref_cell_keys.npz      probe positions (incl. values within a few ulps of a
                         cell face) and the reference's int3 cell keys
  ref_support_pairs.npz  particle pairs at |r| ~ h whose membership depends on
                         how |r| is rounded, and the reference's answer
  ref_neighbors_*.npz    within-support neighbour sets (fluid and boundary) of
                         a perfect and a jittered lattice
  ref_traj_*.npz         initial state, boundary, Akinci volumes, final state
                         and per-step stats of short runs of each solver
  ref_example_scene.npz  the reference's own example scene
                         (data/example_scene/example_scene.json): sampled
                         fluid/boundary positions, volumes, one DFSPH step
  ref_example_wcsph.npz  the same scene stepped twice by WCSPH (no DFSPH
                         gates): a tight check of the force passes
  ref_emit_*.npz         emitter scenes (sph2o_amd.scenes.emitter_scenes) run
                         from the reference's own sampling: the state after
                         the first emission and after a longer run, with the
                         active count after every step

  ref_w8k.npz            the W8k dam break (8,000 particles, WCSPH): the state
                         after 20 steps and box aggregates after 2,000
  ref_d1m.npz            the D1M dam break (1M particles, DFSPH): per-step
                         stats of 3 steps, 4,000 sampled particles after
                         steps 1 and 3
  ref_b64_smallest.npz   the smallest B64 scene alone: 1 step and 10 steps
  ref_frames_wcsph.npz   frame output of SPHSimulation::step() (the WCSPH
                         emitter scene at 1000 fps, 120 steps): the steps that
                         wrote a frame and every frame file's name and fields

  ref_sphere_*.npz       Akinci sphere boundaries (two boundary objects, box
                         and sphere each): sampled boundary, volumes, the
                         state after one step and after a short run

  ref_pcisph_boundary_samples.npz
                         PCISPH boundary objects as the reference samples
                         them: positions and normals of boxes (non-integer
                         corners, a truncated extent), spheres, and the
                         degenerate sub-metre box
  ref_pcib_*.npz         PCISPH boundary runs (WCSPH and PCISPH, a container
                         box, a sphere obstacle, an inverted sphere): state
                         and boundary densities after 1 step and a short run

  ref_opt_*.npz          Akinci adhesion and cube-boundary scenes
                         (sph2o_amd.scenes.boundary_option_scenes): the state
                         after one step and after a short run

``python tests/golden/make_golden.py emitters`` (or ``options``, ``w8k``,
``d1m``, ``b64``, ``frames``, ``spheres``,
``pcisph_boundary``) writes only those files.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from oracle import sphref as S  # noqa: E402
from sph2o_amd import scenes as SC  # noqa: E402

f32 = np.float32


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


def scene_json(sc):
    return np.array(json.dumps(sc))


def cell_keys():
    sc = SC.small(2, n=10)
    base = S.run_reference(sc, 0)
    n = len(base["x0"])
    h = f32(4) * f32(0.01)
    rng = np.random.default_rng(11)
    pts = []
    # coordinates within a few ulps of cell faces k*h, on every axis
    for k in range(1, 10):
        face = f32(k) * h
        for d in range(-6, 7):
            v = face
            for _ in range(abs(d)):
                v = np.nextafter(v, f32(np.inf) if d > 0 else f32(-np.inf))
            pts.append(v)
    pts = np.array(pts, f32)
    x = np.zeros((n, 3), f32)
    m = len(pts)
    x[:m, 0] = pts
    x[:m, 1] = rng.uniform(0.05, 0.25, m)
    x[:m, 2] = rng.uniform(0.05, 0.35, m)
    x[m:2 * m, 1] = pts
    x[m:2 * m, 0] = rng.uniform(0.05, 0.55, m)
    x[m:2 * m, 2] = rng.uniform(0.05, 0.35, m)
    x[2 * m:3 * m, 2] = pts
    x[2 * m:3 * m, 0] = rng.uniform(0.05, 0.55, m)
    x[2 * m:3 * m, 1] = rng.uniform(0.05, 0.25, m)
    rest = n - 3 * m
    x[3 * m:] = rng.uniform([0.0, 0.0, 0.0], [0.6, 0.3, 0.4], (rest, 3)).astype(f32)
    out = S.run_reference(sc, 0, x=x, neighbors=True)
    save("ref_cell_keys.npz", scene=scene_json(sc), x=x, keys=out["keys"])


def support_pairs():
    """Pairs whose |r| <= h outcome differs between plausible roundings."""
    from fractions import Fraction

    def fmaf(a, b, c):
        return f32(float(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))))

    h = f32(4) * f32(0.01)
    sc = SC.small(2, n=10)
    base = S.run_reference(sc, 0)
    n = len(base["x0"])
    rng = np.random.default_rng(2)
    pairs = []
    while len(pairs) < n // 2:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        k = len(pairs)
        c = np.array([0.03 + 0.054 * (k % 10), 0.03 + 0.024 * ((k // 10) % 10), 0.03 + 0.068 * (k // 100)], f32)
        b = (c + d * float(h)).astype(f32)
        for ax in range(3):
            b[ax] = np.nextafter(b[ax], f32(np.inf) if rng.random() < 0.5 else f32(-np.inf))
        r = (c - b).astype(f32)
        plain = np.sqrt(f32(f32(r[0] * r[0] + r[1] * r[1]) + r[2] * r[2])) <= h
        fz = np.sqrt(fmaf(r[2], r[2], fmaf(r[1], r[1], f32(r[0] * r[0])))) <= h
        fx = np.sqrt(fmaf(r[0], r[0], fmaf(r[2], r[2], f32(r[1] * r[1])))) <= h
        if len({bool(plain), bool(fz), bool(fx)}) > 1:
            pairs.append((c, b))
    x = base["x0"].copy()
    for k, (a, b) in enumerate(pairs):
        x[2 * k] = a
        x[2 * k + 1] = b
    out = S.run_reference(sc, 0, x=x, neighbors=True)
    off, ids = out["nbr_off"], out["nbr_ids"]
    member = np.array([(2 * k + 1) in set(ids[off[2 * k]:off[2 * k + 1]].tolist()) for k in range(len(pairs))])
    a = np.array([p[0] for p in pairs], f32)
    b = np.array([p[1] for p in pairs], f32)
    save("ref_support_pairs.npz", a=a, b=b, h=np.array(h), member=member)


def neighbors(tag, jit):
    sc = SC.small(2, n=10, surface_tension_method=3, surface_tension=0.1)
    base = S.run_reference(sc, 0)
    x = base["x0"] if jit == 0 else SC.jitter(base["x0"], jit * 0.02, 7)
    out = S.run_reference(sc, 0, x=x, neighbors=True)
    save("ref_neighbors_%s.npz" % tag, scene=scene_json(sc), x=x, bx=out["bx"], keys=out["keys"],
         off=out["nbr_off"], ids=out["nbr_ids"], boff=out["bnbr_off"], bids=out["bnbr_ids"])


def trajectory(tag, sc, nsteps, jit=0.01, seed=1):
    base = S.run_reference(sc, 0)
    x0 = SC.jitter(base["x0"], jit * 2 * sc["Configuration"]["particleRadius"], seed)
    out = S.run_reference(sc, nsteps, x=x0)
    save("ref_traj_%s.npz" % tag, scene=scene_json(sc), nsteps=np.array(nsteps), x0=x0, bx=out["bx"],
         bvol=out["bvol"], rho0=out["rho0"], alpha0=out["alpha0"], x=out["x"], v=out["v"], rho=out["rho"],
         alpha=out["alpha"], stats=out["stats"], meta=out["meta"])


def example_scene():
    path = "/root/reference/data/example_scene/example_scene.json"
    with open(path) as f:
        sc = json.load(f)
    out = S.run_reference(sc, 1)
    save("ref_example_scene.npz", scene=scene_json(sc), x0=out["x0"], bx=out["bx"], bvol=out["bvol"],
         rho0=out["rho0"], alpha0=out["alpha0"], x=out["x"], v=out["v"], rho=out["rho"], alpha=out["alpha"],
         stats=out["stats"])


def example_scene_wcsph():
    """The example scene stepped by WCSPH: no DFSPH gates, so the force
    passes (artificial viscosity with boundary term, surface tension) can be
    compared tightly on the reference's own lattice."""
    path = "/root/reference/data/example_scene/example_scene.json"
    with open(path) as f:
        sc = json.load(f)
    c = sc["Configuration"]
    c.update(simulationMethod=0, stiffness=1000.0, gamma=7.0, timeStep=1e-4)
    out = S.run_reference(sc, 2)
    save("ref_example_wcsph.npz", scene=scene_json(sc), nsteps=np.array(2), x0=out["x0"], bx=out["bx"],
         bvol=out["bvol"], x=out["x"], v=out["v"], rho=out["rho"], stats=out["stats"])


def emitters():
    """Emitter scenes: state right after the first emission (step `first`) and
    after `nsteps`, from the reference's own lattice (no jitter: the driver's
    --init needs the full particle count, emitted ones included)."""
    # DFSPH runs to just past its third emission (step 25): at step 29 the
    # divergence loop of the two builds stops one iteration apart (a count
    # decided at the tolerance), after which dt and the states part ways
    plan = {"dfsph": (1, 26), "wcsph": (3, 120), "pcisph": (3, 100)}
    for tag, sc in SC.emitter_scenes().items():
        first, nsteps = plan[tag]
        a = S.run_reference(sc, first)
        b = S.run_reference(sc, nsteps)
        save("ref_emit_%s.npz" % tag, scene=scene_json(sc), first=np.array(first), nsteps=np.array(nsteps),
             x0=a["x0"], bx=a["bx"], x1=a["x"], v1=a["v"], nact1=a["nact"], x=b["x"], v=b["v"], rho=b["rho"],
             stats=b["stats"], nact=b["nact"], meta=b["meta"])


def boundary_options():
    """Adhesion and cube-boundary scenes (sph2o_amd.scenes.boundary_option_scenes)
    from the reference's own lattice: the state after 1 step and after a run."""
    plan = {"adh_wcsph": 20, "adh_dfsph": 20, "cube_wcsph": 20, "cube_dfsph": 40}
    for tag, sc in SC.boundary_option_scenes().items():
        a = S.run_reference(sc, 1)
        b = S.run_reference(sc, plan[tag])
        save("ref_opt_%s.npz" % tag, scene=scene_json(sc), nsteps=np.array(plan[tag]), x0=a["x0"], bx=a["bx"],
             x1=a["x"], v1=a["v"], a1=a["a"], rho1=a["rho"], x=b["x"], v=b["v"], rho=b["rho"], stats=b["stats"])


def spheres():
    """Sphere-boundary scenes (sph2o_amd.scenes.sphere_scenes) from the
    reference's own lattice: the sampled boundary (positions in object order,
    Akinci volumes), the state after 1 step and after a run."""
    plan = {"wcsph": 20, "dfsph": 10}
    for tag, sc in SC.sphere_scenes().items():
        a = S.run_reference(sc, 1)
        b = S.run_reference(sc, plan[tag])
        save("ref_sphere_%s.npz" % tag, scene=scene_json(sc), nsteps=np.array(plan[tag]), x0=a["x0"], bx=a["bx"],
             bvol=a["bvol"], rho0=a["rho0"], x1=a["x"], v1=a["v"], rho1=a["rho"], x=b["x"], v=b["v"], rho=b["rho"],
             stats=b["stats"])


def pcisph_boundary():
    """Boundary method 1 (PCISPHBoundaryObject). Samples: boxes with
    non-integer corners (the extents truncate to whole metres), a box with
    two spheres in one object, and the sub-metre box the reference samples as
    one particle at 0/0. Runs (sph2o_amd.scenes.pcisph_boundary_scenes, the
    reference's own lattice): the state and boundary densities after 1 step
    and after a short run."""
    samples = {
        "offset": SC.pcisph_boundary_scene(1, [((0.13, 0.07, 0.011), (1.37, 1.21, 1.49), False)]),
        "wide": SC.pcisph_boundary_scene(1, [((-0.3, 0.05, 0.2), (2.0, 1.2, 3.37), False)]),
        "truncated": SC.pcisph_boundary_scene(1, [((0.0, 0.0, 0.0), (1.5, 1.2, 1.0), True)]),
        "spheres": SC.pcisph_boundary_scene(1, [((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), True)],
                                            [((0.5, 0.5, 0.5), 0.1, False), ((0.3, 0.3, 0.3), 0.05, True)]),
        "degenerate": SC.pcisph_boundary_scene(1, [((0.0, 0.0, 0.0), (0.6, 0.3, 0.4), True)]),
    }
    arrays = {}
    for tag, sc in samples.items():
        r = S.run_reference(sc, 0)
        arrays[tag + "_scene"] = scene_json(sc)
        arrays[tag + "_bx"] = r["pbx"]
        arrays[tag + "_bn"] = r["pbn"]
    save("ref_pcisph_boundary_samples.npz", **arrays)
    plan = {"wcsph_box": 20, "pcisph_box": 20, "wcsph_obstacle": 20, "pcisph_sphere": 40, "dfsph_box": 10,
            "dfsph_sphere_st": 20}
    for tag, sc in SC.pcisph_boundary_scenes().items():
        a = S.run_reference(sc, 1)
        b = S.run_reference(sc, plan[tag])
        save("ref_pcib_%s.npz" % tag, scene=scene_json(sc), nsteps=np.array(plan[tag]), x0=a["x0"], bx=a["pbx"],
             bn=a["pbn"], x1=a["x"], v1=a["v"], rho1=a["rho"], brho1=a["pbrho"], stats1=a["stats"], x=b["x"],
             v=b["v"], rho=b["rho"], brho=b["pbrho"], stats=b["stats"])


def _box_aggregates(x, v, rho, box):
    """inside count, kinetic sum |v|^2, centre of mass and mean density of the
    particles still inside the box."""
    inside = np.all((x >= 0) & (x <= np.asarray(box, np.float32)), axis=1)
    return np.array([inside.sum(), (v.astype(np.float64) ** 2).sum(), *x[inside].astype(np.float64).mean(0),
                     rho[inside].astype(np.float64).mean()])


def w8k():
    """The W8k workload (sph2o_amd.scenes.w8k: 20^3 jittered lattice flush in the
    corner of a 0.8 x 0.4 x 0.4 m box, WCSPH k = 1000, XSPH 0.01, dt = 1e-4):
    the state after 20 steps, and the box aggregates after the full 2,000."""
    sc, x0, _ = SC.w8k()
    a = S.run_reference(sc, 20, x=x0, threads=8)
    b = S.run_reference(sc, 2000, x=x0, threads=8, timeout=3600)
    save("ref_w8k.npz", scene=scene_json(sc), x0=x0, bx=a["bx"], bvol=a["bvol"], nsteps=np.array(20), x=a["x"],
         v=a["v"], rho=a["rho"], long_steps=np.array(2000),
         long_aggregates=_box_aggregates(b["x"], b["v"], b["rho"], (0.8, 0.4, 0.4)))


def d1m(nsteps=3, nsample=4000):
    """The D1M dam break (sph2o_amd.scenes.d1m: 1,000,000 fluid + 400,002
    boundary particles, DFSPH with surface tension, CFL dt): per-step stats of
    the first `nsteps` steps and the state of `nsample` random particles
    (their sampled positions and initial densities, and `nsample` sampled
    boundary positions) after step 1 and after step `nsteps` (8 OpenMP
    threads)."""
    sc, _, _ = SC.d1m()
    a = S.run_reference(sc, 1, threads=8, timeout=3600)
    b = S.run_reference(sc, nsteps, threads=8, timeout=3600)
    idx = np.sort(np.random.default_rng(12).choice(len(a["x0"]), nsample, replace=False))
    save("ref_d1m.npz", scene=scene_json(sc), nsteps=np.array(nsteps), threads=np.array(8), idx=idx,
         x0=a["x0"][idx], bx=a["bx"][idx % len(a["bx"])], bidx=idx % len(a["bx"]), rho0=a["rho0"][idx], x1=a["x"][idx], v1=a["v"][idx], rho1=a["rho"][idx], stats1=a["stats"],
         x=b["x"][idx], v=b["v"][idx], rho=b["rho"][idx], stats=b["stats"],
         mean_density_error=np.array(np.mean(np.maximum(b["rho"].astype(np.float64) / 1000.0 - 1.0, 0.0))))


def b64_smallest(nsteps=10):
    """The smallest scene of the B64 sweep (sph2o_amd.scenes.b64) run alone:
    the state after one step and the stats and state after `nsteps`."""
    scenes = SC.b64()
    sizes = []
    for sc in scenes:
        fb = sc["Fluid"]["FluidModel"][0]["fluidBlocks"][0]
        sizes.append(round((fb["max"][0] - fb["min"][0]) / 0.01) ** 3)
    k = int(np.argmin(sizes))
    sc = scenes[k]
    a = S.run_reference(sc, 1)
    b = S.run_reference(sc, nsteps)
    save("ref_b64_smallest.npz", scene=scene_json(sc), index=np.array(k), nsteps=np.array(nsteps), x0=a["x0"],
         bx=a["bx"], x1=a["x"], v1=a["v"], stats1=a["stats"], x=b["x"], v=b["v"], rho=b["rho"], stats=b["stats"])


def frame_output(nsteps=120):
    """SPHSimulation::step() with frame output on (SceneLoader::writeFluid) for
    the WCSPH emitter scene at 1000 frames per second: per step whether a
    frame was written and the frame counter, and every frame's file name,
    header fields, positions and velocities."""
    sc = SC.emitter_scenes()["wcsph"]
    sc["Configuration"]["fps"] = 1000.0
    r = S.run_reference(sc, nsteps, frames=True)
    names = sorted(r["frames"])
    fr = [r["frames"][n] for n in names]
    save("ref_frames_wcsph.npz", scene=scene_json(sc), nsteps=np.array(nsteps), saved=r["saved"],
         names=np.array(names), num_particles=np.array([f["numParticles"] for f in fr]),
         num_active=np.array([f["numActiveParticles"] for f in fr]),
         time=np.array([f["time"] for f in fr]), time_step=np.array([f["timeStep"] for f in fr]),
         positions=np.array([f["positions"] for f in fr], np.float32),
         velocities=np.array([f["velocities"] for f in fr], np.float32))


def main():
    if not S.reference_available():
        raise SystemExit("compiled reference missing: make -C oracle ref")
    if sys.argv[1:] == ["emitters"]:
        emitters()
        return
    if sys.argv[1:] == ["options"]:
        boundary_options()
        return
    if sys.argv[1:] == ["w8k"]:
        w8k()
        return
    if sys.argv[1:] == ["d1m"]:
        d1m()
        return
    if sys.argv[1:] == ["b64"]:
        b64_smallest()
        return
    if sys.argv[1:] == ["frames"]:
        frame_output()
        return
    if sys.argv[1:] == ["spheres"]:
        spheres()
        return
    if sys.argv[1:] == ["pcisph_boundary"]:
        pcisph_boundary()
        return
    cell_keys()
    support_pairs()
    neighbors("lattice", 0.0)
    neighbors("jitter", 0.01)
    visc = dict(viscosity_method=2, viscosity=0.01, surface_tension_method=3, surface_tension=0.1)
    trajectory("wcsph20", SC.small(0, **visc), 20)
    trajectory("pcisph20", SC.small(1, viscosity_method=2, viscosity=0.01), 20)
    trajectory("dfsph1", SC.small(2, **visc), 1)
    trajectory("dfsph5", SC.small(2, **visc), 5)
    trajectory("dfsph40", SC.small(2, **visc), 40)
    trajectory("dfsph20_art", SC.small(2, viscosity_method=1, viscosity=0.03, boundary_viscosity=0.002,
                                       surface_tension_method=3, surface_tension=0.15), 20)
    trajectory("wcsph20_std", SC.small(0, viscosity_method=0, viscosity=0.5, boundary_viscosity=0.1), 20)
    example_scene()
    example_scene_wcsph()
    emitters()
    boundary_options()
    w8k()
    d1m()
    b64_smallest()
    frame_output()
    spheres()
    pcisph_boundary()


if __name__ == "__main__":
    main()
