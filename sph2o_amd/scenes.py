"""Scene builders for the benchmark and parity workloads.

Each builder returns a scene dict in the reference's JSON schema
(SceneLoader.cpp:40-209) plus, where the workload needs it, explicit initial
positions/velocities (jitter, clusters). Random numbers come from a
counter-based SplitMix64 stream so every workload is fixed by its seed.

Workloads (sizes as in the project's benchmark plan):
  W8k   WCSPH dam break, 20^3 fluid at r = 0.01 in a 0.8 x 0.4 x 0.4 m box
  D1M   DFSPH dam break, 100^3 fluid at r = 0.005 in a 4 x 2 x 2 m box
  P4M   PCISPH neighbour-search stress case, ~3.9M particles in a 2 m cube
  B64   64 DFSPH dam breaks of log-uniform size 2k..200k
  A100k 100k-particle scene for the differentiable step
"""
import json

import numpy as np

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n):
    """n uniform doubles in [0, 1) from SplitMix64 with the given seed."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = (np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)) & MASK64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & MASK64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & MASK64
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n):
    """Standard normals by Box-Muller on the SplitMix64 stream."""
    u = splitmix64(seed, 2 * ((n + 1) // 2))
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


def scene_dict(method, radius, fluid_blocks, boxes, dt=1e-3, max_dt=1e-3, min_dt=0.0, fps=60.0, cfl=1.0,
               gravity=(0.0, 0.0, -9.81), stiffness=1000.0, gamma=7.0, eta=1e-4, eta_v=1e-3, min_it=2,
               max_it=100, min_it_v=1, max_it_v=100, viscosity_method=-1, viscosity=0.0, boundary_viscosity=0.0,
               surface_tension_method=-1, surface_tension=0.0, density0=1000.0, start=0.0, end=1.0e6):
    c = {"startTime": start, "endTime": end, "timeStep": dt, "fps": fps, "minTimeStep": min_dt,
         "maxTimeStep": max_dt, "gravity": list(gravity), "particleRadius": radius, "simulationMethod": method,
         "boundaryMethod": 2, "cflFactor": cfl}
    if method == 0:
        c.update(stiffness=stiffness, gamma=gamma)
    elif method == 1:
        c.update(eta=eta, minIterations=min_it, maxIterations=max_it)
    else:
        c.update(eta=eta, etaV=eta_v, minIterations=min_it, maxIterations=max_it, minIterationsV=min_it_v,
                 maxIterationsV=max_it_v)
    fm = {"density0": density0, "fluidBlocks": [{"min": list(b[0]), "max": list(b[1])} for b in fluid_blocks]}
    if viscosity_method > -1:
        fm.update(viscosity=viscosity, boundaryViscosity=boundary_viscosity)
    if surface_tension_method > -1:
        fm.update(surfaceTension=surface_tension)
    return {"Configuration": c,
            "Fluid": {"viscosityMethod": viscosity_method, "surfaceTensionMethod": surface_tension_method,
                      "adhesionMethod": -1, "FluidModel": [fm]},
            "Boundary": [{"box": [{"min": list(b[0]), "max": list(b[1])} for b in boxes], "sphere": [],
                          "geometry": []}]}


def add_emitter(scene, position, rotation=(1.0, 0.0, 0.0, 0.0), velocity=1.0, num_particles=70, start_time=0.0,
                type=1, width=0.03, height=0.03, spacing=1.0):
    """Append an emitter of the reference schema (SceneLoader.cpp:125-141) to
    FluidModel[0]; rotation is the quaternion (w, x, y, z). Returns the scene."""
    fm = scene["Fluid"]["FluidModel"][0]
    fm.setdefault("emitters", []).append(
        {"position": list(position), "rotation": list(rotation), "velocity": velocity, "numParticles": num_particles,
         "startTime": start_time, "type": type, "width": width, "height": height, "spacing": spacing})
    return scene


def emitter_scenes():
    """Small emitter scenes (10^3-particle block at r = 0.01 in a 0.6 x 0.3 x 0.4 m
    box, like `small`): a disc emitter shooting along -x (quaternion of a -90
    degree turn about y) and a square one shooting along +y (+90 degrees about x),
    keyed by the solver they run."""
    q_mx = (0.70710678, 0.0, -0.70710678, 0.0)
    q_py = (0.70710678, -0.70710678, 0.0, 0.0)
    dfsph = add_emitter(small(2, viscosity_method=2, viscosity=0.01, surface_tension_method=3, surface_tension=0.1),
                        (0.5, 0.15, 0.25), q_mx, velocity=1.0, num_particles=35, type=1, width=0.03)
    wcsph = add_emitter(small(0, viscosity_method=2, viscosity=0.01), (0.45, 0.05, 0.3), q_py, velocity=2.0,
                        num_particles=24, type=0, width=0.06, height=0.04, spacing=1.0, start_time=2e-4)
    pcisph = add_emitter(small(1, viscosity_method=2, viscosity=0.01), (0.5, 0.15, 0.25), q_mx, velocity=1.0,
                         num_particles=21, type=1, width=0.03, start_time=1e-3)
    return {"dfsph": dfsph, "wcsph": wcsph, "pcisph": pcisph}


def boundary_option_scenes():
    """Small scenes (10^3 block at r = 0.01, like `small`) for the force and
    boundary options around the step: Akinci adhesion (method 4, beta = 5) with
    WCSPH and DFSPH, and cube boundary objects (boundary method 0, no boundary
    particles) with WCSPH (a wall at x = 0.2 cutting the block's last layer)
    and DFSPH (a floor at z = 0.025 under the falling block)."""
    out = {}
    for tag, method in (("adh_wcsph", 0), ("adh_dfsph", 2)):
        extra = dict(surface_tension_method=3, surface_tension=0.1) if method == 2 else {}
        sc = small(method, viscosity_method=2, viscosity=0.01, **extra)
        sc["Fluid"]["adhesionMethod"] = 4
        sc["Fluid"]["FluidModel"][0]["adhesion"] = 5.0
        out[tag] = sc
    for tag, method, lo, hi in (("cube_wcsph", 0, (0.0, 0.0, 0.025), (0.2, 0.3, 0.4)),
                                ("cube_dfsph", 2, (0.0, 0.0, 0.025), (0.6, 0.3, 0.4))):
        sc = small(method, viscosity_method=2, viscosity=0.01)
        sc["Configuration"]["boundaryMethod"] = 0
        sc["Boundary"] = [{"box": [{"min": list(lo), "max": list(hi)}], "sphere": [], "geometry": []}]
        out[tag] = sc
    return out


def sphere_scenes():
    """Akinci sphere boundaries (AkinciBoundaryObject::sampleSphere): a 10^3
    block at r = 0.01 starting 2 cm above a sphere of radius 5 cm, inside a
    0.5 x 0.3 x 0.4 m box; a second boundary object holds a pillar box and a
    sphere of radius 3 cm, so the sampling order (object by object, boxes
    before spheres) is part of what is checked. WCSPH (XSPH 0.01) and DFSPH
    (XSPH 0.01, surface tension 0.1)."""
    out = {}
    for tag, method in (("wcsph", 0), ("dfsph", 2)):
        extra = dict(surface_tension_method=3, surface_tension=0.1) if method == 2 else {}
        sc = small(method, viscosity_method=2, viscosity=0.01, **extra)
        fb = sc["Fluid"]["FluidModel"][0]["fluidBlocks"][0]
        fb["min"], fb["max"] = [0.1, 0.05, 0.14], [0.3, 0.25, 0.34]
        sc["Boundary"] = [
            {"box": [{"min": [0.0, 0.0, 0.0], "max": [0.5, 0.3, 0.4]}],
             "sphere": [{"pos": [0.2, 0.15, 0.07], "radius": 0.05}], "geometry": []},
            {"box": [{"min": [0.38, 0.1, 0.0], "max": [0.44, 0.2, 0.1]}],
             "sphere": [{"pos": [0.41, 0.15, 0.2], "radius": 0.03}], "geometry": []},
        ]
        out[tag] = sc
    return out


def pcisph_boundary_scene(method, boxes=(), spheres=(), block=((0.0, 0.0, 0.0), (0.2, 0.2, 0.2)), normal_fct=0.3,
                          tang_fct=0.95, **kw):
    """A scene of boundary method 1 (PCISPHBoundaryObject): one boundary
    object with the given boxes ((min, max, inverted)) and spheres
    ((centre, radius, inverted)), a block at r = 0.01, XSPH 0.01."""
    sc = small(method, viscosity_method=2, viscosity=0.01, **kw)
    sc["Configuration"]["boundaryMethod"] = 1
    fb = sc["Fluid"]["FluidModel"][0]["fluidBlocks"][0]
    fb["min"], fb["max"] = list(block[0]), list(block[1])
    sc["Boundary"] = [{"normalFct": normal_fct, "tangFct": tang_fct,
                       "box": [{"min": list(lo), "max": list(hi), "inverted": inv} for lo, hi, inv in boxes],
                       "sphere": [{"pos": list(c), "radius": rad, "inverted": inv} for c, rad, inv in spheres],
                       "geometry": []}]
    return sc


def pcisph_boundary_scenes():
    """Boundary method 1 with its defined samples (boxes of at least 1 m per
    axis, spheres): a 1 m container box with the block flush in its corner
    (WCSPH) or one particle radius off the walls (PCISPH: flush, the
    reference's first PCISPH step throws particles out at 1,156 m/s), the
    block resting against a sphere obstacle in the same object as the box
    (WCSPH), and a block inside an inverted sphere of radius 0.2 m (PCISPH);
    DFSPH in the box one particle radius off the walls (flush, the
    reference's DFSPH explodes as well) and in the
    sphere with surface tension."""
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), True)
    return {
        "wcsph_box": pcisph_boundary_scene(0, [box]),
        "pcisph_box": pcisph_boundary_scene(1, [box], block=((0.01, 0.01, 0.01), (0.21, 0.21, 0.21))),
        "wcsph_obstacle": pcisph_boundary_scene(0, [box], [((0.57, 0.5, 0.1), 0.065, False)],
                                                block=((0.3, 0.4, 0.0), (0.5, 0.6, 0.2))),
        "pcisph_sphere": pcisph_boundary_scene(1, [], [((0.2, 0.2, 0.2), 0.2, True)],
                                               block=((0.13, 0.13, 0.03), (0.27, 0.27, 0.17))),
        "dfsph_box": pcisph_boundary_scene(2, [box], block=((0.01, 0.01, 0.01), (0.21, 0.21, 0.21))),
        "dfsph_sphere_st": pcisph_boundary_scene(2, [], [((0.2, 0.2, 0.2), 0.2, True)],
                                                 block=((0.13, 0.13, 0.03), (0.27, 0.27, 0.17)),
                                                 surface_tension_method=3, surface_tension=0.1),
    }


def write_json(scene, path):
    with open(path, "w") as f:
        json.dump(scene, f)


def lattice(mn, counts, spacing, offset):
    """Positions mn + offset + spacing*(i,j,k), k fastest (FluidObject.cpp:44-58 order)."""
    idx = np.stack(np.meshgrid(*[np.arange(c) for c in counts], indexing="ij"), -1).reshape(-1, 3)
    return (np.asarray(mn, np.float64) + offset + spacing * idx).astype(np.float32)


def jitter(x, amp, seed):
    u = splitmix64(seed, x.size).reshape(x.shape)
    return (x.astype(np.float64) + (2.0 * u - 1.0) * amp).astype(np.float32)


# ---------------------------------------------------------------- workloads
def w8k():
    r = 0.01
    sc = scene_dict(0, r, [((0, 0, 0), (0.4, 0.4, 0.4))], [((0, 0, 0), (0.8, 0.4, 0.4))], dt=1e-4,
                    viscosity_method=2, viscosity=0.01)
    x = lattice((0, 0, 0), (20, 20, 20), 2 * r, r)
    return sc, jitter(x, 0.01 * 2 * r, 1), None


def d1m(n=100, surface_tension=0.1, gap=None):
    """DFSPH dam break: n^3 fluid at r = 0.005 in a 4 x 2 x 2 m box (n=100: 1M).

    The block sits in a corner one lattice spacing s = 2r away from the three
    walls (gap=None), so its first layer is 1.5 s from the boundary particles.
    Flush against the walls (gap=0: first layer at r) the reference itself
    diverges: its first step stops at 100 density iterations with 1.07% error
    and dt collapses to 4.7e-5 s in step 2."""
    r = 0.005
    s = 2 * r
    g = s if gap is None else gap
    L = n * s
    sc = scene_dict(2, r, [((g, g, g), (g + L, g + L, g + L))], [((0, 0, 0), (4.0 * L, 2.0 * L, 2.0 * L))], dt=1e-3,
                    max_dt=1e-3, eta=1e-4, eta_v=1e-3, min_it=2, max_it=100, min_it_v=1, max_it_v=100,
                    surface_tension_method=3, surface_tension=surface_tension)
    return sc, None, None


def p4m(n_side=156, cluster=71):
    """PCISPH stress case: jittered lattice in a 2 m cube with a 1.5x-dense cluster."""
    r = 0.0064
    s = 2 * r
    box = 2.0 * n_side / 156.0
    x = lattice((0, 0, 0), (n_side,) * 3, s, r).astype(np.float64)
    x = x + (2.0 * splitmix64(3, x.size).reshape(x.shape) - 1.0) * 0.2 * s
    lo, hi = 0.3 * box, 0.7 * box
    keep = ~np.all((x >= lo) & (x < hi), axis=1)
    x = x[keep]
    cs = s / 1.5 ** (1.0 / 3.0)
    c = lattice((lo, lo, lo), (cluster,) * 3, cs, cs / 2).astype(np.float64)
    c = c + (2.0 * splitmix64(4, c.size).reshape(c.shape) - 1.0) * 0.2 * cs
    x = np.concatenate([x, c]).astype(np.float32)
    v = (0.5 * normal(5, x.size)).reshape(x.shape).astype(np.float32)
    sc = scene_dict(1, r, [((0, 0, 0), (r, r, r))], [((0, 0, 0), (box, box, box))], dt=5e-4, eta=0.01, min_it=3,
                    max_it=50)
    return sc, x, v


def b64(seed=6, count=64, lo=2e3, hi=2e5):
    """Batched sweep: `count` DFSPH dam breaks with log-uniform sizes."""
    u = splitmix64(seed, 4 * count).reshape(count, 4)
    out = []
    r = 0.005
    s = 2 * r
    for k in range(count):
        N = np.exp(np.log(lo) + u[k, 0] * (np.log(hi) - np.log(lo)))
        n = max(2, int(round(N ** (1.0 / 3.0))))
        L = n * s
        nu = float(np.exp(np.log(1e-3) + u[k, 1] * (np.log(1e-1) - np.log(1e-3))))
        gam = float(0.2 * u[k, 2])
        # block one spacing from the walls (see d1m); the box keeps ~25% filling
        sc = scene_dict(2, r, [((s, s, s), (s + L, s + L, s + L))], [((0, 0, 0), (2 * L + 2 * s, 2 * L + 2 * s, L + 2 * s))],
                        dt=1e-3, max_dt=1e-3,
                        eta=1e-4, eta_v=1e-3, viscosity_method=2, viscosity=nu, surface_tension_method=3,
                        surface_tension=gam)
        out.append(sc)
    return out


def a100k(method=0, n=(50, 50, 40), dt=5e-4):
    """Differentiable-rollout workload: an n[0] x n[1] x n[2] block (default
    100,000) at r = 0.005, one spacing from the walls of a 1 x 1 x 0.6 m box,
    fixed dt, XSPH viscosity 0.01 and Akinci surface tension 0.1. method 0:
    WCSPH (k = 1000, gamma = 7); method 2: DFSPH with 8 density and 8
    divergence iterations and the CFL rule off."""
    r = 0.005
    s = 2 * r
    hi = tuple(s + s * c for c in n)
    box = (max(1.0, hi[0] + s), max(1.0, hi[1] + s), max(0.6, hi[2] + s))
    sc = scene_dict(method, r, [((s, s, s), hi)], [((0, 0, 0), box)], dt=dt, max_dt=dt, min_dt=dt,
                    min_it=8, max_it=8, min_it_v=8, max_it_v=8, viscosity_method=2, viscosity=0.01,
                    surface_tension_method=3, surface_tension=0.1)
    return sc, None, None


def adjoint_target(x0, seed=7):
    """Loss target y = x0 + N(0, (0.5 s)^2) per coordinate, s = 0.01."""
    return (np.asarray(x0, np.float64) + 0.005 * normal(seed, x0.size).reshape(x0.shape))


def small(method, n=10, radius=0.01, box=(0.6, 0.3, 0.4), gap=0.02, **kw):
    """Small test scene: n^3 lattice starting `gap` from the walls."""
    L = gap + 2 * radius * n
    args = dict(dt=1e-4 if method == 0 else (5e-4 if method == 1 else 1e-3), max_dt=5e-3,
                eta=1e-3 if method == 2 else 1e-2, eta_v=1e-3, min_it=2 if method == 2 else 3,
                max_it=100 if method == 2 else 50)
    args.update(kw)
    return scene_dict(method, radius, [((gap, gap, gap), (L, L, L))], [((0, 0, 0), tuple(box))], **args)
