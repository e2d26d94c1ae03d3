"""ctypes front end of the sphref oracle and of the compiled-reference driver.

TEST INFRASTRUCTURE: imported only by tests/, ``__graft_entry__.smoke()`` and
bench.py's CPU-baseline leg, as the checker. The product package
(``sph2o_amd``) never imports this module.

* :class:`Oracle` drives ``libsphref.so`` (oracle/sphref.hpp), the clean-room
  FP64/FP32 restatement of the reference step.
* :func:`run_reference` runs ``oracle/_ref/sph2o_ref``, the original SPH2O
  sources compiled with a headless driver (oracle/ref_driver.cpp), when that
  binary exists.
"""
import ctypes
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsphref.so")
REF_BIN = os.path.join(HERE, "_ref", "sph2o_ref")


class Params(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32), ("boundary_method", ctypes.c_int32),
        ("viscosity_method", ctypes.c_int32), ("st_method", ctypes.c_int32),
        ("min_it", ctypes.c_int32), ("max_it", ctypes.c_int32),
        ("min_it_v", ctypes.c_int32), ("max_it_v", ctypes.c_int32),
        ("adaptive_dt", ctypes.c_int32), ("pad0", ctypes.c_int32),
        ("radius", ctypes.c_double), ("density0", ctypes.c_double),
        ("gx", ctypes.c_double), ("gy", ctypes.c_double), ("gz", ctypes.c_double),
        ("dt0", ctypes.c_double), ("min_dt", ctypes.c_double), ("max_dt", ctypes.c_double),
        ("fps", ctypes.c_double), ("cfl", ctypes.c_double),
        ("stiffness", ctypes.c_double), ("gamma", ctypes.c_double),
        ("eta", ctypes.c_double), ("eta_v", ctypes.c_double),
        ("viscosity", ctypes.c_double), ("bviscosity", ctypes.c_double),
        ("surface_tension", ctypes.c_double),
        ("adhesion_on", ctypes.c_int32), ("n_cubes", ctypes.c_int32),
        ("adhesion", ctypes.c_double),
        ("cubes", (ctypes.c_double * 8) * 4),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.sphref_create.restype = ctypes.c_void_p
        L.sphref_create.argtypes = [ctypes.c_int, P(Params), ctypes.c_int, P(ctypes.c_float), P(ctypes.c_float),
                                    ctypes.c_int, P(ctypes.c_float)]
        L.sphref_destroy.argtypes = [ctypes.c_void_p]
        L.sphref_init.argtypes = [ctypes.c_void_p]
        L.sphref_step.restype = ctypes.c_double
        L.sphref_step.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_double)]
        L.sphref_get.argtypes = [ctypes.c_void_p] + [P(ctypes.c_double)] * 6
        L.sphref_get_boundary.argtypes = [ctypes.c_void_p, P(ctypes.c_double)]
        L.sphref_scalars.argtypes = [ctypes.c_void_p, P(ctypes.c_double)]
        L.sphref_cell_keys.argtypes = [ctypes.c_void_p, P(ctypes.c_int32)]
        L.sphref_neighbors.restype = ctypes.c_int64
        L.sphref_neighbors.argtypes = [ctypes.c_void_p, P(ctypes.c_int64), P(ctypes.c_int32), ctypes.c_int64,
                                       P(ctypes.c_int64), P(ctypes.c_int32), ctypes.c_int64, P(ctypes.c_int64)]
        L.sphref_kernel_eval.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         P(ctypes.c_double), P(ctypes.c_double)]
        L.sphref_sample_block.restype = ctypes.c_int64
        L.sphref_sample_block.argtypes = [P(ctypes.c_float), P(ctypes.c_float), ctypes.c_float, P(ctypes.c_float),
                                          ctypes.c_int64]
        L.sphref_sample_box.restype = ctypes.c_int64
        L.sphref_sample_box.argtypes = L.sphref_sample_block.argtypes
        L.sphref_sample_sphere.restype = ctypes.c_int64
        L.sphref_sample_sphere.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_float, ctypes.c_float,
                                           ctypes.POINTER(ctypes.c_float), ctypes.c_int64]
        L.sphref_set_state64.argtypes = [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_double)]
        L.sphref_cohesion_print.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64)]
        L.sphref_set_pressure_branch.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_uint8)]
        L.sphref_set_cohesion_positions.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_float)]
        L.sphref_set_dfsph_branch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [P(ctypes.c_float)] * 4
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t)) if a is not None else None


def f32(x):
    """Value as the reference stores it (Real = float), widened to double."""
    return float(np.float32(x))


def params_from_scene(scene, adaptive_dt=True):
    """Oracle parameters from a scene dict of the reference JSON schema
    (SceneLoader.cpp:40-146). Every value is rounded to FP32 first, as the
    reference stores it, so the FP64 oracle simulates the same constants."""
    c = scene["Configuration"]
    fl = scene["Fluid"]
    fm = fl["FluidModel"][0]
    p = Params()
    p.method = int(c["simulationMethod"])
    p.boundary_method = int(c["boundaryMethod"])
    p.viscosity_method = int(fl["viscosityMethod"])
    p.st_method = int(fl["surfaceTensionMethod"])
    p.min_it = int(c.get("minIterations", 3))
    p.max_it = int(c.get("maxIterations", 100))
    p.min_it_v = int(c.get("minIterationsV", 1))
    p.max_it_v = int(c.get("maxIterationsV", 100))
    p.adaptive_dt = 1 if adaptive_dt else 0
    p.radius = f32(c["particleRadius"])
    p.density0 = f32(fm["density0"])
    p.gx, p.gy, p.gz = (f32(g) for g in c["gravity"])
    p.dt0 = f32(c["timeStep"])
    p.min_dt = f32(c["minTimeStep"])
    p.max_dt = f32(c["maxTimeStep"])
    p.fps = f32(c["fps"])
    p.cfl = f32(c["cflFactor"])
    p.stiffness = f32(c.get("stiffness", 0.0))
    p.gamma = f32(c.get("gamma", 1.0))
    p.eta = f32(c.get("eta", 0.01))
    p.eta_v = f32(c.get("etaV", 0.01))
    p.viscosity = f32(fm.get("viscosity", 0.0)) if p.viscosity_method > -1 else 0.0
    p.bviscosity = f32(fm.get("boundaryViscosity", 0.0)) if p.viscosity_method > -1 else 0.0
    p.surface_tension = f32(fm.get("surfaceTension", 0.0)) if p.st_method > -1 else 0.0
    # adhesion exists for the Akinci method (4) with a nonzero coefficient (SPHSimulation.cpp:396-397,
    # FluidObject::setAdhesionForce); cube objects: the first box of every Boundary entry, both
    # factors 0 as the reference's loader leaves them (SceneLoader.cpp:168-172, SPHSimulation.cpp:457-474)
    adh = f32(fm.get("adhesion", 0.0)) if int(fl.get("adhesionMethod", -1)) > -1 else 0.0
    p.adhesion_on = int(int(fl.get("adhesionMethod", -1)) == 4 and adh != 0.0)
    p.adhesion = adh
    if p.boundary_method == 0:
        boxes = [b["box"][0] for b in scene.get("Boundary", []) if b.get("box")]
        p.n_cubes = len(boxes)
        for k, b in enumerate(boxes[:4]):
            p.cubes[k][:] = [f32(q) for q in list(b["min"]) + list(b["max"])] + [0.0, 0.0]
    return p


def sample_block(mn, mx, radius):
    mn = np.asarray(mn, np.float32)
    mx = np.asarray(mx, np.float32)
    cap = 1
    while True:
        out = np.zeros((cap, 3), np.float32)
        n = lib().sphref_sample_block(_p(mn, ctypes.c_float), _p(mx, ctypes.c_float), np.float32(radius),
                                      _p(out, ctypes.c_float), cap)
        if n <= cap:
            return out[:n]
        cap = int(n)


def sample_box(mn, mx, radius):
    mn = np.asarray(mn, np.float32)
    mx = np.asarray(mx, np.float32)
    cap = 1
    while True:
        out = np.zeros((cap, 3), np.float32)
        n = lib().sphref_sample_box(_p(mn, ctypes.c_float), _p(mx, ctypes.c_float), np.float32(radius),
                                    _p(out, ctypes.c_float), cap)
        if n <= cap:
            return out[:n]
        cap = int(n)


def sample_sphere(centre, R, radius):
    c = np.asarray(centre, np.float32)
    cap = 1
    while True:
        out = np.zeros((cap, 3), np.float32)
        n = lib().sphref_sample_sphere(_p(c, ctypes.c_float), np.float32(R), np.float32(radius),
                                       _p(out, ctypes.c_float), cap)
        if n <= cap:
            return out[:n]
        cap = int(n)


def kernel_eval(bits, h, mode, r):
    r = np.ascontiguousarray(r, np.float64)
    out = np.zeros_like(r)
    lib().sphref_kernel_eval(bits, float(h), mode, len(r), _p(r, ctypes.c_double), _p(out, ctypes.c_double))
    return out


class Oracle:
    """One scene in the CPU restatement. bits=64 is the FP64 oracle."""

    def __init__(self, params, xf, vf, xb, bits=64):
        self.bits = bits
        self.params = params
        self.nf = len(xf)
        self.nb = len(xb)
        xf = np.ascontiguousarray(xf, np.float32).reshape(-1, 3)
        vf = np.ascontiguousarray(vf if vf is not None else np.zeros_like(xf), np.float32).reshape(-1, 3)
        xb = np.ascontiguousarray(xb, np.float32).reshape(-1, 3)
        if self.nb == 0:
            xb = np.zeros((1, 3), np.float32)
        self.h = lib().sphref_create(bits, ctypes.byref(params), self.nf, _p(xf, ctypes.c_float),
                                     _p(vf, ctypes.c_float), self.nb, _p(xb, ctypes.c_float))
        self.last_step_seconds = 0.0

    def close(self):
        if self.h:
            lib().sphref_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self):
        lib().sphref_init(self.h)

    def step(self, n=1):
        st = np.zeros((n, 6), np.float64)
        self.last_step_seconds = lib().sphref_step(self.h, n, _p(st, ctypes.c_double))
        return st

    def state(self):
        n = self.nf
        x = np.zeros((n, 3)); v = np.zeros((n, 3)); a = np.zeros((n, 3))
        rho = np.zeros(n); p = np.zeros(n); al = np.zeros(n)
        lib().sphref_get(self.h, *(_p(q, ctypes.c_double) for q in (x, v, a, rho, p, al)))
        return dict(x=x, v=v, a=a, rho=rho, p=p, alpha=al)

    def set_state64(self, x=None, v=None):
        """Overwrite positions / velocities (N x 3) at FP64 precision."""
        xx = None if x is None else np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        vv = None if v is None else np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        lib().sphref_set_state64(self.h, None if xx is None else _p(xx, ctypes.c_double),
                                 None if vv is None else _p(vv, ctypes.c_double))

    def set_pressure_branch(self, masks):
        """WCSPH: per step (rows of an nsteps x N bool array) which particles
        take p = B((rho/rho0)^g - 1) instead of max(0, .); None clears."""
        if masks is None:
            lib().sphref_set_pressure_branch(self.h, 0, None)
            return
        m = np.ascontiguousarray(masks, np.uint8).reshape(-1, self.nf)
        lib().sphref_set_pressure_branch(self.h, len(m), _p(m, ctypes.c_uint8))

    def set_cohesion_positions(self, xs):
        """Per step (nsteps x N x 3 float32) positions whose FP32 pair lengths
        decide the cohesion branch at h/2; None clears."""
        if xs is None:
            lib().sphref_set_cohesion_positions(self.h, 0, None)
            return
        q = np.ascontiguousarray(xs, np.float32).reshape(-1, self.nf, 3)
        lib().sphref_set_cohesion_positions(self.h, len(q), _p(q, ctypes.c_float))

    def set_dfsph_branch(self, kD, sD, kV, sV):
        """DFSPH: branches of the next steps from another run, each array
        [nsteps][nit + 1][N] float32 (pass nit = warm start): k and rho* of the
        density solve, k and clamped divergence of the divergence solve."""
        arrs = [np.ascontiguousarray(a, np.float32) for a in (kD, sD, kV, sV)]
        nsteps, nit1 = arrs[0].shape[0], arrs[0].shape[1]
        lib().sphref_set_dfsph_branch(self.h, nsteps, nit1 - 1, *(_p(a, ctypes.c_float) for a in arrs))

    def cohesion_print(self):
        """(hash, count) of the pairs that took the inner cohesion branch so far."""
        out = np.zeros(2, np.uint64)
        lib().sphref_cohesion_print(self.h, _p(out, ctypes.c_uint64))
        return int(out[0]), int(out[1])

    def boundary_volumes(self):
        out = np.zeros(max(self.nb, 1))
        lib().sphref_get_boundary(self.h, _p(out, ctypes.c_double))
        return out[: self.nb]

    def scalars(self):
        out = np.zeros(5)
        lib().sphref_scalars(self.h, _p(out, ctypes.c_double))
        return dict(t=out[0], dt=out[1], mass=out[2], h=out[3], pcisph_delta=out[4])

    def cell_keys(self):
        out = np.zeros((self.nf, 3), np.int32)
        lib().sphref_cell_keys(self.h, _p(out, ctypes.c_int32))
        return out

    def neighbors(self):
        off = np.zeros(self.nf + 1, np.int64)
        boff = np.zeros(self.nf + 1, np.int64)
        cap = 64 * self.nf + 1
        bcap = 64 * self.nf + 1
        while True:
            ids = np.zeros(cap, np.int32)
            bids = np.zeros(bcap, np.int32)
            bt = ctypes.c_int64(0)
            tot = lib().sphref_neighbors(self.h, _p(off, ctypes.c_int64), _p(ids, ctypes.c_int32), cap,
                                         _p(boff, ctypes.c_int64), _p(bids, ctypes.c_int32), bcap, ctypes.byref(bt))
            if tot <= cap and bt.value <= bcap:
                return off, ids[:tot], boff, bids[: bt.value]
            cap, bcap = max(cap, tot), max(bcap, bt.value)


def write_scene(scene, path):
    sc = json.loads(json.dumps(scene))
    for b in sc.get("Boundary", []):
        # the reference reads these arrays unconditionally (SceneLoader.cpp:160-162)
        b.setdefault("box", [])
        b.setdefault("sphere", [])
        b.setdefault("geometry", [])
    with open(path, "w") as f:
        json.dump(sc, f)


def reference_available():
    return os.path.exists(REF_BIN)


def run_reference(scene, nsteps, x=None, v=None, neighbors=False, threads=1, timeout=600, frames=False):
    """Run the compiled reference on `scene` for `nsteps` steps. Returns a dict
    of arrays (see oracle/ref_driver.cpp); with frames=True the steps go
    through SPHSimulation::step() with frame output on, and the dict also
    holds "frames": {file name: parsed JSON} of the frames it wrote."""
    if not reference_available():
        raise RuntimeError("compiled reference not built (make -C oracle ref)")
    with tempfile.TemporaryDirectory() as td:
        sp = os.path.join(td, "scene.json")
        write_scene(scene, sp)
        args = [REF_BIN, sp, str(int(nsteps)), td]
        if x is not None:
            x = np.asarray(x, np.float32).reshape(-1, 3)
            v = np.zeros_like(x) if v is None else np.asarray(v, np.float32).reshape(-1, 3)
            ip = os.path.join(td, "init.bin")
            with open(ip, "wb") as f:
                np.array([len(x)], np.int32).tofile(f)
                x.tofile(f)
                v.tofile(f)
            args += ["--init", ip]
        if neighbors:
            args.append("--nbr")
        if frames:
            args.append("--frames")
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run(args, env=env, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError("reference run failed: %s %s" % (r.stdout[-2000:], r.stderr[-2000:]))
        out = {}
        for fn in os.listdir(td):
            if fn.endswith(".npy"):
                out[fn[:-4]] = np.load(os.path.join(td, fn))
        if frames:
            out["frames"] = {}
            for fn in sorted(os.listdir(td)):
                if fn.startswith("scene_") and fn.endswith(".json"):
                    with open(os.path.join(td, fn)) as f:
                        out["frames"][fn] = json.load(f)
        return out
