"""Host-side mirror of the reference's simulation interface over the C ABI.

:class:`SPHSimulation` keeps the method names and meaning of the reference's
``SPHSimulation`` (inc/SPH2O/SPHSimulation.h:208-365, SPHSimulation.cpp):
``importScene`` / ``init`` / ``step`` / ``run`` and the getters, with the
same time-window rule in ``step()`` (SPHSimulation.cpp:157-191). Each call
goes through ``libsph2o.so``; there is no CPU fallback. Unlike the reference
(a process-wide singleton with static kernel coefficients) several instances
may coexist, and one instance may hold many scenes (batched stepping).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import SPHError, SceneParams, StepStats


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None


class Context:
    """One sph_ctx: scenes on one GPU and one HIP stream."""

    def __init__(self, device=0, stream=None):
        self.L = _lib.load()
        h = ctypes.c_void_p()
        rc = self.L.sph_create(ctypes.byref(h), int(device), ctypes.c_void_p(stream) if stream else None)
        if rc != 0:
            raise SPHError(rc, "sph_create failed (no HIP device?)")
        self.h = h
        self.initialized = False

    def _check(self, rc):
        if rc < 0:
            raise SPHError(rc, self.L.sph_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.sph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- scenes
    def load_scene_json(self, path):
        return self._check(self.L.sph_load_scene_json(self.h, str(path).encode()))

    def add_scene(self, params, x, v=None, xb=None):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        v = None if v is None else np.ascontiguousarray(v, np.float32).reshape(-1, 3)
        xb = np.zeros((0, 3), np.float32) if xb is None else np.ascontiguousarray(xb, np.float32).reshape(-1, 3)
        return self._check(self.L.sph_add_scene(self.h, ctypes.byref(params), len(x), _fp(x), _fp(v), len(xb),
                                                _fp(xb) if len(xb) else None))

    def num_scenes(self):
        return self.L.sph_num_scenes(self.h)

    def params(self, scene=0):
        p = SceneParams()
        self._check(self.L.sph_get_params(self.h, scene, ctypes.byref(p)))
        return p

    def num_fluid(self, scene=0):
        return self._check(self.L.sph_num_fluid(self.h, scene))

    def save_state(self, path):
        """Whole-context state file for an exact resume (sph_save_state)."""
        self._check(self.L.sph_save_state(self.h, str(path).encode()))

    def load_state(self, path):
        """Resume from a state file written by save_state on a context built from
        the same scenes (after init)."""
        self._check(self.L.sph_load_state(self.h, str(path).encode()))

    def num_active(self, scene=0):
        """Active fluid particles (the first num_active of the num_fluid)."""
        return self._check(self.L.sph_num_active(self.h, scene))

    def add_emitter(self, scene, type, num_particles, position, rotation=(1.0, 0.0, 0.0, 0.0), velocity=1.0,
                    start_time=0.0, width=0.1, height=0.1, spacing=1.0):
        """FluidObject::addEmitter before init: type 0 square / 1 circle,
        rotation the quaternion (w, x, y, z)."""
        e = _lib.EmitterParams()
        e.type = int(type)
        e.num_particles = int(num_particles)
        e.position[:] = [float(c) for c in position]
        e.rotation[:] = [float(c) for c in rotation]
        e.velocity, e.start_time, e.width, e.height, e.spacing = velocity, start_time, width, height, spacing
        return self._check(self.L.sph_add_emitter(self.h, scene, ctypes.byref(e)))

    def num_boundary(self, scene=0):
        return self._check(self.L.sph_num_boundary(self.h, scene))

    # -- lifecycle
    def init(self):
        self._check(self.L.sph_init(self.h))
        self.initialized = True

    def step(self, nsteps=1):
        ns = self.num_scenes()
        st = (StepStats * (max(1, nsteps) * ns))()
        self._check(self.L.sph_step(self.h, int(nsteps), st))
        return [[st[k * ns + s] for s in range(ns)] for k in range(nsteps)]

    def synchronize(self):
        self._check(self.L.sph_synchronize(self.h))

    # -- state
    def set_fluid_state(self, scene, x=None, v=None):
        n = self.num_fluid(scene)
        x = None if x is None else np.ascontiguousarray(x, np.float32).reshape(n, 3)
        v = None if v is None else np.ascontiguousarray(v, np.float32).reshape(n, 3)
        self._check(self.L.sph_set_fluid_state(self.h, scene, n, _fp(x), _fp(v)))

    def fluid_state(self, scene=0):
        n = self.num_fluid(scene)
        x = np.zeros((n, 3), np.float32)
        v = np.zeros((n, 3), np.float32)
        rho = np.zeros(n, np.float32)
        p = np.zeros(n, np.float32)
        self._check(self.L.sph_get_fluid_state(self.h, scene, _fp(x), _fp(v), _fp(rho), _fp(p)))
        return dict(x=x, v=v, rho=rho, p=p)

    def boundary(self, scene=0, volumes=True):
        nb = self.num_boundary(scene)
        xb = np.zeros((nb, 3), np.float32)
        vol = np.zeros(nb, np.float32) if volumes else None
        self._check(self.L.sph_get_boundary(self.h, scene, _fp(xb) if nb else None, _fp(vol) if (volumes and nb) else None))
        return xb, vol

    def set_boundary_normals(self, scene, normals, normal_fct, tangential_fct):
        """The PCISPH boundary object of a scene added by add_scene (boundary
        method 1): one normal per boundary particle and the object's factors."""
        nrm = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        self._check(self.L.sph_set_boundary_normals(self.h, scene, len(nrm), _fp(nrm), float(normal_fct),
                                                    float(tangential_fct)))

    def boundary_state(self, scene=0):
        """Densities and pressures of a PCISPH boundary object's particles."""
        nb = self.num_boundary(scene)
        rho = np.zeros(max(1, nb), np.float32)
        p = np.zeros(max(1, nb), np.float32)
        self._check(self.L.sph_get_boundary_state(self.h, scene, _fp(rho), _fp(p)))
        return rho[:nb], p[:nb]

    FIELDS = {"alpha": (0, 1), "a": (1, 3), "aux": (2, 1), "k_ws": (3, 1), "kv_ws": (4, 1), "normal": (5, 3)}

    def field(self, name, scene=0):
        fid, comp = self.FIELDS[name]
        n = self.num_fluid(scene)
        out = np.zeros((n, comp) if comp > 1 else n, np.float32)
        self._check(self.L.sph_get_field(self.h, scene, fid, _fp(out)))
        return out

    def viewer_feed_into(self, scene, pos4_ptr, speed_ptr):
        """sph_viewer_feed into caller-owned device buffers (integer device
        addresses, 0 to skip one): vec4 positions (w = 1) and |v| of the active
        particles in original order, ordered on the context's stream."""
        self._check(self.L.sph_viewer_feed(self.h, scene, ctypes.c_void_p(pos4_ptr or None),
                                           ctypes.c_void_p(speed_ptr or None)))

    def viewer_feed(self, scene=0):
        """The viewer's per-frame data as host arrays: (n, 4) positions with
        w = 1 and (n,) |v| (SPHSimulation::updateScalarField) of the active
        particles in original order, written on the GPU into device buffers
        and copied back."""
        from ._hip import DeviceBuffer
        n = self.num_active(scene)
        pos4 = np.zeros((n, 4), np.float32)
        speed = np.zeros(n, np.float32)
        with DeviceBuffer(max(1, pos4.nbytes)) as dp, DeviceBuffer(max(1, speed.nbytes)) as ds:
            self.viewer_feed_into(scene, dp.ptr, ds.ptr)
            self.synchronize()
            dp.copy_to(pos4)
            ds.copy_to(speed)
        return pos4, speed

    def time(self, scene=0):
        t = ctypes.c_double()
        dt = ctypes.c_double()
        self._check(self.L.sph_get_time(self.h, scene, ctypes.byref(t), ctypes.byref(dt)))
        return t.value, dt.value

    def cells(self, scene=0):
        n = self.num_fluid(scene)
        out = np.zeros((n, 3), np.int32)
        self._check(self.L.sph_dump_cells(self.h, scene, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    def cells_noalloc(self, scene=0):
        """The same keys through sph_dump_cells_noalloc (no device allocation)."""
        n = self.num_fluid(scene)
        out = np.zeros((n, 3), np.int32)
        self._check(self.L.sph_dump_cells_noalloc(self.h, scene, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    def neighbors(self, scene=0):
        """Within-support sets (CSR, original ids, unsorted rows)."""
        n = self.num_fluid(scene)
        off = np.zeros(n + 1, np.int64)
        boff = np.zeros(n + 1, np.int64)
        cap, bcap = 48 * n + 1, 32 * n + 1
        P64 = ctypes.POINTER(ctypes.c_int64)
        P32 = ctypes.POINTER(ctypes.c_int32)
        while True:
            ids = np.zeros(cap, np.int32)
            bids = np.zeros(bcap, np.int32)
            rc = self.L.sph_dump_neighbors(self.h, scene, off.ctypes.data_as(P64), ids.ctypes.data_as(P32), cap,
                                           boff.ctypes.data_as(P64), bids.ctypes.data_as(P32), bcap)
            if rc == _lib.SPH_E_CAPACITY:
                cap, bcap = int(off[-1]) + 1, int(boff[-1]) + 1
                continue
            self._check(rc)
            return off, ids[: off[-1]], boff, bids[: boff[-1]]

    # -- gradients of a WCSPH rollout
    def set_coefficients(self, scene, viscosity, surface_tension):
        self._check(self.L.sph_set_coefficients(self.h, scene, float(viscosity), float(surface_tension)))

    def forward_record(self, nsteps, checkpoint_every):
        """Advance nsteps like step() and record the rollout for adjoint()."""
        ns = self.num_scenes()
        st = (StepStats * (max(1, nsteps) * ns))()
        self._check(self.L.sph_forward_record(self.h, int(nsteps), int(checkpoint_every), st))
        return [[st[k * ns + s] for s in range(ns)] for k in range(nsteps)]

    def adjoint(self, seeds):
        """seeds: {scene: (dL/dx_T, dL/dv_T)} (N x 3 arrays or None). Returns
        {scene: dict(x0=dL/dx_0, v0=dL/dv_0, viscosity=..., surface_tension=...)}."""
        for sc, (gx, gv) in seeds.items():
            n = self.num_fluid(sc)
            gx = None if gx is None else np.ascontiguousarray(gx, np.float32).reshape(n, 3)
            gv = None if gv is None else np.ascontiguousarray(gv, np.float32).reshape(n, 3)
            self._check(self.L.sph_adjoint_seed(self.h, sc, _fp(gx), _fp(gv)))
        self._check(self.L.sph_adjoint_run(self.h))
        out = {}
        for sc in seeds:
            n = self.num_fluid(sc)
            gx0 = np.zeros((n, 3), np.float32)
            gv0 = np.zeros((n, 3), np.float32)
            gnu = ctypes.c_double()
            gst = ctypes.c_double()
            self._check(self.L.sph_adjoint_result(self.h, sc, _fp(gx0), _fp(gv0), ctypes.byref(gnu),
                                                  ctypes.byref(gst)))
            out[sc] = dict(x0=gx0, v0=gv0, viscosity=gnu.value, surface_tension=gst.value)
        return out

    RECORDED = {"k_density": 0, "rho_star": 1, "k_divergence": 2, "divergence": 3, "x": 4, "pressure": 5}

    def recorded(self, name, step, scene=0, pass_=0):
        """A field of the recorded rollout at `step` (see sph_recorded_field)."""
        fid = self.RECORDED[name]
        n = self.num_fluid(scene)
        out = np.zeros((n, 3) if fid == 4 else n, np.float32)
        self._check(self.L.sph_recorded_field(self.h, scene, int(step), fid, int(pass_), _fp(out)))
        return out

    # -- timing
    def last_step_ms(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self.L.sph_last_step_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def set_profiling(self, on):
        """True/False for every kernel, or an int bit mask of kernel ids."""
        mask = int(on) if isinstance(on, int) and not isinstance(on, bool) else (1 if on else 0)
        self._check(self.L.sph_set_profiling(self.h, mask))

    def kernel_stats(self):
        out = {}
        for k in range(self.L.sph_kernel_count()):
            ms = ctypes.c_double()
            n = ctypes.c_int64()
            self._check(self.L.sph_kernel_stats(self.h, k, ctypes.byref(ms), ctypes.byref(n)))
            if n.value:
                out[self.L.sph_kernel_name(k).decode()] = (ms.value, n.value)
        return out

    def reset_kernel_stats(self):
        self._check(self.L.sph_reset_kernel_stats(self.h))

    def set_search_lanes(self, lanes):
        """Lanes per particle of the search's list build: 1, 2, 4 or 8 pins
        it, 0 times each on the first searches and keeps the fastest (every
        count builds the same lists)."""
        self._check(self.L.sph_set_search_lanes(self.h, int(lanes)))

    def search_lanes(self):
        """(lanes in use, tuned, {lanes: fastest build ms}) of the search."""
        lanes = ctypes.c_int()
        tuned = ctypes.c_int()
        ms = (ctypes.c_float * 4)()
        self._check(self.L.sph_get_search_lanes(self.h, ctypes.byref(lanes), ctypes.byref(tuned), ms))
        return lanes.value, bool(tuned.value), {c: float(m) for c, m in zip((1, 2, 4, 8), ms)}


class SPHSimulation:
    """Drop-in for the reference's ``SPHSimulation`` on one scene
    (SPHSimulation.cpp:117-199): ``importScene(path)``, ``init()``,
    ``step()`` (advances only while startTime <= t <= endTime and returns
    whether a frame is due, TimeHelper.h:100-110), ``run()`` (which, unlike
    SPHSimulation.cpp:193-199, re-reads the time each iteration so it ends)."""

    def __init__(self, device=0):
        self.ctx = Context(device)
        self.scene = None
        self.name = ""
        self.frame = 1
        self.last_save = np.float32(0.0)
        self.active_save = False

    def importScene(self, path):
        try:
            self.scene = self.ctx.load_scene_json(path)
        except SPHError:
            return False
        self.name = str(path)
        return True

    def activateSave(self, active):
        """Write a frame file at every frame step (SPHSimulation.h:131)."""
        self.active_save = bool(active)

    def getName(self):
        return self.name

    def writeFluid(self):
        """SceneLoader::writeFluid (SceneLoader.cpp:211-275): <scene>_NNNN.json
        (frame number zero-padded to 4 digits, SceneLoader.h:20) with
        numParticles, numActiveParticles, time, timeStep and every particle's
        position and velocity."""
        import json
        num = "%04d" % self.frame
        path = self.name[:-len(".json")] + "_" + num + ".json" if self.name.endswith(".json") else \
            self.name + "_" + num + ".json"
        st = self.ctx.fluid_state(self.scene)
        t, dt = self.ctx.time(self.scene)
        state = {"numParticles": int(self.ctx.num_fluid(self.scene)),
                 "numActiveParticles": int(self.ctx.num_active(self.scene)),
                 "time": float(np.float32(t)), "timeStep": float(np.float32(dt)),
                 "positions": st["x"].astype(np.float64).tolist(),
                 "velocities": st["v"].astype(np.float64).tolist()}
        with open(path, "w") as f:
            json.dump(state, f)
        return path

    def init(self):
        self.ctx.init()
        self.frame = 1
        self.last_save = np.float32(0.0)

    def getTime(self):
        return self.ctx.time(self.scene)[0]

    def getTimeStep(self):
        return self.ctx.time(self.scene)[1]

    def getStartTime(self):
        return self.ctx.params(self.scene).start_time

    def getEndTime(self):
        return self.ctx.params(self.scene).end_time

    def getParticleRadius(self):
        return self.ctx.params(self.scene).particle_radius

    def getSupportRadius(self):
        return np.float32(4.0) * np.float32(self.getParticleRadius())

    def getFrame(self):
        return self.frame

    def step(self):
        t = np.float32(self.getTime())
        p = self.ctx.params(self.scene)
        if np.float32(p.start_time) <= t <= np.float32(p.end_time):
            self.last_stats = self.ctx.step(1)[0][self.scene]
            t = np.float32(self.getTime())
            # TimeHelper::hasToSave: time - lastTimeSave in FP32 against 1.0 / fps
            # in FP64; lastTimeSave = (Real) frame / fps in FP32
            if float(np.float32(t - self.last_save)) >= 1.0 / float(np.float32(p.fps)):
                self.last_save = np.float32(np.float32(self.frame) / np.float32(p.fps))
                if self.active_save:
                    self.writeFluid()
                self.frame += 1
                return True
        return False

    def run(self, max_steps=None):
        n = 0
        p = self.ctx.params(self.scene)
        while np.float32(p.start_time) <= np.float32(self.getTime()) <= np.float32(p.end_time):
            self.step()
            n += 1
            if max_steps is not None and n >= max_steps:
                break
        return n

    def getPositions(self):
        """Positions as vec4 with w = 1 (SimulationObject::getPositions, GUI instancing)."""
        x = self.ctx.fluid_state(self.scene)["x"]
        return np.concatenate([x, np.ones((len(x), 1), np.float32)], axis=1)

    def getVelocities(self):
        return self.ctx.fluid_state(self.scene)["v"]

    def getDensities(self):
        return self.ctx.fluid_state(self.scene)["rho"]
