"""ctypes binding of libsph2o.so (the C ABI declared in include/sph2o.h).

The library is the only compute path: importing this module fails loudly when
the HIP extension has not been built, and there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsph2o.so")

SPH_OK = 0
SPH_E_ARG = -1
SPH_E_SCENE = -2
SPH_E_UNSUPPORTED = -3
SPH_E_HIP = -4
SPH_E_NONFINITE = -5
SPH_E_OUT_OF_GRID = -6
SPH_E_STATE = -7
SPH_E_CAPACITY = -8

WCSPH, PCISPH, DFSPH = 0, 1, 2

# every function declared in include/sph2o.h
EXPORTS = [
    "sph_create", "sph_destroy", "sph_last_error", "sph_version", "sph_add_scene", "sph_load_scene_json",
    "sph_scene_from_json",
    "sph_num_scenes", "sph_get_params", "sph_num_fluid", "sph_num_boundary", "sph_init", "sph_step",
    "sph_synchronize", "sph_set_fluid_state", "sph_get_fluid_state", "sph_get_boundary", "sph_get_field",
    "sph_get_time", "sph_dump_cells", "sph_dump_neighbors", "sph_last_step_timing", "sph_set_profiling",
    "sph_kernel_count", "sph_kernel_name", "sph_kernel_stats", "sph_reset_kernel_stats",
    "sph_set_coefficients", "sph_forward_record", "sph_adjoint_seed", "sph_adjoint_run", "sph_adjoint_result",
    "sph_recorded_field", "sph_add_emitter", "sph_num_active", "sph_save_state", "sph_load_state",
    "sph_add_cube_boundary", "sph_viewer_feed", "sph_dump_cells_noalloc",
    "sph_set_boundary_normals", "sph_scene_boundary_normals_from_json", "sph_get_boundary_state",
    "sph_set_search_lanes", "sph_get_search_lanes",
]


class SceneParams(ctypes.Structure):
    """sph_scene_params (include/sph2o.h)."""
    _fields_ = [
        ("simulation_method", ctypes.c_int32), ("boundary_method", ctypes.c_int32),
        ("viscosity_method", ctypes.c_int32), ("surface_tension_method", ctypes.c_int32),
        ("adhesion_method", ctypes.c_int32),
        ("min_iterations", ctypes.c_uint32), ("max_iterations", ctypes.c_uint32),
        ("min_iterations_v", ctypes.c_uint32), ("max_iterations_v", ctypes.c_uint32),
        ("adaptive_time_step", ctypes.c_int32),
        ("particle_radius", ctypes.c_float), ("density0", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
        ("time_step", ctypes.c_float), ("min_time_step", ctypes.c_float), ("max_time_step", ctypes.c_float),
        ("fps", ctypes.c_float), ("cfl_factor", ctypes.c_float),
        ("start_time", ctypes.c_float), ("end_time", ctypes.c_float),
        ("stiffness", ctypes.c_float), ("gamma", ctypes.c_float),
        ("eta", ctypes.c_float), ("eta_v", ctypes.c_float),
        ("viscosity", ctypes.c_float), ("boundary_viscosity", ctypes.c_float),
        ("surface_tension", ctypes.c_float), ("adhesion", ctypes.c_float),
        ("reserved", ctypes.c_int32 * 8),
    ]


class EmitterParams(ctypes.Structure):
    """sph_emitter_params (include/sph2o.h)."""
    _fields_ = [
        ("type", ctypes.c_int32), ("reserved0", ctypes.c_int32), ("num_particles", ctypes.c_int64),
        ("position", ctypes.c_float * 3), ("rotation", ctypes.c_float * 4), ("velocity", ctypes.c_float),
        ("start_time", ctypes.c_float), ("width", ctypes.c_float), ("height", ctypes.c_float),
        ("spacing", ctypes.c_float), ("reserved", ctypes.c_int32 * 4),
    ]


class StepStats(ctypes.Structure):
    """sph_step_stats (include/sph2o.h)."""
    _fields_ = [
        ("t", ctypes.c_double), ("dt", ctypes.c_double),
        ("iterations", ctypes.c_int32), ("iterations_v", ctypes.c_int32),
        ("avg_density_error", ctypes.c_double), ("avg_divergence_error", ctypes.c_double),
        ("max_velocity", ctypes.c_double),
        ("capped_density", ctypes.c_int32), ("capped_divergence", ctypes.c_int32),
        ("scene", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


_lib = None


def load():
    """Load libsph2o.so (built in-tree by ``__graft_entry__.build()``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsph2o.so is not built (run __graft_entry__.build() or make -C sph2o_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    fp = P(ctypes.c_float)
    sig = {
        "sph_create": (ctypes.c_int, [P(vp), ctypes.c_int, vp]),
        "sph_destroy": (ctypes.c_int, [vp]),
        "sph_last_error": (ctypes.c_char_p, [vp]),
        "sph_version": (ctypes.c_char_p, []),
        "sph_add_scene": (ctypes.c_int, [vp, P(SceneParams), ctypes.c_int64, fp, fp, ctypes.c_int64, fp]),
        "sph_load_scene_json": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "sph_scene_from_json": (ctypes.c_int, [ctypes.c_char_p, P(SceneParams), P(ctypes.c_int64), fp, ctypes.c_int64,
                                               P(ctypes.c_int64), fp, ctypes.c_int64]),
        "sph_num_scenes": (ctypes.c_int, [vp]),
        "sph_get_params": (ctypes.c_int, [vp, ctypes.c_int, P(SceneParams)]),
        "sph_num_fluid": (ctypes.c_int64, [vp, ctypes.c_int]),
        "sph_num_active": (ctypes.c_int64, [vp, ctypes.c_int]),
        "sph_save_state": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "sph_add_cube_boundary": (ctypes.c_int, [vp, ctypes.c_int, fp, fp, ctypes.c_float, ctypes.c_float]),
        "sph_set_boundary_normals": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int64, fp, ctypes.c_float,
                                                    ctypes.c_float]),
        "sph_scene_boundary_normals_from_json": (ctypes.c_int, [ctypes.c_char_p, P(ctypes.c_int64), fp, ctypes.c_int64,
                                                                fp, fp]),
        "sph_get_boundary_state": (ctypes.c_int, [vp, ctypes.c_int, fp, fp]),
        "sph_load_state": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "sph_add_emitter": (ctypes.c_int, [vp, ctypes.c_int, P(EmitterParams)]),
        "sph_num_boundary": (ctypes.c_int64, [vp, ctypes.c_int]),
        "sph_init": (ctypes.c_int, [vp]),
        "sph_step": (ctypes.c_int, [vp, ctypes.c_int, P(StepStats)]),
        "sph_synchronize": (ctypes.c_int, [vp]),
        "sph_set_fluid_state": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int64, fp, fp]),
        "sph_get_fluid_state": (ctypes.c_int, [vp, ctypes.c_int, fp, fp, fp, fp]),
        "sph_get_boundary": (ctypes.c_int, [vp, ctypes.c_int, fp, fp]),
        "sph_get_field": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, fp]),
        "sph_get_time": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double)]),
        "sph_viewer_feed": (ctypes.c_int, [vp, ctypes.c_int, vp, vp]),
        "sph_dump_cells": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_int32)]),
        "sph_dump_cells_noalloc": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_int32)]),
        "sph_dump_neighbors": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_int64), P(ctypes.c_int32), ctypes.c_int64,
                                              P(ctypes.c_int64), P(ctypes.c_int32), ctypes.c_int64]),
        "sph_last_step_timing": (ctypes.c_int, [vp, P(ctypes.c_double), P(ctypes.c_int64)]),
        "sph_set_profiling": (ctypes.c_int, [vp, ctypes.c_int]),
        "sph_set_search_lanes": (ctypes.c_int, [vp, ctypes.c_int]),
        "sph_get_search_lanes": (ctypes.c_int, [vp, P(ctypes.c_int), P(ctypes.c_int), fp]),
        "sph_kernel_count": (ctypes.c_int, []),
        "sph_kernel_name": (ctypes.c_char_p, [ctypes.c_int]),
        "sph_kernel_stats": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_int64)]),
        "sph_reset_kernel_stats": (ctypes.c_int, [vp]),
        "sph_set_coefficients": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_float, ctypes.c_float]),
        "sph_forward_record": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, P(StepStats)]),
        "sph_adjoint_seed": (ctypes.c_int, [vp, ctypes.c_int, fp, fp]),
        "sph_adjoint_run": (ctypes.c_int, [vp]),
        "sph_adjoint_result": (ctypes.c_int, [vp, ctypes.c_int, fp, fp, P(ctypes.c_double), P(ctypes.c_double)]),
        "sph_recorded_field": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class SPHError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def scene_from_json(path):
    """Parse a scene JSON and sample its particles without a GPU.
    Returns (SceneParams, fluid positions (N,3), boundary positions (Nb,3))."""
    import numpy as np
    L = load()
    p = SceneParams()
    nf = ctypes.c_int64(0)
    nb = ctypes.c_int64(0)
    fpp = ctypes.POINTER(ctypes.c_float)
    rc = L.sph_scene_from_json(str(path).encode(), ctypes.byref(p), ctypes.byref(nf), None, 0, ctypes.byref(nb),
                               None, 0)
    if rc not in (SPH_OK, SPH_E_CAPACITY):
        raise SPHError(rc, "sph_scene_from_json failed for %s" % path)
    x = np.zeros((max(1, nf.value), 3), np.float32)
    xb = np.zeros((max(1, nb.value), 3), np.float32)
    rc = L.sph_scene_from_json(str(path).encode(), ctypes.byref(p), ctypes.byref(nf), x.ctypes.data_as(fpp),
                               len(x), ctypes.byref(nb), xb.ctypes.data_as(fpp), len(xb))
    if rc != SPH_OK:
        raise SPHError(rc, "sph_scene_from_json failed for %s" % path)
    return p, x[: nf.value], xb[: nb.value]


def boundary_normals_from_json(path):
    """The PCISPH boundary's normals (Nb, 3) and (normalFct, tangFct) of a
    scene JSON of boundary method 1, without a GPU."""
    import numpy as np
    L = load()
    nb = ctypes.c_int64(0)
    nf, tf = ctypes.c_float(0), ctypes.c_float(0)
    fpp = ctypes.POINTER(ctypes.c_float)
    rc = L.sph_scene_boundary_normals_from_json(str(path).encode(), ctypes.byref(nb), None, 0, ctypes.byref(nf),
                                                ctypes.byref(tf))
    if rc not in (SPH_OK, SPH_E_CAPACITY):
        raise SPHError(rc, "sph_scene_boundary_normals_from_json failed for %s" % path)
    out = np.zeros((max(1, nb.value), 3), np.float32)
    rc = L.sph_scene_boundary_normals_from_json(str(path).encode(), ctypes.byref(nb), out.ctypes.data_as(fpp), len(out),
                                                ctypes.byref(nf), ctypes.byref(tf))
    if rc != SPH_OK:
        raise SPHError(rc, "sph_scene_boundary_normals_from_json failed for %s" % path)
    return out[: nb.value], (nf.value, tf.value)
