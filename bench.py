"""Benchmark: DFSPH particle-steps/s on the 1M-particle dam break (one MI355X).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload (the largest configuration that fits one GPU and the one the metric
is quoted on): DFSPH dam break, 100^3 = 1,000,000 FP32 fluid particles
(r = 0.005, h = 0.02, lattice in [0,1]^3) plus 400,002 Akinci boundary
particles sampled on a 4 x 2 x 2 m box; CFL-adaptive dt (max 1e-3); density
tolerance 0.01%, divergence tolerance 0.1%, 2..100 iterations; Akinci
surface tension gamma = 0.1. A step is one DFSPH step of every fluid
particle; inputs are resident in HBM when the timed region starts.

Prints ONE JSON line with value = fluid particle-steps per second over K
timed steps (after W untimed warmup steps), the roofline of the dominant
kernel (its algorithmic bytes per launch over its average launch time,
measured with HIP events on the context's stream inside the timed region)
and a CPU baseline: the original SPH2O code compiled with OpenMP
(oracle/_ref/sph2o_ref) timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DFSPH particle-steps/s at 1M (ρ err≤0.1%); batched gain; adjoint/fwd cost"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per fluid particle per launch (streaming model: each
# pass reads/writes each particle's own fields once, x counted as 12 B, FP32;
# neighbour re-reads hit LDS/L2; boundary records (x 12 B + volume 4 B) read
# once per boundary-reading pass, i.e. + 16 Nb/Nf bytes per particle).
# name -> (own-field bytes, reads boundary)
ALGO_BYTES = {
    "rho_star": (48, True),     # x, v, rho, alpha, k_ws in; rho*, k, k_ws out
    "velocity": (40, True),     # x, k, v in; v out
    "divergence": (48, True),   # x, v, rho, alpha, kv_ws in; d, kv, kv_ws out
    "density": (20, True),      # x in; rho, alpha out
    "forces": (28, False),      # x, rho in; normal out (surface-tension pass 1)
    "surface_tension": (52, False),  # x, normal, rho, a in; a out
    "build_neighbors": (28, True),   # x, cell key in; ranges (model cost of the grid build)
    "predict_velocity": (36, False),
    "keys": (36, False),
    "reorder": (56, False),
    "vmax": (24, False),
    "scale": (8, False),
}


def calibration_copy(nbytes):
    """One device-to-device copy of `nbytes` by the HIP runtime's copy kernel:
    the reference point that scales rocprofv3's FETCH_SIZE / WRITE_SIZE
    counters to bytes (it reads and writes exactly `nbytes`)."""
    import ctypes

    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(ctypes.byref(ptr), nbytes) != 0:
            raise RuntimeError("hipMalloc(%d) failed" % nbytes)
    try:
        if hip.hipMemset(src, 1, nbytes) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(dst, src, nbytes, 3) != 0 or hip.hipDeviceSynchronize() != 0:  # device to device
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    log("calibration copy: %d bytes read, %d bytes written" % (nbytes, nbytes))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/d1m_hbm_traffic.md,
    lines `traffic <kernel> <bytes_per_launch>`), or None."""
    path = os.path.join(ROOT, "profiles", "d1m_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.strip().strip("`").split()
            if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel:
                return float(parts[2])
    return None


# VALU issue peak of MI355X: 256 CUs, each issuing one wave64 VALU instruction
# per cycle (4 SIMDs, 4 cycles per wave64 instruction), at 2.4 GHz
VALU_PEAK_WAVE_INSTS = 256 * 2.4e9


def measured_valu(kernel):
    """VALU wave-instructions per launch of `kernel` on this workload, from the
    committed rocprofv3 SQ_INSTS_VALU summary
    (profiles/d1m_neighbour_pass_counters.md, lines `valu <kernel> <per_launch>`),
    or None."""
    path = os.path.join(ROOT, "profiles", "d1m_neighbour_pass_counters.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.strip().strip("`").split()
            if len(parts) == 3 and parts[0] == "valu" and parts[1] == kernel:
                return float(parts[2])
    return None


def measured_ta_busy(kernel):
    """Busy fraction of the texture-address unit (one per CU, every vector
    memory instruction passes it) over `kernel`'s own dispatches in the
    committed rocprofv3 TA_BUSY_avr run of this workload
    (profiles/d1m_vmem_pipeline_counters.md, lines `ta_busy_frac <kernel> <f>`),
    or None."""
    path = os.path.join(ROOT, "profiles", "d1m_vmem_pipeline_counters.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.strip().strip("`").split()
            if len(parts) == 3 and parts[0] == "ta_busy_frac" and parts[1] == kernel:
                return float(parts[2])
    return None


def search_info(ctx):
    """The search's lanes per particle in a context and the fastest list build
    its tuning rounds saw per count (ms; None for a count not timed)."""
    lanes, tuned, ms = ctx.search_lanes()
    return {"lanes": lanes, "tuned": tuned,
            "build_ms": {str(k): (float(v) if np.isfinite(v) else None) for k, v in ms.items()}}


def pair_throughput(ctx, sid, its, steps_per_s):
    """True neighbour pairs evaluated per second (every pass's sum over i of
    |S_i|, fluid and boundary, at the mean list lengths of the final state)
    and 27-cell candidate tests per second of the search. DFSPH passes per
    step: normals and surface tension (fluid), rho* and velocity per density
    iteration and warm start, density + alpha, divergence and velocity per
    divergence iteration and warm start (fluid + boundary); the search tests
    every fluid and boundary candidate of the 27 cells once."""
    off, _, boff, _ = ctx.neighbors(sid)
    nf_mean = float(np.diff(off).mean())
    nb_mean = float(np.diff(boff).mean())
    cells = ctx.cells(sid)
    xb, _ = ctx.boundary(sid, volumes=False)
    h = 4 * ctx.params(sid).particle_radius
    bcells = np.floor(xb * np.float32(1.0 / h)).astype(np.int64)
    pack = lambda c: (c[:, 0] + 2 ** 20) * 2 ** 42 + (c[:, 1] + 2 ** 20) * 2 ** 21 + (c[:, 2] + 2 ** 20)
    uf, cf = np.unique(pack(cells.astype(np.int64)), return_counts=True)
    ub, cb = np.unique(pack(bcells), return_counts=True)
    cand = 0
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                q = pack(cells.astype(np.int64) + np.array([dx, dy, dz]))
                for u, c in ((uf, cf), (ub, cb)):
                    k = np.minimum(np.searchsorted(u, q), len(u) - 1)
                    cand += int(np.where(u[k] == q, c[k], 0).sum())
    i_d, i_v = float(its[:, 0].mean()), float(its[:, 1].mean())
    n = len(cells)
    fluid_passes = 2 + 2 * (1 + i_d) + 1 + 2 * (1 + i_v)
    both_passes = 2 * (1 + i_d) + 1 + 2 * (1 + i_v)
    per_step = n * (fluid_passes * nf_mean + both_passes * nb_mean)
    return {"mean_fluid_neighbours": nf_mean, "mean_boundary_neighbours": nb_mean,
            "pairs_per_step": per_step, "pairs_per_s": per_step * steps_per_s,
            "candidate_tests_per_step": cand, "candidate_tests_per_s": cand * steps_per_s}


def host_info():
    """CPU model, logical CPUs and the compiler that built the CPU baselines."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cc = subprocess.run(["g++", "--version"], capture_output=True, text=True, timeout=10).stdout.splitlines()[0]
    except (OSError, IndexError, subprocess.SubprocessError):
        cc = "unknown"
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "compiler": cc}


def cpu_baseline(scene_path, nf, steps, threads):
    """Time the compiled reference (the original sources, built here by
    oracle/Makefile with the shipped -O3 -ffast-math flags and OpenMP) on
    steps 1..`steps` of the workload, step() calls only."""
    from oracle import sphref as S

    if not S.reference_available():
        return None
    with tempfile.TemporaryDirectory() as td:
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        r = subprocess.run([S.REF_BIN, scene_path, str(steps), td], env=env, capture_output=True, text=True,
                           timeout=900)
        if r.returncode != 0:
            log("cpu baseline failed: " + r.stderr[-500:])
            return None
        meta = np.load(os.path.join(td, "meta.npy"))
    secs = float(meta[3])
    return {"value": nf * steps / secs, "unit": "particle-steps/s", "cores": threads, "kind": "reference",
            "sample": "original SPH2O DFSPH step (%s, -O3 -ffast-math) on the same 1M-particle scene, "
                      "steps 1-%d, step() calls only (%.1f s)"
                      % ("OpenMP, %d threads" % threads if threads > 1 else "1 thread", steps, secs)}


def cpu_port(scene_path, nf, steps, threads):
    """The oracle's FP32 restatement of the reference step (oracle/sphref.hpp:
    the same hash grid and pass sequence, -O2, OpenMP) on the same steps."""
    from oracle import sphref as S
    from sph2o_amd import _lib

    os.environ["OMP_NUM_THREADS"] = str(threads)
    p, x, xb = _lib.scene_from_json(scene_path)
    with open(scene_path) as f:
        sc = json.load(f)
    o = S.Oracle(S.params_from_scene(sc), x, None, xb, bits=32)
    o.init()
    t0 = time.perf_counter()
    o.step(steps)
    secs = time.perf_counter() - t0
    o.close()
    return {"value": nf * steps / secs, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "FP32 restatement (oracle/sphref.hpp) of the DFSPH step on the same 1M-particle scene, "
                      "steps 1-%d (%.1f s)" % (steps, secs)}


def adjoint_cost(steps, every, method, warmup=5):
    """A100k: gradient of L = mean |x_T - y|^2 w.r.t. x0, v0, viscosity and
    surface tension through `steps` steps of 100,000 particles (XSPH 0.01,
    Akinci surface tension 0.1). WCSPH (k = 1000, gamma = 7) at dt = 2e-4: the
    largest of 5e-4 / 2.5e-4 / 2e-4 that stays stable at r = 0.005. DFSPH with
    8 density and 8 divergence iterations at a fixed dt = 5e-4. Times the plain
    forward, the recording forward (checkpoint every `every` steps) and the
    reverse sweep (segment recompute + adjoint passes) from the same state.
    The gradient is computed twice: the first call allocates the tapes (two
    segments' worth) and is reported as `first_call`; the second, like the
    forward after its warmup steps, is the measurement, and must give the
    same gradient bit for bit."""
    from sph2o_amd import scenes as SC
    from sph2o_amd.simulation import Context

    dt = 2e-4 if method == 0 else 5e-4
    sc, _, _ = SC.a100k(method, dt=dt)
    td = tempfile.mkdtemp()
    path = os.path.join(td, "a100k.json")
    SC.write_json(sc, path)
    ctx = Context(0)
    sid = ctx.load_scene_json(path)
    ctx.init()
    x0 = ctx.fluid_state(sid)["x"]
    v0 = np.zeros_like(x0)
    n = len(x0)
    ctx.step(warmup)
    ctx.set_fluid_state(sid, x0, v0)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.synchronize()
    t_fwd = time.perf_counter() - t0
    y = SC.adjoint_target(x0)
    from sph2o_amd._lib import SPH_E_NONFINITE, SPHError

    def gradient():
        ctx.set_fluid_state(sid, x0, v0)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.forward_record(steps, every)
        ctx.synchronize()
        t_rec = time.perf_counter() - t0
        xT = ctx.fluid_state(sid)["x"].astype(np.float64)
        seed = (2.0 * (xT - y) / n).astype(np.float32)
        t0 = time.perf_counter()
        try:
            g = ctx.adjoint({sid: (seed, None)})[sid]
        except SPHError as e:   # the sweep completed; its gradient is not finite
            if e.code != SPH_E_NONFINITE:
                raise
            g = None
        return t_rec, time.perf_counter() - t0, xT, g

    rec1, adj1, _, g1 = gradient()
    t_rec, t_adj, xT, g = gradient()
    ctx.close()
    same = (g is None) == (g1 is None) and (g is None or all(
        np.array_equal(g[k], g1[k]) for k in ("x0", "v0", "viscosity", "surface_tension")))
    name = "WCSPH" if method == 0 else "DFSPH 8+8 iterations"
    finite = lambda q: float(q) if q is not None and np.isfinite(q) else None   # strict JSON has no inf / NaN
    gnorm = float(np.linalg.norm(g["v0"].astype(np.float64))) if g is not None else None
    return {"workload": "A100k: %s, %d fluid particles, %d steps at dt=%g, checkpoint every %d"
                        % (name, n, steps, dt, every),
            "forward_s": t_fwd, "record_s": t_rec, "adjoint_s": t_adj, "record_over_forward": t_rec / t_fwd,
            "adjoint_over_forward": t_adj / t_fwd,
            # the whole gradient: the recording forward plus the reverse sweep
            "gradient_over_forward": (t_rec + t_adj) / t_fwd,
            "first_call": {"record_s": rec1, "adjoint_s": adj1, "gradient_over_forward": (rec1 + adj1) / t_fwd},
            "repeat_bit_identical": bool(same),
            "loss": float(np.mean(np.sum((xT - y) ** 2, axis=1))),
            "gradient_finite": g is not None and bool(np.isfinite(gnorm)),
            "dL_dviscosity": finite(g["viscosity"] if g else None),
            "dL_dsurface_tension": finite(g["surface_tension"] if g else None),
            "dL_dv0_norm": finite(gnorm)}


def ref_leg(sc, steps, threads, x=None, v=None, what="step"):
    """The compiled reference on `steps` steps of a scene (step() calls only),
    optionally from given positions / velocities, or None where it is absent."""
    from oracle import sphref as S

    if not S.reference_available():
        return None
    r = S.run_reference(sc, steps, x=x, v=v, threads=threads, timeout=900)
    n = int(r["meta"][0])
    secs = float(r["meta"][3])
    return {"value": n * steps / secs, "unit": "particle-steps/s", "cores": threads, "kind": "reference",
            "sample": "original SPH2O %s (%s, -O3 -ffast-math), steps 1-%d, step() calls only (%.1f s)"
                      % (what, "OpenMP, %d threads" % threads if threads > 1 else "1 thread", steps, secs)}


def port_leg(sc, steps, threads, x, v=None, xb=None):
    """The oracle's FP32 restatement on the same steps."""
    from oracle import sphref as S

    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = S.Oracle(S.params_from_scene(sc), x, v, xb, bits=32)
    o.init()
    t0 = time.perf_counter()
    o.step(steps)
    secs = time.perf_counter() - t0
    o.close()
    return {"value": len(x) * steps / secs, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "FP32 restatement (oracle/sphref.hpp), steps 1-%d (%.1f s)" % (steps, secs)}


def gpu_leg(sc, x, v, windows):
    """One context on `sc` from positions x (velocities v); times each
    (first, last) step window in turn (the windows follow each other)."""
    from sph2o_amd import _lib, scenes as SC
    from sph2o_amd.simulation import Context

    td = tempfile.mkdtemp()
    path = os.path.join(td, "s.json")
    SC.write_json(sc, path)
    p, xs, xb = _lib.scene_from_json(path)
    ctx = Context(0)
    sid = ctx.add_scene(p, xs if x is None else x, v, xb)
    ctx.init()
    n = ctx.num_fluid(sid)
    out = []
    done = 0
    for a, b in windows:
        if a - 1 > done:
            ctx.step(a - 1 - done)
        ctx.synchronize()
        t0 = time.perf_counter()
        st = ctx.step(b - a + 1)
        ctx.synchronize()
        secs = time.perf_counter() - t0
        done = b
        its = float(np.mean([row[sid].iterations for row in st]))
        out.append({"steps": "%d-%d" % (a, b), "value": n * (b - a + 1) / secs, "unit": "particle-steps/s",
                    "ms_per_step": 1000.0 * secs / (b - a + 1), "iterations_per_step": its,
                    "search_lanes": search_info(ctx)})
    ctx.close()
    return n, out


def d1m_flush_leg(threads, steps=20, ref_steps=3):
    """D1M with the fluid block flush against the walls (no gap): the per-step
    statistics of the first `ref_steps` steps on the GPU and in the compiled
    reference, and the GPU's time and mean density error over `steps`."""
    from oracle import sphref as S
    from sph2o_amd import scenes as SC
    from sph2o_amd.simulation import Context

    sc, _, _ = SC.d1m(gap=0.0)
    td = tempfile.mkdtemp()
    path = os.path.join(td, "d1m_flush.json")
    SC.write_json(sc, path)
    from sph2o_amd._lib import SPH_E_NONFINITE, SPHError

    ctx = Context(0)
    sid = ctx.load_scene_json(path)
    ctx.init()
    n = ctx.num_fluid(sid)
    ctx.synchronize()
    row = lambda s: [s.t, s.dt, s.iterations, s.iterations_v, s.avg_density_error, s.avg_divergence_error]
    rows, nonfinite, secs, rho = [], None, 0.0, None
    for k in range(steps):
        t0 = time.perf_counter()
        try:
            rows.append(row(ctx.step(1)[0][sid]))
        except SPHError as e:
            if e.code != SPH_E_NONFINITE:
                raise
            nonfinite = k + 1
            break
        ctx.synchronize()
        secs += time.perf_counter() - t0
        rho = ctx.fluid_state(sid)["rho"]
    ctx.close()
    done = len(rows)
    out = {"workload": "D1M flush: the D1M dam break with the block flush against the walls (first layer r from the "
                       "boundary particles), %d fluid particles" % n, "particles": n,
           "gpu": {"steps": "1-%d" % done, "value": n * done / secs if done else None, "unit": "particle-steps/s",
                   "ms_per_step": 1000.0 * secs / done if done else None,
                   "nonfinite_at_step": nonfinite,
                   "mean_density_error": float(np.mean(np.maximum(rho / 1000.0 - 1.0, 0.0))) if rho is not None else None,
                   "steps_stats": rows},
           "reference_first_steps": None}
    if S.reference_available():
        r = S.run_reference(sc, ref_steps, threads=threads, timeout=900)
        out["reference_first_steps"] = r["stats"].tolist()
    return out


def configs_table(threads):
    """Every workload's GPU throughput next to the CPU legs that fit the
    box's time budget (the figures of the measured-results table), one row
    at a time."""
    from sph2o_amd import scenes as SC

    from sph2o_amd import _lib

    sc, x, _ = SC.w8k()
    td = tempfile.mkdtemp()
    path = os.path.join(td, "w8k.json")
    SC.write_json(sc, path)
    _, _, xb = _lib.scene_from_json(path)
    log("configs: W8k GPU, 2,000 steps")
    n, g = gpu_leg(sc, x, None, [(1, 2000)])
    log("configs: W8k CPU legs")
    yield ({"workload": "W8k: WCSPH dam break, 8,000 particles, dt = 1e-4, 2,000 steps", "particles": n,
                 "gpu": g, "cpu_reference_all_threads": ref_leg(sc, 200, threads, x=x, what="WCSPH step"),
                 "cpu_reference_1_thread": ref_leg(sc, 200, 1, x=x, what="WCSPH step"),
                 "cpu_port_all_threads": port_leg(sc, 200, threads, x, xb=xb)})
    sc, x, v = SC.p4m()
    log("configs: P4M GPU, steps 1-20 and 21-200")
    n, g = gpu_leg(sc, x, v, [(1, 20), (21, 200)])
    log("configs: P4M CPU leg")
    yield ({"workload": "P4M: PCISPH neighbour-search stress case, %d particles (jittered lattice + 1.5x "
                             "cluster, Gaussian velocities)" % n, "particles": n, "gpu": g,
                 "cpu_reference_all_threads": ref_leg(sc, 2, threads, x=x, v=v, what="PCISPH step")})
    log("configs: D1M flush, GPU 20 steps, reference 3 steps")
    yield d1m_flush_leg(threads)


def batched_gain(steps, warmup=12):
    """B64: 64 DFSPH scenes (2k..200k particles) advanced together in one
    context vs the same scenes as 64 sequential single-scene contexts. The
    warmup steps cover each context's search tuning (12 timed searches: the
    one at init and one per step)."""
    from sph2o_amd import scenes as SC
    from sph2o_amd.simulation import Context

    td = tempfile.mkdtemp()
    paths = []
    for k, sc in enumerate(SC.b64()):
        p = os.path.join(td, "b%02d.json" % k)
        SC.write_json(sc, p)
        paths.append(p)
    batch = Context(0)
    for p in paths:
        batch.load_scene_json(p)
    batch.init()
    total = sum(batch.num_fluid(s) for s in range(len(paths)))
    batch.step(warmup)
    batch.synchronize()
    t0 = time.perf_counter()
    st = batch.step(steps)
    batch.synchronize()
    tb = time.perf_counter() - t0
    its_b = float(np.mean([[x.iterations + x.iterations_v for x in row] for row in st]))
    search_b = search_info(batch)
    batch.close()
    ts = 0.0
    its_s = []
    lanes_s = {}
    for p in paths:
        c = Context(0)
        c.load_scene_json(p)
        c.init()
        c.step(warmup)
        c.synchronize()
        t0 = time.perf_counter()
        st = c.step(steps)
        c.synchronize()
        ts += time.perf_counter() - t0
        its_s.append(np.mean([row[0].iterations + row[0].iterations_v for row in st]))
        ln = str(c.search_lanes()[0])
        lanes_s[ln] = lanes_s.get(ln, 0) + 1
        c.close()
    b_ps = total * steps / tb
    s_ps = total * steps / ts
    return {"scenes": len(paths), "fluid_particles": total, "steps": steps, "warmup": warmup,
            "batched_particle_steps_per_s": b_ps, "sequential_particle_steps_per_s": s_ps, "gain": b_ps / s_ps,
            "mean_iterations_per_scene_step": {"batched": its_b, "sequential": float(np.mean(its_s))},
            "search_lanes": {"batched": search_b, "sequential_contexts_by_lanes": lanes_s}}


def b64_timed(steps):
    """The B64 context alone: 3 warmup steps, a 0.5 s pause (a marker in a
    kernel trace), then `steps` timed steps. Prints the wall time."""
    from sph2o_amd import scenes as SC
    from sph2o_amd.simulation import Context

    td = tempfile.mkdtemp()
    ctx = Context(0)
    for k, sc in enumerate(SC.b64()):
        p = os.path.join(td, "b%02d.json" % k)
        SC.write_json(sc, p)
        ctx.load_scene_json(p)
    ctx.init()
    ctx.step(3)
    ctx.synchronize()
    time.sleep(0.5)
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.synchronize()
    t = time.perf_counter() - t0
    ctx.close()
    print(json.dumps({"steps": steps, "wall_s": t, "ms_per_step": 1000.0 * t / steps}), flush=True)


def a100k_sweep(steps):
    """A100k DFSPH (8 + 8 iterations, checkpoint every 20): an untimed
    40-step gradient (the tapes' allocation), the recording forward of
    `steps` steps, a 0.5 s pause (a marker in a kernel trace), then the
    reverse sweep. Prints the wall times."""
    from sph2o_amd import scenes as SC
    from sph2o_amd.simulation import Context
    from sph2o_amd._lib import SPH_E_NONFINITE, SPHError

    sc, _, _ = SC.a100k(2, dt=5e-4)
    td = tempfile.mkdtemp()
    path = os.path.join(td, "a100k.json")
    SC.write_json(sc, path)
    ctx = Context(0)
    sid = ctx.load_scene_json(path)
    ctx.init()
    x0 = ctx.fluid_state(sid)["x"]
    ctx.forward_record(40, 20)
    ctx.adjoint({sid: (np.zeros_like(x0), None)})
    ctx.set_fluid_state(sid, x0, np.zeros_like(x0))
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.forward_record(steps, 20)
    ctx.synchronize()
    t_rec = time.perf_counter() - t0
    xT = ctx.fluid_state(sid)["x"].astype(np.float64)
    seed = (2.0 * (xT - SC.adjoint_target(x0)) / len(x0)).astype(np.float32)
    time.sleep(0.5)
    t0 = time.perf_counter()
    try:
        ctx.adjoint({sid: (seed, None)})
    except SPHError as e:
        if e.code != SPH_E_NONFINITE:
            raise
    t_adj = time.perf_counter() - t0
    ctx.close()
    print(json.dumps({"steps": steps, "record_s": t_rec, "reverse_s": t_adj}), flush=True)


def _union(iv):
    """Total length of the union of [start, end) intervals (sorted by start)."""
    if not iv:
        return 0
    busy, (cs, ce) = 0, iv[0][:2]
    for s_, e_ in (q[:2] for q in iv[1:]):
        if s_ > ce:
            busy += ce - cs
            cs, ce = s_, e_
        else:
            ce = max(ce, e_)
    return busy + ce - cs


def kernel_busy(trace_dir, steps):
    """From a rocprofv3 --kernel-trace CSV of `--b64-steps` or `--a100k-sweep`:
    the dispatches after the marker pause, their span, the union of their
    intervals, and per stream the union of its own (a sum of those above the
    overall union is time with kernels of two streams running at once)."""
    import csv
    import glob
    from collections import Counter, defaultdict

    rows = list(csv.DictReader(open(glob.glob(os.path.join(trace_dir, "*kernel_trace.csv"))[0])))
    iv = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"], r.get("Stream_Id", "0"))
                for r in rows)
    gap, k = max((iv[q + 1][0] - iv[q][1], q) for q in range(len(iv) - 1))
    seg = iv[k + 1:]
    busy = _union(seg)
    span = seg[-1][1] - seg[0][0]
    streams = defaultdict(list)
    for q in seg:
        streams[q[3]].append(q)
    per = {sid: _union(v) / 1e6 for sid, v in streams.items()}
    print(json.dumps({"marker_gap_ms": gap / 1e6, "dispatches": len(seg), "span_ms": span / 1e6,
                      "kernel_busy_ms": busy / 1e6, "busy_fraction": busy / span,
                      "dispatches_per_step": len(seg) / float(steps),
                      "streams": {sid: {"dispatches": len(streams[sid]), "kernel_busy_ms": per[sid]} for sid in per},
                      "concurrent_ms": sum(per.values()) - busy / 1e6,
                      "top": Counter(n.split("(")[0] for _, _, n, _ in seg).most_common(8)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=100, help="fluid lattice edge (100 -> 1M particles)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-steps-serial", type=int, default=2, help="steps of the 1-thread reference leg (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=20, help="B64 batched-gain steps (0 = skip)")
    ap.add_argument("--adjoint-steps", type=int, default=200, help="A100k adjoint-cost steps (0 = skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="finish with a device-to-device copy of this size (HBM counter calibration under rocprofv3)")
    ap.add_argument("--b64-steps", type=int, default=0,
                    help="only time this many steps of the B64 context after a marker pause (for a kernel trace)")
    ap.add_argument("--configs", action="store_true",
                    help="print the W8k, P4M and flush-D1M GPU / CPU figures instead of the bench line")
    ap.add_argument("--a100k-sweep", type=int, default=0,
                    help="only record this many A100k DFSPH steps, pause, and run the reverse sweep (for a kernel trace)")
    ap.add_argument("--kernel-busy", metavar="TRACE_DIR",
                    help="summarise a rocprofv3 kernel trace of --b64-steps or --a100k-sweep (no GPU use)")
    args = ap.parse_args()
    if args.kernel_busy:
        kernel_busy(args.kernel_busy, args.b64_steps or args.a100k_sweep or 300)
        return
    if args.a100k_sweep > 0:
        a100k_sweep(args.a100k_sweep)
        return
    if args.b64_steps > 0:
        b64_timed(args.b64_steps)
        return
    if args.configs:
        threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
        for row in configs_table(threads):
            print(json.dumps(row, allow_nan=False), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    from sph2o_amd import scenes as SC
    from sph2o_amd.simulation import Context

    sc, _, _ = SC.d1m(args.n)
    td = tempfile.mkdtemp()
    scene_path = os.path.join(td, "d1m.json")
    SC.write_json(sc, scene_path)

    ctx = Context(0)
    sid = ctx.load_scene_json(scene_path)
    nf, nb = ctx.num_fluid(sid), ctx.num_boundary(sid)
    log("workload: %d fluid + %d boundary particles" % (nf, nb))
    t0 = time.time()
    ctx.init()
    log("init %.2f s" % (time.time() - t0))

    # warmup (untimed); the last few warmup steps are profiled to find the dominant kernel
    prof_steps = min(5, args.warmup)
    if args.warmup - prof_steps:
        ctx.step(args.warmup - prof_steps)
    ctx.set_profiling(True)
    ctx.step(prof_steps)
    ks = ctx.kernel_stats()
    ctx.set_profiling(False)
    ctx.reset_kernel_stats()
    dominant = max(ks.items(), key=lambda kv: kv[1][0])[0] if ks else "velocity"
    names = [ctx.L.sph_kernel_name(k).decode() for k in range(ctx.L.sph_kernel_count())]
    kid = names.index(dominant)
    log("dominant kernel: %s (%s)" % (dominant, ", ".join("%s %.2f ms" % (k, v[0]) for k, v in
                                                          sorted(ks.items(), key=lambda kv: -kv[1][0])[:6])))

    # timed region: only the dominant kernel carries an event pair
    ctx.set_profiling(1 << kid if kid > 0 else 1)
    ctx.synchronize()
    t0 = time.perf_counter()
    stats = ctx.step(args.steps)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    kms, klaunch = ctx.kernel_stats().get(dominant, (0.0, 0))
    ctx.set_profiling(False)

    st = ctx.fluid_state(sid)
    rho_err = float(np.mean(np.maximum(st["rho"] / 1000.0 - 1.0, 0.0)))
    its = np.array([[s[sid].iterations, s[sid].iterations_v] for s in stats])
    pairs = pair_throughput(ctx, sid, its, args.steps / elapsed)
    value = nf * args.steps / elapsed
    own, reads_b = ALGO_BYTES.get(dominant, (0, False))
    bytes_per_launch = nf * (own + (16.0 * nb / nf if reads_b else 0.0))
    avg_ms = kms / klaunch if klaunch else float("nan")
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if klaunch else None
    traffic = measured_traffic(dominant)
    valu = measured_valu(dominant)
    valu_frac = valu / (avg_ms * 1e-3) / VALU_PEAK_WAVE_INSTS if (valu and klaunch) else None
    ta_frac = measured_ta_busy(dominant)
    search = search_info(ctx)

    ctx.close()
    if args.calibrate_copy > 0:
        calibration_copy(args.calibrate_copy << 20)
    batched = None
    if args.batch_steps > 0:
        log("batched gain: B64, %d steps" % args.batch_steps)
        batched = batched_gain(args.batch_steps)

    adjoint = None
    if args.adjoint_steps > 0:
        log("adjoint cost: A100k WCSPH and DFSPH, %d steps" % args.adjoint_steps)
        adjoint = {"wcsph": adjoint_cost(args.adjoint_steps, 20, 0),
                   "dfsph": adjoint_cost(args.adjoint_steps, 20, 2)}

    cpu = None
    cpu_all = None
    if not args.no_cpu:
        # the host CPUs one GPU's job may use: OMP_NUM_THREADS (16 per GPU on
        # the MI355X hosts, whatever os.cpu_count() reports there)
        threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
        log("cpu baseline: %d steps of the reference on %d threads, %d on 1 thread, %d of the FP32 port"
            % (args.cpu_steps, threads, args.cpu_steps_serial, args.cpu_steps))
        cpu = cpu_baseline(scene_path, nf, args.cpu_steps, threads)
        cpu_all = {"host": host_info(), "reference_all_threads": cpu,
                   "reference_1_thread": cpu_baseline(scene_path, nf, args.cpu_steps_serial, 1)
                   if args.cpu_steps_serial > 0 else None,
                   "port_all_threads": cpu_port(scene_path, nf, args.cpu_steps, threads)}

    line = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000.0 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "D1M: DFSPH dam break, %d fluid + %d Akinci boundary particles, r=0.005 m, "
                               "h=0.02 m, 4x2x2 m box, CFL dt<=1e-3, eta=0.01%%, etaV=0.1%%, surface tension 0.1"
                               % (nf, nb),
                   "fluid_particles": nf, "boundary_particles": nb},
        "mean_density_error": rho_err,
        "density_error_ok": rho_err <= 1e-3,
        "iterations_per_step": {"density": float(its[:, 0].mean()), "divergence": float(its[:, 1].mean())},
        "neighbour_pairs": pairs,
        "search_lanes": search,
        "sim_time_s": float(stats[-1][sid].t),
        "roofline": {"kernel": dominant, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                     "avg_launch_ms": avg_ms, "launches": klaunch, "algorithmic_bytes_per_launch": bytes_per_launch,
                     # SQ_INSTS_VALU per launch (counter run of this workload) over this run's launch time,
                     # against the VALU issue peak
                     "valu_issue": {"wave_insts_per_launch": valu, "peak_wave_insts_per_s": VALU_PEAK_WAVE_INSTS,
                                    "frac": valu_frac},
                     # TA_BUSY_avr over the kernel's dispatch time in the counter run of this workload: the
                     # vector-memory address unit, the busiest unit of the neighbour passes (scattered 64-lane
                     # gathers, ~37 L1 lines each)
                     "vmem": {"ta_busy_frac": ta_frac, "source": "profiles/d1m_vmem_pipeline_counters.md"}},
        "batched_gain": batched["gain"] if batched else None,
        "batched": batched,
        # (recording forward + reverse sweep) / plain forward, the larger of the two methods
        "adjoint_fwd_cost": max(a["gradient_over_forward"] for a in adjoint.values()) if adjoint else None,
        "adjoint": adjoint,
        "cpu_baseline": cpu,
        "cpu_baselines": cpu_all,
        "profiled_warmup": {"steps": prof_steps,
                            "kernel_ms": {k: round(v[0], 4) for k, v in sorted(ks.items(), key=lambda kv: -kv[1][0])},
                            "launches": {k: v[1] for k, v in ks.items()}},
    }
    print(json.dumps(line, allow_nan=False), flush=True)


if __name__ == "__main__":
    main()
