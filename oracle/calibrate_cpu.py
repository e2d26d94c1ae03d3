"""CPU-baseline calibration (test infrastructure; runs only where the
reference tree and its headless build exist, i.e. not on the GPU hosts).

Times the compiled reference (oracle/_ref/sph2o_ref: the original sources,
shipped -O3 -ffast-math flags, OpenMP) against this repository's clean-room
FP32 restatement (oracle/libsphref.so) on the same scenes and thread counts,
step() calls only:

    python oracle/calibrate_cpu.py [threads]

W8k: the WCSPH dam break, steps 1-200; D125k: the DFSPH dam break at
50^3 = 125,000 particles, steps 1-10. Prints one JSON object per case with
both particle-steps/s figures and their ratio.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import sphref as S  # noqa: E402
from sph2o_amd import _lib, scenes as SC  # noqa: E402


def _scene_particles(sc):
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "s.json")
        SC.write_json(sc, path)
        _, x, xb = _lib.scene_from_json(path)
    return x, xb


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    cases = []
    sc, x, _ = SC.w8k()
    cases.append(("W8k", sc, x, 200))
    sc, _, _ = SC.d1m(50)
    cases.append(("D125k", sc, None, 10))
    for name, sc, x, steps in cases:
        xs, xb = _scene_particles(sc)
        x = xs if x is None else x
        r = S.run_reference(sc, steps, x=x, threads=threads, timeout=3600)
        ref_secs = float(r["meta"][3])
        o = S.Oracle(S.params_from_scene(sc), x, None, r["bx"], bits=32)
        o.init()
        t0 = time.perf_counter()
        o.step(steps)
        port_secs = time.perf_counter() - t0
        o.close()
        n = len(x)
        print(json.dumps({"case": name, "particles": n, "steps": steps, "threads": threads,
                          "reference_particle_steps_per_s": n * steps / ref_secs,
                          "port_particle_steps_per_s": n * steps / port_secs,
                          "port_over_reference": ref_secs / port_secs}), flush=True)


if __name__ == "__main__":
    main()
