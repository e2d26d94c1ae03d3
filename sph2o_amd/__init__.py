"""MI355X-native SPH particle step (WCSPH / PCISPH / DFSPH with Akinci
boundaries) behind the C ABI of include/sph2o.h.

The compute path is libsph2o.so (hand-written HIP kernels for gfx950); this
package only binds it (``_lib``), mirrors the reference's host interface
(``simulation.SPHSimulation``) and builds workloads (``scenes``).
"""
from ._lib import SPHError, SceneParams, StepStats, load  # noqa: F401
from .simulation import Context, SPHSimulation  # noqa: F401

__all__ = ["Context", "SPHSimulation", "SPHError", "SceneParams", "StepStats", "load"]
