"""Differentiable WCSPH and fixed-iteration DFSPH rollouts as a
``torch.autograd.Function``.

    x_T, v_T = rollout(context, scene, x0, v0, viscosity, surface_tension,
                       nsteps, checkpoint_every=20)

The forward pass runs ``nsteps`` steps of the scene's solver (WCSPH, or
DFSPH with a fixed time step and min = max iterations) on the HIP path from (x0, v0)
through ``sph_forward_record`` (checkpoints every ``checkpoint_every`` steps);
the backward pass runs the adjoint sweep of libsph2o.so (``sph_adjoint_run``)
and returns dL/dx0, dL/dv0, dL/d(viscosity) and dL/d(surface tension). The
state crosses the C ABI as N x 3 FP32 host arrays in the original particle
order; tensors on a GPU are copied. There is no CPU fallback: forward and
backward are the HIP kernels. One recorded rollout per context at a time: a
second forward on the same context before the backward of the first makes
that backward fail with ``SPH_E_STATE``.

The reference (WCSPHSolver.cpp:14-78, DFSPHSolver.cpp:33-107) has no
gradients; the forward step is the same as ``Context.step``. A non-finite
gradient raises ``SPHError`` (``SPH_E_NONFINITE``) in the backward pass.
"""
import torch

__all__ = ["SPHRollout", "WCSPHRollout", "rollout"]


class SPHRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, viscosity, surface_tension, context, scene, nsteps, checkpoint_every):
        context.set_fluid_state(scene, x0.detach().cpu().numpy(), v0.detach().cpu().numpy())
        context.set_coefficients(scene, float(viscosity), float(surface_tension))
        context.forward_record(int(nsteps), int(checkpoint_every))
        st = context.fluid_state(scene)
        ctx.context = context
        ctx.scene = scene
        ctx.like = (x0, v0, viscosity, surface_tension)
        xT = torch.from_numpy(st["x"]).to(device=x0.device, dtype=x0.dtype)
        vT = torch.from_numpy(st["v"]).to(device=v0.device, dtype=v0.dtype)
        return xT, vT

    @staticmethod
    def backward(ctx, gx, gv):
        seed = (None if gx is None else gx.detach().float().cpu().numpy(),
                None if gv is None else gv.detach().float().cpu().numpy())
        res = ctx.context.adjoint({ctx.scene: seed})[ctx.scene]
        x0, v0, nu, gam = ctx.like
        return (torch.from_numpy(res["x0"]).to(device=x0.device, dtype=x0.dtype),
                torch.from_numpy(res["v0"]).to(device=v0.device, dtype=v0.dtype),
                torch.tensor(res["viscosity"], dtype=nu.dtype, device=nu.device),
                torch.tensor(res["surface_tension"], dtype=gam.dtype, device=gam.device),
                None, None, None, None)


def rollout(context, scene, x0, v0, viscosity, surface_tension, nsteps, checkpoint_every=20):
    """x_T, v_T after `nsteps` steps; differentiable in x0, v0 (N x 3
    tensors) and the viscosity / surface-tension coefficients (scalars or
    0-d tensors)."""
    nu = viscosity if torch.is_tensor(viscosity) else torch.tensor(float(viscosity), dtype=torch.float64)
    gam = surface_tension if torch.is_tensor(surface_tension) else torch.tensor(float(surface_tension),
                                                                               dtype=torch.float64)
    return SPHRollout.apply(x0, v0, nu, gam, context, scene, nsteps, checkpoint_every)


# the name under which the WCSPH-only first version was published
WCSPHRollout = SPHRollout
