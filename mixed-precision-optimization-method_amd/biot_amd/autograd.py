"""torch.autograd binding of the checkpointed forward run and its hand-written adjoint.

`traces(params, solver, nt, every)` returns the receiver traces of the parameter set as a
torch tensor whose backward runs biot_adjoint on the GPU: the adjoint source is the
incoming trace gradient, the coefficient gradients are summed in FP64 on the device and
mapped to (lambda, mu, alpha, M, rho, rho_f) by the exact chain rule of the parameter
builder.  `misfit_gradient` is the same in one call with the misfit reduced on the device.
The reference has no adjoint; nothing here has a counterpart there.
"""
import numpy as np
import torch

from ._lib import FIXED_PARAM_NAMES, PARAM_NAMES


def _np(t):
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


class BiotTraces(torch.autograd.Function):
    """Inputs: the six inverted parameter tensors (nz, ny, nx), then the solver (a Solver
    with source and receivers set, any storage policy), the fixed arrays (dict with C1, tphi),
    nt and the checkpoint interval.  Output: traces (ncomp, nt_max, nrec), float64 for an
    FP64 solver, else float32."""

    @staticmethod
    def forward(ctx, lam, mu, alpha, M, rho, rhof, solver, fixed, nt, every):
        vals = [lam, mu, alpha, M, rho, rhof]
        params = {k: _np(v) for k, v in zip(PARAM_NAMES, vals)}
        params.update({k: np.asarray(fixed[k], np.float64) for k in FIXED_PARAM_NAMES})
        solver.set_params(params)
        solver.forward_checkpointed(nt, every)
        d = solver.traces_f64() if solver.precision == "fp64" else solver.traces()
        ctx.solver = solver
        ctx.generation = solver.ckpt_generation
        ctx.dtypes = [v.dtype for v in vals]
        ctx.device = lam.device
        return torch.from_numpy(d).to(lam.device)

    @staticmethod
    def backward(ctx, g):
        # the adjoint recomputes from the solver's checkpoints, which belong to its latest
        # checkpointed run: a later run on the same solver would give that run's gradient
        if ctx.solver.ckpt_generation != ctx.generation:
            raise RuntimeError("the solver ran another checkpointed forward after these traces were "
                               "computed; call backward before reusing the solver, or use one solver "
                               "per differentiable run")
        adj = g.detach().to("cpu", torch.float32).contiguous().numpy()
        grads, _ = ctx.solver.adjoint(adj, params=True)
        out = [torch.from_numpy(grads[k]).to(ctx.device, dt) for k, dt in zip(PARAM_NAMES, ctx.dtypes)]
        return (*out, None, None, None, None)


def traces(params, solver, nt, every=50):
    """Differentiable receiver traces; params: dict of the six tensors and the fixed arrays."""
    fixed = {k: params[k] for k in FIXED_PARAM_NAMES}
    return BiotTraces.apply(*[params[k] for k in PARAM_NAMES], solver, fixed, nt, every)


def misfit_gradient(params, solver, observed, nt, every=50):
    """(J, gradients): J = 1/2 sum (d - observed)^2 reduced in FP64 on the device and the
    adjoint gradients w.r.t. the six parameters (dict of float64 arrays)."""
    p = {k: np.asarray(v.detach().cpu().numpy() if torch.is_tensor(v) else v, np.float64)
         for k, v in params.items()}
    solver.set_params(p)
    solver.forward_checkpointed(nt, every)
    J, res = solver.misfit(observed)
    grads, _ = solver.adjoint(res, params=True)
    return J, grads
