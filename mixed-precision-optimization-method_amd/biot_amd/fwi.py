"""Full-waveform inversion over torch.autograd: the summed L2 misfit of a Batch of shots as a
differentiable function of the six parameter tensors (its backward is the batched adjoint),
and a steepest-descent loop with a backtracking step on relative parameter perturbations.
The reference is forward-only; nothing here has a counterpart there."""
import numpy as np
import torch

from ._lib import FIXED_PARAM_NAMES, PARAM_NAMES


class BatchMisfit(torch.autograd.Function):
    """Inputs: the six model-grid parameter tensors (nz, ny, nx), the Batch, the fixed arrays
    (C1, tphi), the observed traces (one array per shot), nt and the checkpoint interval.
    Output: J (float64 scalar)."""

    @staticmethod
    def forward(ctx, lam, mu, alpha, M, rho, rhof, batch, fixed, observed, nt, every):
        vals = [lam, mu, alpha, M, rho, rhof]
        p = {k: v.detach().to("cpu", torch.float64).numpy() for k, v in zip(PARAM_NAMES, vals)}
        p.update({k: np.asarray(fixed[k], np.float64) for k in FIXED_PARAM_NAMES})
        batch.set_model(p)
        J, g = batch.gradient(nt, observed, every=every)
        ctx.grads = [torch.from_numpy(g[k]).to(v.device, v.dtype) for k, v in zip(PARAM_NAMES, vals)]
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        return (*[gJ * g for g in ctx.grads], None, None, None, None, None)


def misfit(model, batch, observed, nt, every=50):
    """J of a model dict (six tensors that may require grad, plus the fixed C1, tphi)."""
    fixed = {k: model[k] for k in FIXED_PARAM_NAMES}
    return BatchMisfit.apply(*[model[k] for k in PARAM_NAMES], batch, fixed, observed, nt, every)


def forward_misfit(batch, model, observed, nt):
    """J of a model without its gradient (a forward run and the host sum)."""
    batch.set_model({k: np.asarray(v.detach().numpy() if torch.is_tensor(v) else v, np.float64)
                     for k, v in model.items()})
    batch.forward(nt, graph_steps=0)
    return 0.5 * sum(float(np.sum((batch.traces(i).astype(np.float64) - o) ** 2))
                     for i, o in enumerate(observed))


def invert(batch, model0, observed, nt, every=50, iters=5, step=0.01, params=("lambda", "mu", "M"),
           tries=4, log=None):
    """Steepest descent on x (p = p0 (1 + x)) for the selected parameters: the gradient from
    the batched adjoint through torch.autograd, a step that moves the largest |x| component by
    `step`, halved until the misfit falls (at most `tries` times).  Returns the final model
    (dict of numpy arrays) and the misfit history."""
    base = {k: torch.as_tensor(np.asarray(v, np.float64)) for k, v in model0.items()}
    obs = [np.asarray(o, np.float64) for o in observed]
    x = {k: torch.zeros_like(base[k], requires_grad=True) for k in params}

    def model_of(xv):
        m = {k: base[k] * (1 + xv[k]) if k in xv else base[k] for k in PARAM_NAMES}
        m.update({k: base[k] for k in FIXED_PARAM_NAMES})
        return m

    history = []
    for it in range(iters):
        for v in x.values():
            v.grad = None
        J = misfit(model_of(x), batch, observed, nt, every)
        J.backward()
        J0 = J.item()
        history.append(J0)
        gmax = max(float(v.grad.abs().max()) for v in x.values())
        if gmax == 0:
            break
        alpha = step / gmax
        accepted = False
        for _ in range(tries):
            trial = {k: (v - alpha * v.grad).detach() for k, v in x.items()}
            Jt = forward_misfit(batch, model_of(trial), obs, nt)
            if log:
                log(f"iteration {it}: J {J0:.6e}, trial step {alpha * gmax:.2e} -> {Jt:.6e}")
            if Jt < J0:
                with torch.no_grad():
                    for k in x:
                        x[k].copy_(trial[k])
                accepted = True
                break
            alpha *= 0.5
        if not accepted:
            break
    final = {k: v.detach().numpy() for k, v in model_of({k: v.detach() for k, v in x.items()}).items()}
    return final, history
