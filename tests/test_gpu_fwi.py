"""GPU check of the inversion loop over torch.autograd (biot_amd.fwi): on three shots of a
small layered model with a +6 % Kb block in the true model, steepest descent on lambda, mu
and M from the model without the block lowers the misfit at every iteration, and the
autograd misfit's gradient equals the batch gradient it wraps."""
import numpy as np
import pytest
import torch

from biot_amd import configs, fwi, media

pytestmark = pytest.mark.gpu


def _cfg(medium):
    return configs.batch16(nshots=3, nt=150, model=(72, 72, 60), side=(40, 56), npml=8,
                           rec_stride=5, medium=medium)


def test_fwi_lowers_the_misfit():
    print("fwi 3 shots", flush=True)
    shape = (60, 72, 72)
    start = media.layered(shape)
    true = {k: v.copy() for k, v in start.items()}
    true["Kb"][20:40, 24:48, 24:48] *= 1.06
    cfg = _cfg(start)
    b = configs.make_batch(cfg, precision="fp32")
    p_true = media.params_from_medium(true)
    b.set_model(p_true)
    b.forward(cfg["nt"], graph_steps=0)
    obs = [b.traces(i) for i in range(3)]
    p0 = media.params_from_medium(start)
    # the autograd misfit and its gradient are the batch's own
    model = {k: torch.tensor(v, requires_grad=k in ("lambda", "M")) for k, v in p0.items()}
    J = fwi.misfit(model, b, obs, cfg["nt"], every=30)
    J.backward()
    b.set_model(p0)
    J2, g = b.gradient(cfg["nt"], obs, every=30)
    assert abs(J.item() - J2) <= 1e-12 * J2
    ga = model["lambda"].grad.numpy()  # FP64 atomics reorder the sums: equal to rounding
    assert np.abs(ga - g["lambda"]).max() <= 1e-12 * np.abs(g["lambda"]).max()
    final, hist = fwi.invert(b, p0, obs, cfg["nt"], every=30, iters=3, step=0.02,
                             log=lambda s: print("  " + s, flush=True))
    print("  misfit history: " + " ".join(f"{j:.4e}" for j in hist))
    b.close()
    assert len(hist) == 3
    assert all(b2 < a2 for a2, b2 in zip(hist, hist[1:]))
