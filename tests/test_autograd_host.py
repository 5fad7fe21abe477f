"""CPU check of the torch.autograd binding's host logic (biot_amd.autograd.BiotTraces) with a
stand-in solver object in place of a GPU handle: the backward hands the incoming trace
gradient to the solver's adjoint and returns its six parameter gradients in the inputs'
dtypes, and a backward after the same solver ran another checkpointed forward raises instead
of returning the later run's gradient."""
import numpy as np
import pytest
import torch

from biot_amd import autograd
from biot_amd._lib import FIXED_PARAM_NAMES, PARAM_NAMES

SHAPE = (3, 4, 5)


class StandInSolver:
    """The calls BiotTraces makes on a Solver. Traces are linear in the parameters,
    d[c, t, r] = w[c, t, r] * sum_k a_k * sum(param_k), so the adjoint is known exactly."""

    precision = "fp32"

    def __init__(self):
        self.ckpt_generation = 0
        rng = np.random.default_rng(5)
        self.w = rng.standard_normal((2, 6, 3)).astype(np.float32)
        self.a = np.arange(1.0, 7.0)
        self.params = None
        self.adjoint_calls = 0

    def set_params(self, params):
        self.params = {k: np.array(v) for k, v in params.items()}

    def forward_checkpointed(self, nt, every):
        self.ckpt_generation += 1
        self.run_params = self.params

    def traces(self):
        s = sum(a * self.run_params[k].sum() for a, k in zip(self.a, PARAM_NAMES))
        return (self.w * s).astype(np.float32)

    def adjoint(self, adj, params=True):
        self.adjoint_calls += 1
        g = float((adj.astype(np.float64) * self.w).sum())
        return {k: np.full(SHAPE, a * g) for a, k in zip(self.a, PARAM_NAMES)}, None


def _params(scale, dtype=torch.float64):
    p = {k: torch.full(SHAPE, scale * (i + 1), dtype=dtype, requires_grad=True)
         for i, k in enumerate(PARAM_NAMES)}
    p.update({k: np.ones(SHAPE) for k in FIXED_PARAM_NAMES})
    return p


def test_backward_returns_the_adjoint_of_the_linear_map():
    s = StandInSolver()
    p = _params(0.5, torch.float32)
    d = autograd.traces(p, s, nt=6, every=2)
    assert d.dtype == torch.float32 and d.shape == (2, 6, 3)
    y = torch.from_numpy(np.random.default_rng(1).standard_normal(d.shape).astype(np.float32))
    (d * y).sum().backward()
    g = float((y.numpy().astype(np.float64) * s.w).sum())
    for a, k in zip(s.a, PARAM_NAMES):
        assert p[k].grad.dtype == torch.float32
        assert np.allclose(p[k].grad.numpy(), a * g, rtol=1e-6)
    assert s.adjoint_calls == 1


def test_backward_after_another_checkpointed_run_raises():
    s = StandInSolver()
    p1, p2 = _params(1.0), _params(2.0)
    d1 = autograd.traces(p1, s, nt=6, every=2)
    d2 = autograd.traces(p2, s, nt=6, every=2)
    with pytest.raises(RuntimeError, match="another checkpointed forward"):
        d1.sum().backward()
    assert s.adjoint_calls == 0
    d2.sum().backward()              # the latest run still differentiates
    assert s.adjoint_calls == 1 and p2["lambda"].grad is not None
