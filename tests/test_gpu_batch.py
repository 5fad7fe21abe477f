"""GPU checks of the batched shots (biot_batch_* through the C-ABI).

* The persistent batched launch reproduces every shot's single-grid run (traces equal to
  1e-7 relative, FP32 and FP16), eagerly and from a captured HIP graph; the graph is
  replayed several times with eager runs in between and every replay is checked.
* The batched gradient equals the sum of the single-grid adjoint gradients of the shots,
  added into the model grid on the host (FP32: 1e-9 relative; FP64 atomics only reorder the
  sums).
* FP16 wavefield storage: the batched gradient against the FP32 one (bar in the test).
"""
import numpy as np
import pytest

from biot_amd import BiotError, configs, media
from biot_amd.batch import Batch
from biot_amd.solver import Solver
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _small(nt=120, precision_seed=3):
    m = media.perturbed((80, 96, 96), corr_cells=4.0, seed=precision_seed)
    return configs.batch16(nshots=3, nt=nt, model=(96, 96, 80), side=(32, 56), npml=8,
                           rec_stride=6, medium=m)


def _single(cfg, shot, precision):
    wx, wy, wz = shot["size"]
    s = Solver(wx, wy, wz, cfg["dx"], cfg["dt"], order=cfg["order"], precision=precision,
               npml=cfg["npml"])
    s.set_params(configs.window_params(cfg, shot))
    from biot_amd import ricker

    si, sj, sk = shot["src"]
    s.set_source(s.cell(si, sj, sk), ricker(cfg["f0"], cfg["dt"], cfg["nt"]), cfg["src_fields"])
    rec = np.asarray(shot["rec"])
    s.set_receivers(rec[:, 0] + rec[:, 1] * wx + rec[:, 2] * wx * wy, cfg["rec_fields"], cfg["nt"])
    return s


@pytest.mark.parametrize("precision", ["fp32", "fp16", "fp16t", "fp16c", "fp16m"])
def test_batch_forward_matches_single_grids_eager_and_graph(precision):
    print(f"batch forward {precision}", flush=True)
    cfg = _small()
    nt = cfg["nt"]
    b = configs.make_batch(cfg, precision=precision)
    b.set_launch(64, 4, 8, 64)
    refs = []
    for sh in cfg["shots"]:
        s = _single(cfg, sh, precision)
        s.run(nt)
        refs.append(s.traces())
        s.close()
    b.forward(nt, graph_steps=0)
    eager = [b.traces(i) for i in range(len(refs))]
    worst = max(rel_l2(e, r) for e, r in zip(eager, refs))
    print(f"  eager vs single grids: {worst:.3e}")
    assert worst <= 1e-7
    for replay in range(3):
        b.forward(nt, graph_steps=16)   # 7 graph launches + 8 single steps
        got = [b.traces(i) for i in range(len(refs))]
        diff = max(float(np.abs(g - e).max()) for g, e in zip(got, eager))
        print(f"  graph replay {replay}: max |graph - eager| = {diff:.3e}")
        assert diff == 0.0
        b.forward(nt, graph_steps=0)   # eager work between replays
    b.close()


def test_batch_gradient_equals_sum_of_single_grid_gradients():
    print("batch gradient fp32", flush=True)
    cfg = _small(nt=120)
    nt, every = cfg["nt"], 25
    true = media.perturbed((80, 96, 96), corr_cells=4.0, seed=3)
    true["Kb"] = true["Kb"] * 1.05
    tcfg = dict(cfg, medium=true)
    obs = []
    for sh in cfg["shots"]:
        s = _single(tcfg, sh, "fp32")
        s.run(nt)
        obs.append(s.traces())
        s.close()
    b = configs.make_batch(cfg, precision="fp32")
    J, g = b.gradient(nt, obs, every=every)
    b.close()
    ref = {k: np.zeros((80, 96, 96)) for k in g}
    Jref = 0.0
    for sh, o in zip(cfg["shots"], obs):
        s = _single(cfg, sh, "fp32")
        s.forward_checkpointed(nt, every)
        Js, res = s.misfit(o)
        gs, _ = s.adjoint(res)
        s.close()
        Jref += Js
        x0, y0, z0 = sh["origin"]
        wx, wy, wz = sh["size"]
        for k in ref:
            ref[k][z0:z0 + wz, y0:y0 + wy, x0:x0 + wx] += gs[k]
    print(f"  J batch {J:.9e} single {Jref:.9e}")
    assert abs(J - Jref) <= 1e-9 * Jref
    for k in ref:
        e = rel_l2(g[k], ref[k])
        print(f"  {k}: rel L2 {e:.2e}")
        assert e <= 1e-9


def test_batch_fp16_gradient_against_fp32():
    print("batch gradient fp16 vs fp32", flush=True)
    cfg = _small(nt=120)
    nt, every = cfg["nt"], 30
    true = media.perturbed((80, 96, 96), corr_cells=4.0, seed=3)
    true["Kb"] = true["Kb"] * 1.05
    tcfg = dict(cfg, medium=true)
    obs = []
    for sh in cfg["shots"]:
        s = _single(tcfg, sh, "fp32")
        s.run(nt)
        obs.append(s.traces())
        s.close()
    grads = {}
    for prec in ("fp32", "fp16"):
        b = configs.make_batch(cfg, precision=prec)
        grads[prec] = b.gradient(nt, obs, every=every)
        b.close()
    (J32, g32), (J16, g16) = grads["fp32"], grads["fp16"]
    print(f"  J fp32 {J32:.6e} fp16 {J16:.6e}")
    errs = {k: rel_l2(g16[k], g32[k]) for k in g32}
    print("  fp16 vs fp32 gradient rel L2: " + " ".join(f"{k} {e:.2e}" for k, e in errs.items()))
    assert abs(J16 - J32) <= 0.05 * J32
    assert max(errs.values()) <= 0.1


def test_batch_argument_errors():
    b = Batch(64, 64, 64, 5.0, 4e-4, order=4, precision="fp32", npml=8)
    p = media.params_from_medium(media.homogeneous((32, 32, 32)))
    w = np.zeros(10, np.float32)
    with pytest.raises(BiotError):   # window outside the model
        b.add_shot((40, 0, 0), (32, 32, 32), p, (16, 16, 16), w, [(16, 16, 8)], ("vx",))
    with pytest.raises(BiotError):   # no shots
        b.forward(10)
    b.add_shot((0, 0, 0), (32, 32, 32), p, (16, 16, 16), w, [(16, 16, 8)], ("vx",))
    with pytest.raises(BiotError):   # more steps than the receivers hold
        b.gradient(20, [np.zeros((1, 10, 1), np.float32)], every=5)
    b.forward(10, graph_steps=4)
    assert np.all(b.traces(0) == 0)
    b.close()
