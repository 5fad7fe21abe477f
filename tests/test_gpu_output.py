"""GPU check of the streamed copies (biot_run_stream_traces, biot_run_stream): the traces and
xz snapshots copied to the host on a side stream while the loop runs equal the ones read after
a plain run, bit for bit, for chunk sizes that do and do not divide the step count, for
snapshot steps inside and at the ends of chunks, and across two calls."""
import numpy as np
import pytest

from biot_amd import configs
from biot_amd.solver import BiotError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("chunk", [7, 64, 500])
def test_streamed_traces_equal_plain_run(chunk):
    print(f"stream traces chunk {chunk}", flush=True)
    cfg = configs.fwd64(order=4, nt=150)
    s = configs.make_solver(cfg, precision="fp32")
    s.run(cfg["nt"])
    ref = s.traces()
    s.reset()
    a = s.run_stream_traces(100, chunk)       # steps 0..99
    b = s.run_stream_traces(50, chunk)        # steps 100..149 (rows 0..99 untouched: zeros)
    s.close()
    assert np.array_equal(a[:, :100], ref[:, :100])
    assert not a[:, 100:].any()
    assert np.array_equal(b[:, 100:], ref[:, 100:])


@pytest.mark.parametrize("precision,chunk",
                         [("fp32", 7), ("fp32", 64), ("fp16", 50), ("fp16c", 500), ("fp16t", 64), ("fp16m", 64)])
def test_streamed_snapshots_equal_plain_run(precision, chunk):
    print(f"stream snapshots {precision} chunk {chunk}", flush=True)
    cfg = configs.fwd64(order=4, nt=150)
    steps = [0, 9, 49, 63, 149]
    s = configs.make_solver(cfg, precision=precision)
    ref_snaps, done = [], 0
    for t in steps:                     # plain run, read after each listed step
        s.run(t + 1 - done)
        done = t + 1
        ref_snaps.append(s.snapshot_xz("tzz", 20))
    ref_tr = s.traces()
    s.reset()
    tr, snaps = s.run_stream(150, chunk, snap_field="tzz", snap_j=20, snap_steps=steps)
    assert s.step() == 150
    assert np.array_equal(tr, ref_tr)
    assert snaps.shape == (len(steps), s.nz, s.nx)
    for a, b in zip(snaps, ref_snaps):
        assert np.array_equal(a, b)
    assert np.abs(snaps[-1]).max() > 0
    # a second call counts its snapshot steps from its own first step; no traces asked for
    s.reset()
    s.run(100)
    ref = s.snapshot_xz("vx", 32)
    s.reset()
    s.run(40)
    tr2, snaps2 = s.run_stream(60, chunk, snap_field="vx", snap_j=32, snap_steps=[59], traces=False)
    s.close()
    assert tr2 is None
    assert np.array_equal(snaps2[0], ref)


def test_streamed_snapshot_arguments():
    cfg = configs.fwd64(order=2, nt=20)
    s = configs.make_solver(cfg, precision="fp32")
    for kw in (dict(snap_steps=[5, 5]), dict(snap_steps=[7, 3]), dict(snap_steps=[10]),
               dict(snap_steps=[-1]), dict(snap_steps=[2], snap_j=64), dict(snap_steps=[2], snap_field=13)):
        with pytest.raises(BiotError):
            s.run_stream(10, 4, **kw)
        assert s.step() == 0            # rejected before any step ran
    tr, snaps = s.run_stream(10, 4)     # no snapshots
    assert snaps.shape == (0, s.nz, s.nx) and s.step() == 10
    s.close()


def test_text_files_of_a_run(tmp_path):
    cfg = configs.fwd64(order=2, nt=60)
    s = configs.make_solver(cfg, precision="fp32")
    s.run(60)
    s.write_traces_text(str(tmp_path / "sisx.dat"), comp=0)
    s.write_snapshot_text(str(tmp_path / "txx.dat"), "txx")
    rows = [r for r in (tmp_path / "sisx.dat").read_text().split("\n") if r]
    tr = s.traces()
    assert len(rows) == tr.shape[2]
    assert np.allclose(np.array(rows[5].split(), np.float32), tr[0, :, 5], rtol=1e-6, atol=0)
    snap = s.snapshot_xz("txx", s.ny // 2)
    rows = [r for r in (tmp_path / "txx.dat").read_text().split("\n") if r]
    assert np.allclose(np.array(rows[32].split(), np.float32), snap[32], rtol=1e-6, atol=0)
    s.close()


@pytest.mark.parametrize("precision", ["fp32", "fp16", "fp16m"])
def test_nonfinite_watchdog(precision):
    """A NaN sample in the wavelet (forced from the host; NaN arithmetic does not fault) reaches
    the receiver at the source cell at that step: biot_run reports BIOT_ERR_NONFINITE naming the
    step, so does the streamed loop, and biot_reset clears the flag."""
    print(f"watchdog {precision}", flush=True)
    cfg = configs.fwd64(order=2, nt=20)
    s = configs.make_solver(cfg, precision=precision, nt=20)
    cell = int(configs.cells(cfg, cfg["src"])[0])
    w = np.ones(20, np.float32)
    w[5] = np.nan
    s.set_source(cell, w, ("txx",))
    s.set_receivers([cell], ("txx",), 20)
    with pytest.raises(BiotError) as e:
        s.run(20)
    assert e.value.status == 4 and "at step 5" in str(e.value)
    s.reset()
    s.run(4)  # steps 0..3 are clean again after the reset
    with pytest.raises(BiotError) as e:
        s.run_stream_traces(16, 5)
    assert e.value.status == 4 and "at step 5" in str(e.value)
    w[5] = 1.0
    s.set_source(cell, w, ("txx",))
    s.reset()
    s.run(20)
    assert np.isfinite(s.traces()).all()
    s.close()
