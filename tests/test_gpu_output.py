"""GPU check of the streamed trace copy (biot_run_stream_traces): the traces copied to the
host on a side stream while the loop runs equal the ones read after a plain run, bit for
bit, for chunk sizes that do and do not divide the step count, and across two calls."""
import numpy as np
import pytest

from biot_amd import configs

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
