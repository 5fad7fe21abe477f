"""The reference's text output format (outputfunc.cpp:35-50, 97-151) and the binary writer,
through the library's host-only writers (no GPU)."""
import numpy as np

from biot_amd import _lib


def test_text_2d_matches_reference_format(tmp_path):
    # every value "%e " (C printf), each row ended by a newline and followed by an empty line
    a = np.array([[1.0, -2.5, 3e-10], [0.0, 65504.0, -1.25e-38]], np.float32)
    p = tmp_path / "sis.dat"
    _lib.write_text_2d(str(p), a)
    expect = ("1.000000e+00 -2.500000e+00 3.000000e-10 \n\n"
              "0.000000e+00 6.550400e+04 -1.250000e-38 \n\n")
    assert p.read_text() == expect


def test_text_xz_takes_the_middle_y_slice(tmp_path):
    nz, ny, nx = 3, 5, 4
    vol = np.arange(nz * ny * nx, dtype=np.float32).reshape(nz, ny, nx)
    p = tmp_path / "snap.dat"
    _lib.write_text_xz(str(p), vol)
    rows = [r for r in p.read_text().split("\n") if r]
    assert len(rows) == nz
    for k, r in enumerate(rows):
        vals = np.array([float(v) for v in r.split()], np.float32)
        assert np.array_equal(vals, vol[k, ny // 2])
    # blank line after every row, values separated by single spaces with a trailing one
    assert p.read_text().count("\n\n") == nz and rows[0].endswith(" ")


def test_binary_round_trip(tmp_path):
    a = np.random.default_rng(1).standard_normal(1000).astype(np.float32)
    p = tmp_path / "tr.bin"
    _lib.write_binary(str(p), a)
    assert np.array_equal(np.fromfile(p, dtype="<f4"), a)
