"""CPU: the native text readers and writers of the reference's file formats
(pisco_read_txt_columns / pisco_write_txt_columns).  Files written by
numpy.savetxt, as scripts/create_cmb.py and scripts/fields2beams.py write
them, are read to the float each string rounds to (the C library's strtof,
called here on the same strings); the writer's output reads back unchanged
and matches numpy.savetxt's text; comments, blank lines and malformed rows
are handled as documented.  Host code only: no GPU call."""
import ctypes
import ctypes.util

import numpy as np
import pytest

from pisco4cmb_amd import pypisco

libc = ctypes.CDLL(ctypes.util.find_library("c"))
libc.strtof.restype = ctypes.c_float
libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def strtof_all(path):
    out = []
    for line in open(path):
        out.append([libc.strtof(tok.encode(), None) for tok in line.split()])
    return np.array(out, np.float32)


def test_reads_numpy_savetxt_sky(tmp_path):
    nside = 8
    rng = np.random.default_rng(1)
    data = np.vstack([rng.normal(0, 100, 768), rng.normal(0, 3, 768), rng.normal(0, 3, 768),
                      np.zeros(768)])
    path = tmp_path / "cmb.txt"
    np.savetxt(path, data.T)       # create_cmb.py:96-97
    sky = pypisco.read_sky_txt(path)
    assert sky.get_nside() == nside
    ref = strtof_all(path)
    for k, col in enumerate((sky.sI, sky.sQ, sky.sU, sky.sV)):
        assert np.array_equal(col.view(np.int32), ref[:, k].view(np.int32))
    # within float rounding of numpy's own parse
    assert np.allclose(sky.sI, data[0], rtol=6e-8, atol=0)


def test_large_file_is_split_across_threads(tmp_path):
    rng = np.random.default_rng(2)
    n = 200_000
    cols = [rng.normal(size=n).astype(np.float32) * 10 ** rng.uniform(-30, 30, n).astype(
        np.float32) for _ in range(4)]
    path = tmp_path / "big.txt"
    pypisco.write_txt_columns(path, cols)
    assert path.stat().st_size > (1 << 20)
    back = pypisco.read_txt_columns(path, 4)
    for a, b in zip(cols, back):
        assert np.array_equal(a.view(np.int32), b.view(np.int32))


def test_writer_matches_numpy_savetxt(tmp_path):
    rng = np.random.default_rng(3)
    cols = [rng.normal(size=50).astype(np.float32) for _ in range(4)]
    p1, p2 = tmp_path / "ours.txt", tmp_path / "numpy.txt"
    pypisco.write_txt_columns(p1, cols)
    np.savetxt(p2, np.vstack([c.astype(np.float64) for c in cols]).T)
    assert p1.read_text() == p2.read_text()


def test_tilde_beams_round_trip(tmp_path):
    beam = pypisco.PolBeam(256, 500, 0.0, "p")
    beam.make_unpol_gaussian_elliptical_beams(2.0, 1.8, 10.0)
    path = tmp_path / "hpx_tilde_beams_det_a.txt"
    beam.write_tilde_beams(path, "a")
    b2 = pypisco.PolBeam(256, 500, 0.0, "a")
    b2.read_tilde_beams(path, "a")
    for x, y in zip((beam.Ia, beam.Qa, beam.Ua, beam.Va), (b2.Ia, b2.Qa, b2.Ua, b2.Va)):
        assert np.array_equal(x, y)
    with pytest.raises(ValueError):
        pypisco.PolBeam(256, 499, 0.0, "a").read_tilde_beams(path, "a")


def test_comments_blank_lines_and_errors(tmp_path):
    p = tmp_path / "c.txt"
    p.write_text("# header\n1 2 3 4\n\n   \n5 6 7 8  # trailing\n")
    cols = pypisco.read_txt_columns(p, 4)
    assert [list(c) for c in cols] == [[1, 5], [2, 6], [3, 7], [4, 8]]
    p.write_text("1 2 3\n")
    with pytest.raises(ValueError, match="expected 4"):
        pypisco.read_txt_columns(p, 4)
    p.write_text("1 2 3 4 5\n")
    with pytest.raises(ValueError, match="more than 4"):
        pypisco.read_txt_columns(p, 4)
    with pytest.raises(ValueError, match="cannot open"):
        pypisco.read_txt_columns(tmp_path / "missing.txt", 4)
    p.write_text("1 2 3 4\n" * 10)
    with pytest.raises(ValueError, match="not 12 nside"):
        pypisco.read_sky_txt(p)
