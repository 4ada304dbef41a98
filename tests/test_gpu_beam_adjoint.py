"""The beam adjoint (TOD -> beam maps) against the oracle's FP64 beam adjoint,
relative RMS 1e-6 per column, in each of its forms: visit records written
in band order and summed band by band in LDS as 64-bit fixed-point integers
(adjbeam_records = 1, the default; also with a 1-MiB pool, so that the scan
runs in chunks of 256 samples), the annulus walks over LDS bands
(adjbeam_records = 0; 8 or 64 lanes per sample, small and large work items)
and global atomics (adjbeam_bands = 0).  Shapes: the headline beam (113
bands of records, 81 annulus bands; a sample count that leaves the record
form's last write block of 8 samples partial), a cap reaching the equatorial rings of
a coarse beam map (bands of 300 rows: whole and partial rings), samples near
the poles, small discs, beams of one ring and a last band of one row; the
record form's bits repeat from call to call; and the dot product with the
forward.  The headline shape and the dot
product run on co-polar and cross-polar beams (the gradient's columns do not
depend on the beam values, the dot product does)."""
import math

import numpy as np
import pytest

from pisco4cmb_amd import inputs, pypisco

pytestmark = pytest.mark.gpu
DEG = math.pi / 180


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def conv_for(ns, sky, beam, ra, dec, pa):
    conv = pypisco.GPUConvolver(pypisco.default_run_config())
    conv.configure_execution()
    conv.update_sky(pypisco.Sky(ns, *sky))
    conv.update_beam(beam)
    conv.update_scan(pypisco.Scan(ra.size, ra, dec, pa))
    return conv


DEFAULTS = {"adjbeam_bands": 1, "adjbeam_records": 1, "adjbeam_pool_mb": 8192,
            "adjbeam_chunk": 1024, "adjbeam_lanes": 8}


def forms(chunks):
    """Tuning of each form of the beam adjoint (keys not named: defaults)."""
    out = [{}, {"adjbeam_pool_mb": 1}]
    out += [{"adjbeam_records": 0, "adjbeam_chunk": c} for c in chunks]
    out += [{"adjbeam_records": 0, "adjbeam_chunk": chunks[0], "adjbeam_lanes": 64},
            {"adjbeam_bands": 0}]
    return out


def check_against_oracle(oracle, ns, sky, beam, ra, dec, pa, chunks=(1024,), rows=0):
    rng = np.random.default_rng(81)
    ga, gb = (rng.normal(size=ra.size).astype(np.float32) for _ in range(2))
    beams = list(beam.aBeams) + list(beam.bBeams)
    G = oracle.adjoint_beam(ns, *sky[:3], beam.nside, beams, ra, dec, pa, ga, gb)
    conv = conv_for(ns, sky, beam, ra, dec, pa)
    conv.set_tuning("adjbeam_rows", rows)
    res = []
    for form in forms(chunks):
        for k, v in dict(DEFAULTS, **form).items():
            conv.set_tuning(k, v)
        oa, ob = conv.exec_adjoint_beam(ga, gb)
        for k, (got, col) in enumerate(((oa[0], 0), (oa[1], 1), (oa[2], 2),
                                        (ob[0], 3), (ob[1], 4), (ob[2], 5))):
            assert rel_rms(got, G[:, col]) < 1e-6, (form, k)
        assert np.all(oa[3] == 0) and np.all(ob[3] == 0)
        res.append((form, oa, ob))
    return conv, ga, gb, res


@pytest.mark.parametrize("kind", ["copolar", "crosspolar"])
def test_headline_beam_shape(oracle, kind):
    ns = 512
    sky = inputs.iid_sky(ns, seed=82)
    beam = inputs.make_polbeam(kind, 4096, 5.0, 0.33, 0.05, 0.01, nside_sky=ns)
    # 397 samples: the record form's last write block (8 samples) is partial,
    # and its band-sum parts are cut at multiples of 8 samples
    ra, dec, pa = inputs.ces_scan(397, t0=1300.0)
    conv, ga, gb, res = check_against_oracle(oracle, ns, sky, beam, ra, dec, pa,
                                             chunks=(1024, 7))
    # the record form sums integers: the same bits on every call
    for k, v in DEFAULTS.items():
        conv.set_tuning(k, v)
    again = conv.exec_adjoint_beam(ga, gb)
    for x, y in zip(res[0][1] + res[0][2], again[0] + again[1]):
        assert np.array_equal(x.view(np.int32), y.view(np.int32))
    check_against_oracle(oracle, ns, sky, beam, ra, dec, pa, rows=1000)


def test_coarse_beam_equatorial_rings_and_poles(oracle):
    ns = 32
    sky = inputs.iid_sky(ns, seed=83)
    npix = pypisco.beam_npix_for_radius(32, 60 * DEG)
    b = pypisco.PolBeam(32, npix, 0.02, "p")
    b.make_unpol_gaussian_elliptical_beams(30.0, 24.0, 10.0)
    b.build(ns)
    rng = np.random.default_rng(84)
    n = 300
    ra = rng.uniform(0, 2 * math.pi, n).astype(np.float32)
    dec = np.where(np.arange(n) % 3 == 0, (89.0 * DEG) * np.sign(rng.normal(size=n)),
                   np.arcsin(rng.uniform(-1, 1, n))).astype(np.float32)
    pa = rng.uniform(-math.pi, math.pi, n).astype(np.float32)
    check_against_oracle(oracle, ns, sky, b, ra, dec, pa, chunks=(1024, 5), rows=300)
    check_against_oracle(oracle, ns, sky, b, ra, dec, pa)


def test_small_discs(oracle):
    """Configuration 2's beam shape (about 134 sky pixels per disc)."""
    ns = 256
    sky = inputs.iid_sky(ns, seed=91)
    beam = inputs.crosspolar_polbeam(2048, 1.5, 0.5, 0.05, 0.01, nside_sky=ns)
    ra, dec, pa = inputs.ces_scan(3000, t0=91.0)
    check_against_oracle(oracle, ns, sky, beam, ra, dec, pa)


@pytest.mark.parametrize("kind", ["copolar", "crosspolar"])
def test_dot_product(oracle, kind):
    """<TOD, g> = <beams, A_beam^T g>, with g = TOD + noise so that the inner
    product does not cancel."""
    ns = 256
    sky = inputs.iid_sky(ns, seed=85)
    beam = inputs.make_polbeam(kind, 2048, 1.5, 0.5, 0.05, 0.01, nside_sky=ns)
    ra, dec, pa = inputs.ces_scan(5000, t0=77.0)
    conv = conv_for(ns, sky, beam, ra, dec, pa)
    a = np.zeros(ra.size, np.float32)
    b = np.zeros(ra.size, np.float32)
    conv.exec_convolution(a, b)
    rng = np.random.default_rng(86)
    ga = (a + a.std() * rng.normal(size=ra.size)).astype(np.float32)
    gb = (b + b.std() * rng.normal(size=ra.size)).astype(np.float32)
    oa, ob = conv.exec_adjoint_beam(ga, gb)
    lhs = float(np.dot(a.astype(np.float64), ga) + np.dot(b.astype(np.float64), gb))
    rhs = float(np.sum(np.stack(oa).astype(np.float64) * np.stack(beam.aBeams)) +
                np.sum(np.stack(ob).astype(np.float64) * np.stack(beam.bBeams)))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


@pytest.mark.parametrize("npix,rows", [(4, 0), (60, 0), (301, 300)])
def test_small_beams_and_one_cotangent(oracle, npix, rows):
    """A beam of one ring (4 rows), of a few rings, and a last band of one
    row (301 rows in bands of 300), with g_b absent: against the oracle."""
    ns = 16
    sky = inputs.iid_sky(ns, seed=87)
    b = pypisco.PolBeam(64, npix, 0.0, "a")
    b.make_unpol_gaussian_elliptical_beams(3.0, 2.5, 0.0)
    b.build(ns)
    ra, dec, pa = inputs.raster_scan(500, size_deg=30.0)
    rng = np.random.default_rng(88)
    ga = rng.normal(size=ra.size).astype(np.float32)
    beams = list(b.aBeams) + list(b.bBeams)
    G = oracle.adjoint_beam(ns, *sky[:3], 64, beams, ra, dec, pa, ga, None)
    conv = conv_for(ns, sky, b, ra, dec, pa)
    conv.set_tuning("adjbeam_rows", rows)
    for records, lanes in ((1, 8), (0, 8), (0, 64)):
        conv.set_tuning("adjbeam_records", records)
        conv.set_tuning("adjbeam_lanes", lanes)
        oa, ob = conv.exec_adjoint_beam(ga, None)
        for k in range(3):
            assert rel_rms(oa[k], G[:, k]) < 1e-6
        assert np.all(ob[0] == 0) and np.all(ob[1] == 0) and np.all(ob[2] == 0)


SMALL_ROWS_FIRST = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from pisco4cmb_amd import inputs, pypisco
ns = 128
sky = inputs.iid_sky(ns, seed=89)
ra, dec, pa = inputs.ces_scan(600, t0=40.0)
g = np.random.default_rng(90).normal(size=(2, ra.size)).astype(np.float32)
out = {}
for records, npb_deg, nb, rows in ((0, 3.0, 256, 300), (0, 3.0, 1024, 0), (1, 3.0, 256, 300),
                                   (1, 3.0, 1024, 0)):
    beam = inputs.gaussian_polbeam(nb, npb_deg, 1.0, 0.05, 0.01, nside_sky=ns)
    conv = pypisco.GPUConvolver(pypisco.default_run_config())
    conv.configure_execution()
    conv.update_sky(pypisco.Sky(ns, *sky))
    conv.update_beam(beam)
    conv.update_scan(pypisco.Scan(ra.size, ra, dec, pa))
    conv.set_tuning("adjbeam_rows", rows)
    conv.set_tuning("adjbeam_records", records)
    banded = conv.exec_adjoint_beam(g[0], g[1])
    conv.set_tuning("adjbeam_bands", 0)
    atomic = conv.exec_adjoint_beam(g[0], g[1])
    for x, y in zip(banded[0][:3] + banded[1][:3], atomic[0][:3] + atomic[1][:3]):
        x = x.astype(np.float64); y = y.astype(np.float64)
        assert np.sqrt(np.mean((x - y) ** 2)) <= 1e-6 * np.sqrt(np.mean(y ** 2)), (nb, rows)
    print("rows", rows, "beam rows", beam.nPixels, "ok", flush=True)
    conv.close()
print("SMALL_ROWS_FIRST_OK", flush=True)
"""


def test_small_band_rows_first_in_a_fresh_process():
    """A process whose first banded beam adjoint uses small bands (300 rows)
    and whose second uses the default bands of an 8,580-row beam: the second
    launch needs a larger dynamic-LDS limit than the first set, and must
    raise it (both equal the global-atomic form).  The same for the record
    path's band sums, after the annulus walks."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", SMALL_ROWS_FIRST, root], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "SMALL_ROWS_FIRST_OK" in r.stdout
