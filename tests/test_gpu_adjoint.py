"""The LDS-privatised sky adjoint: against the FP64 oracle, against the
direct-atomic mode, with windows forced to overflow (fallback path), and the
dot-product identity at the headline geometry (NSIDE 512, 5 deg cap).
Every case runs on a co-polar Gaussian beam and on a cross-polar beam (all
beam maps non-zero, so the A2/B2 terms of the adjoint are compared)."""
import math

import numpy as np
import pytest

from pisco4cmb_amd import inputs, pypisco

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def conv_for(nside_sky, sky, beam, ra, dec, pa):
    conv = pypisco.GPUConvolver(pypisco.default_run_config())
    conv.configure_execution()
    conv.update_sky(pypisco.Sky(nside_sky, *sky))
    conv.update_beam(beam)
    conv.update_scan(pypisco.Scan(ra.size, ra, dec, pa))
    return conv


KINDS = ["copolar", "crosspolar"]


@pytest.fixture(scope="module", params=KINDS)
def c3_small(request):
    nside_sky = 512
    sky = inputs.iid_sky(nside_sky, seed=3)
    beam = inputs.make_polbeam(request.param, 4096, 5.0, 0.33, 0.02, 0.01, nside_sky=nside_sky)
    ra, dec, pa = inputs.ces_scan(1500, fs=100.0)
    return nside_sky, sky, beam, ra, dec, pa


def test_adjoint_lds_vs_oracle_and_direct(c3_small, oracle):
    nside_sky, sky, beam, ra, dec, pa = c3_small
    rng = np.random.default_rng(1)
    ga = rng.normal(size=ra.size).astype(np.float32)
    gb = rng.normal(size=ra.size).astype(np.float32)
    conv = conv_for(nside_sky, sky, beam, ra, dec, pa)
    lds = conv.exec_adjoint_map(ga, gb)
    conv.set_tuning("adjmap_chunk", 0)
    direct = conv.exec_adjoint_map(ga, gb)
    beams = list(beam.aBeams) + list(beam.bBeams)
    ref = oracle.adjoint_map(nside_sky, 4096, beams, ra, dec, pa, ga, gb)
    for k in range(3):
        assert rel_rms(lds[k], ref[k]) < 1e-6
        assert rel_rms(direct[k], ref[k]) < 1e-6
    # every pixel with a significant oracle gradient is covered by a window
    sig = np.abs(ref[0]) > 1e-6 * np.abs(ref[0]).max()
    assert np.all(lds[0][sig] != 0)


@pytest.mark.parametrize("chunk,cap", [(1, 6144), (7, 6144), (256, 6144), (64, 500)])
def test_adjoint_lds_shapes_and_fallback(c3_small, oracle, chunk, cap):
    """Chunk sizes (incl. a ragged last chunk) and a window too small for any
    chunk (every chunk takes the direct-atomic fallback)."""
    nside_sky, sky, beam, ra, dec, pa = c3_small
    n = 700
    ra, dec, pa = ra[:n], dec[:n], pa[:n]
    rng = np.random.default_rng(2)
    ga = rng.normal(size=n).astype(np.float32)
    conv = conv_for(nside_sky, sky, beam, ra, dec, pa)
    conv.set_tuning("adjmap_chunk", chunk)
    conv.set_tuning("adjmap_cap", cap)
    got = conv.exec_adjoint_map(ga, None)
    beams = list(beam.aBeams) + list(beam.bBeams)
    ref = oracle.adjoint_map(nside_sky, 4096, beams, ra, dec, pa, ga, None)
    for k in range(3):
        assert rel_rms(got[k], ref[k]) < 1e-6


@pytest.mark.parametrize("kind", KINDS)
def test_auto_chunk_small_discs(oracle, kind):
    """Small discs (sky NSIDE 256, cap 1.5 deg): the default chunk is chosen
    from the disc size (128 samples per window here); the result equals the
    oracle and the 16-sample windows within relative RMS 1e-6."""
    nside_sky = 256
    sky = inputs.iid_sky(nside_sky, seed=31)
    beam = inputs.make_polbeam(kind, 2048, 1.5, 0.5, 0.03, 0.01, nside_sky=nside_sky)
    ra, dec, pa = inputs.ces_scan(20000, t0=300.0)
    rng = np.random.default_rng(32)
    ga, gb = (rng.normal(size=ra.size).astype(np.float32) for _ in range(2))
    conv = conv_for(nside_sky, sky, beam, ra, dec, pa)
    assert conv.get_tuning("adjmap_chunk") == -1
    auto = conv.exec_adjoint_map(ga, gb)
    conv.set_tuning("adjmap_chunk", 16)
    c16 = conv.exec_adjoint_map(ga, gb)
    beams = list(beam.aBeams) + list(beam.bBeams)
    ref = oracle.adjoint_map(nside_sky, 2048, beams, ra, dec, pa, ga, gb)
    for k in range(3):
        assert rel_rms(auto[k], ref[k]) < 1e-6
        assert rel_rms(auto[k], c16[k]) < 1e-6


@pytest.mark.parametrize("kind", KINDS)
def test_adjoint_polar_and_wraparound(oracle, kind):
    """Pointings through both poles and across ra = 0 (ring windows that wrap
    and whole polar rings)."""
    nside_sky = 128
    sky = inputs.iid_sky(nside_sky, seed=4)
    beam = inputs.make_polbeam(kind, 1024, 4.0, 1.5, 0.0, 0.0, nside_sky=nside_sky)
    n = 2000
    t = np.linspace(0, 2 * math.pi, n)
    dec = (math.pi / 2 * np.sin(t)).astype(np.float32)      # meridian pole to pole
    ra = np.where(np.cos(t) >= 0, 0.02, math.pi + 0.02).astype(np.float32)
    ra[::3] = np.float32(2 * math.pi - 0.01)
    pa = (0.3 * t).astype(np.float32)
    rng = np.random.default_rng(5)
    ga = rng.normal(size=n).astype(np.float32)
    gb = rng.normal(size=n).astype(np.float32)
    conv = conv_for(nside_sky, sky, beam, ra, dec, pa)
    got = conv.exec_adjoint_map(ga, gb)
    beams = list(beam.aBeams) + list(beam.bBeams)
    ref = oracle.adjoint_map(nside_sky, 1024, beams, ra, dec, pa, ga, gb)
    for k in range(3):
        assert rel_rms(got[k], ref[k]) < 1e-6


def test_dot_product_headline_geometry(c3_small):
    nside_sky, sky, beam, ra, dec, pa = c3_small
    rng = np.random.default_rng(6)
    ga = rng.normal(size=ra.size).astype(np.float32)
    gb = rng.normal(size=ra.size).astype(np.float32)
    conv = conv_for(nside_sky, sky, beam, ra, dec, pa)
    a = np.zeros(ra.size, np.float32)
    b = np.zeros(ra.size, np.float32)
    conv.exec_convolution(a, b)
    gI, gQ, gU = conv.exec_adjoint_map(ga, gb)
    lhs = float(np.dot(a.astype(np.float64), ga) + np.dot(b.astype(np.float64), gb))
    rhs = sum(float(np.dot(sky[k].astype(np.float64), g.astype(np.float64)))
              for k, g in enumerate((gI, gQ, gU)))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
