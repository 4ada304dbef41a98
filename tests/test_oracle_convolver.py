"""CPU: the oracle's Convolver and PolBeam restatements pinned by physics
invariants (point source, uniform sky, linearity, a/b symmetry, 90 degree
rotation of a circular beam) and closed forms of the Gaussian beam."""
import math

import numpy as np
import pytest

from pisco4cmb_amd import inputs, pypisco

DEG = math.pi / 180


def beams_of(b):
    return list(b.aBeams) + list(b.bBeams)


@pytest.fixture(scope="module")
def setup():
    nside_sky, nside_beam = 64, 512
    beam = inputs.gaussian_polbeam(nside_beam, 6.0, 2.0, 0.0, 0.0, nside_sky=nside_sky)
    rng = np.random.default_rng(0)
    n = 400
    ra = rng.uniform(0, 2 * math.pi, n).astype(np.float32)
    dec = np.arcsin(rng.uniform(-0.9, 0.9, n)).astype(np.float32)
    pa = rng.uniform(-math.pi, math.pi, n).astype(np.float32)
    return nside_sky, nside_beam, beam, ra, dec, pa


def test_point_source_traces_the_beam(oracle, setup):
    """A unit I point source at pixel p0 gives d_a(i) = A0[b] where b is the
    beam pixel of p0 seen from sample i, if p0 is in the disc and b < npix
    (convolver.cpp:98-164); b is computed here from 3-vectors."""
    ns, nb, beam, ra, dec, pa = setup
    npix = 12 * ns * ns
    p0 = 12345
    I = np.zeros(npix, np.float32)
    I[p0] = 1.0
    z = np.zeros(npix, np.float32)
    # point the samples near p0
    th0, ph0 = oracle.pix2ang(ns, np.array([p0], np.int32))
    rng = np.random.default_rng(1)
    n = 300
    ra = (ph0[0] + rng.normal(0, 2 * DEG, n)).astype(np.float32)
    dec = (math.pi / 2 - th0[0] + rng.normal(0, 2 * DEG, n)).astype(np.float32)
    pa = rng.uniform(-math.pi, math.pi, n).astype(np.float32)
    a, b = oracle.convolve(ns, I, z, z, nb, beams_of(beam), ra, dec, pa)
    rho, sigma, _ = oracle.rho_sigma_chi(ra.astype(np.float64), dec.astype(np.float64),
                                         -pa.astype(np.float64), np.full(n, ph0[0]),
                                         np.full(n, math.pi / 2 - th0[0]))
    bp = oracle.ang2pix(nb, rho, sigma)
    inside = (rho <= beam.rhoMax) & (bp < beam.nPixels)
    expect = np.where(inside, beam.aBeams[0][np.minimum(bp, beam.nPixels - 1)], 0.0)
    assert np.mean(np.isclose(a, expect, rtol=1e-6, atol=0)) > 0.99
    assert np.array_equal(a, b)   # identical a/b beams, I only


def test_uniform_sky_sums_the_beam(oracle):
    """I = 1, Q = U = 0: d = sum over the disc of A0[b(p)] ~ (1 + eps) times
    the beam fraction inside the cap, since A0 is normalised so that
    sum A0 * (nside_sky/nside_beam)^2 = 1 + eps (polbeam.cpp:204-268)."""
    ns, nb = 256, 1024
    eps = 0.01
    beam = inputs.gaussian_polbeam(nb, 8.0, 2.0, 0.0, eps, nside_sky=ns)
    npix = 12 * ns * ns
    one = np.ones(npix, np.float32)
    z = np.zeros(npix, np.float32)
    ra, dec, pa = inputs.raster_scan(200, size_deg=40)
    a, b = oracle.convolve(ns, one, z, z, nb, beams_of(beam), ra, dec, pa)
    assert np.all(np.abs(a - (1 + eps)) < 0.01)
    assert np.array_equal(a, b)


def test_linearity(oracle, setup):
    ns, nb, beam, ra, dec, pa = setup
    s1 = inputs.iid_sky(ns, 1)
    s2 = inputs.iid_sky(ns, 2)
    comb = [(0.5 * x + 2.0 * y).astype(np.float32) for x, y in zip(s1[:3], s2[:3])]
    a1, b1 = oracle.convolve(ns, *s1[:3], nb, beams_of(beam), ra, dec, pa, precision64=True)
    a2, b2 = oracle.convolve(ns, *s2[:3], nb, beams_of(beam), ra, dec, pa, precision64=True)
    a3, b3 = oracle.convolve(ns, *comb, nb, beams_of(beam), ra, dec, pa, precision64=True)
    assert np.allclose(a3, 0.5 * a1 + 2.0 * a2, rtol=1e-6, atol=1e-6 * np.abs(a3).max())
    assert np.allclose(b3, 0.5 * b1 + 2.0 * b2, rtol=1e-6, atol=1e-6 * np.abs(b3).max())


def test_pair_polarisation_antisymmetry(oracle, setup):
    """Identical a/b beams and no leakage: d_a + d_b = 2 x (I term) and
    d_a - d_b = 2 x (polarised term) (convolver.cpp:157-164)."""
    ns, nb, beam, ra, dec, pa = setup
    I, Q, U, _ = inputs.iid_sky(ns, 5)
    z = np.zeros_like(I)
    a, b = oracle.convolve(ns, I, Q, U, nb, beams_of(beam), ra, dec, pa, precision64=True)
    ai, _ = oracle.convolve(ns, I, z, z, nb, beams_of(beam), ra, dec, pa, precision64=True)
    ap, bp = oracle.convolve(ns, z, Q, U, nb, beams_of(beam), ra, dec, pa, precision64=True)
    assert np.allclose(a + b, 2 * ai, atol=1e-9 * np.abs(ai).max())
    assert np.allclose(ap, -bp, atol=1e-12)


def test_rotation_by_90_degrees_flips_polarisation(oracle):
    """A circular beam is invariant under sigma -> sigma + pi/2 (the polar
    cap has four-fold symmetry), while chi -> chi - pi/2 flips cos/sin(2chi):
    a Q/U-only sky gives d(pa + 90 deg) = -d(pa)."""
    ns, nb = 64, 512
    beam = inputs.gaussian_polbeam(nb, 6.0, 2.0, 0.0, 0.0, nside_sky=ns)
    z, Q, U = [x for x in inputs.iid_sky(ns, 8)[:3]]
    z = np.zeros_like(z)
    rng = np.random.default_rng(3)
    n = 300
    ra = rng.uniform(0, 2 * math.pi, n).astype(np.float32)
    dec = np.arcsin(rng.uniform(-0.8, 0.8, n)).astype(np.float32)
    pa = rng.uniform(0, 1, n).astype(np.float32)
    a0, _ = oracle.convolve(ns, z, Q, U, nb, beams_of(beam), ra, dec, pa, precision64=True)
    a1, _ = oracle.convolve(ns, z, Q, U, nb, beams_of(beam), ra, dec,
                            (pa + np.float32(math.pi / 2)).astype(np.float32), precision64=True)
    close = np.isclose(a1, -a0, rtol=1e-4, atol=1e-4 * np.abs(a0).max())
    assert np.mean(close) > 0.95


def test_polflag_validation(oracle, setup):
    ns, nb, beam, ra, dec, pa = setup
    I, Q, U, _ = inputs.iid_sky(ns, 5)
    with pytest.raises(ValueError):
        oracle.convolve(ns, I, Q, U, nb, beams_of(beam), ra, dec, pa, polflag="x")
    a, b = oracle.convolve(ns, I, Q, U, nb, beams_of(beam), ra[:50], dec[:50], pa[:50], polflag="a")
    assert np.all(b == 0) and np.any(a != 0)


def test_gaussian_beam_closed_form(oracle):
    """tilde I of make_unpol_gaussian_elliptical_beams at each beam pixel
    centre equals exp(-(a cos^2 s + 2b cos s sin s + c sin^2 s) rho^2)
    (polbeam.cpp:332-357), U = V = 0, Q = I; the built beams are
    (I + eps I) / norm with norm = sum I * (ns/nb)^2."""
    nb, ns, eps = 256, 64, 0.02
    npix = pypisco.beam_npix_for_radius(nb, 10 * DEG)
    fx, fy, phi0 = 3.0, 2.5, 30.0
    out, tilde, f = oracle.gaussian_polbeam(nb, npix, fx, fy, phi0, eps, "p", ns)
    th, ph = oracle.pix2ang(nb, np.arange(npix, dtype=np.int32))
    sx, sy, p0 = fx * DEG / 2.35482, fy * DEG / 2.35482, phi0 * DEG
    a = math.cos(p0) ** 2 / (2 * sx * sx) + math.sin(p0) ** 2 / (2 * sy * sy)
    b = -math.sin(2 * p0) / (4 * sx * sx) + math.sin(2 * p0) / (4 * sy * sy)
    c = math.sin(p0) ** 2 / (2 * sx * sx) + math.cos(p0) ** 2 / (2 * sy * sy)
    expect = np.exp(-(a * np.cos(ph) ** 2 + 2 * b * np.cos(ph) * np.sin(ph)
                      + c * np.sin(ph) ** 2) * th ** 2)
    assert np.allclose(tilde[0], expect, rtol=1e-6, atol=1e-30)
    assert np.array_equal(tilde[0], tilde[1]) and np.all(tilde[2] == 0) and np.all(tilde[3] == 0)
    norm = np.sum(tilde[0].astype(np.float64)) * (ns / nb) ** 2
    assert np.allclose(out[0], (1 + eps) * tilde[0] / norm, rtol=1e-6)
    assert np.allclose(out[1], (1 - eps) * tilde[1] / norm, rtol=1e-6)
    assert float(np.sum(out[0].astype(np.float64)) * (ns / nb) ** 2) == pytest.approx(1 + eps,
                                                                                      rel=1e-6)


def test_product_polbeam_matches_oracle_bit_exact(oracle):
    """pypisco.PolBeam (host code of libpisco) builds the same beams as the
    oracle's PolBeam restatement, bit for bit."""
    nb, ns = 1024, 128
    npix = pypisco.beam_npix_for_radius(nb, 2.5 * DEG)
    bm = pypisco.PolBeam(nb, npix, 0.01, "p")
    bm.make_unpol_gaussian_elliptical_beams(1.0, 0.95, 0.0)
    bm.build(ns)
    out, _, _ = oracle.gaussian_polbeam(nb, npix, 1.0, 0.95, 0.0, 0.01, "p", ns)
    for k in range(4):
        assert np.array_equal(bm.aBeams[k], out[k])
        assert np.array_equal(bm.bBeams[k], out[4 + k])
    assert bm.get_rho_max() == oracle.beam_rho_max(nb, npix)


def test_shared_libm_vs_glibc_index_diagnostic(oracle, oracle_glibc):
    """How far the shared FP64 library moves the oracle from a glibc build:
    beam-pixel indices at the headline geometry (NSIDE 512, 5 deg cap,
    beam NSIDE 4096) over ~6e6 visits."""
    rng = np.random.default_rng(11)
    n = 1000
    ra = rng.uniform(0, 2 * math.pi, n).astype(np.float32)
    dec = np.arcsin(rng.uniform(-1, 1, n)).astype(np.float32)
    pa = rng.uniform(-math.pi, math.pi, n).astype(np.float32)
    off, pix = oracle.query_disc_many(512, ra, dec, 5 * DEG)
    off2, pix2 = oracle_glibc.query_disc_many(512, ra, dec, 5 * DEG)
    assert np.array_equal(off, off2) and np.array_equal(pix, pix2)
    b1, _ = oracle.match(512, 4096, ra, dec, pa, off, pix)
    b2, _ = oracle_glibc.match(512, 4096, ra, dec, pa, off, pix)
    mism = int(np.sum(b1 != b2))
    print("beam-pixel mismatches shared-libm vs glibc: %d of %d" % (mism, b1.size))
    assert mism <= 1e-5 * b1.size


def test_adjoints_satisfy_the_dot_product_identity(oracle, setup):
    """The oracle's adjoints are the transposes of its forward:
    <F x, g> = <x, F^T g> for the sky (I,Q,U) and for the beams (A0..A2,
    B0..B2; the forward does not read V).  The adjoints use cos/sin(2 chi)
    rounded to float as the reference's forward does, the FP64 forward keeps
    them in double: the two sides differ by that rounding (~1e-8 relative),
    while a dropped term, a wrong sign or a transposed operand would move
    them by O(0.1-1).  Tolerance 1e-6 relative."""
    nside_sky, nside_beam, beam, ra, dec, pa = setup
    rng = np.random.default_rng(11)
    npix = 12 * nside_sky ** 2
    I, Q, U = (rng.normal(size=npix).astype(np.float32) for _ in range(3))
    ga, gb = (rng.normal(size=ra.size).astype(np.float32) for _ in range(2))
    beams = beams_of(beam)
    da, db = oracle.convolve(nside_sky, I, Q, U, nside_beam, beams, ra, dec, pa,
                             precision64=True)
    lhs = float(np.dot(da, ga) + np.dot(db, gb))
    gI, gQ, gU = oracle.adjoint_map(nside_sky, nside_beam, beams, ra, dec, pa, ga, gb)
    rhs_sky = float(np.dot(gI, I) + np.dot(gQ, Q) + np.dot(gU, U))
    assert abs(lhs - rhs_sky) <= 1e-6 * abs(lhs)
    G = oracle.adjoint_beam(nside_sky, I, Q, U, nside_beam, beams, ra, dec, pa, ga, gb)
    used = [beams[0], beams[1], beams[2], beams[4], beams[5], beams[6]]
    rhs_beam = float(sum(np.dot(G[:, k], used[k].astype(np.float64)) for k in range(6)))
    assert abs(lhs - rhs_beam) <= 1e-6 * abs(lhs)
