"""CPU: the oracle's HEALPix RING restatement (healpix_cxx semantics) pinned
by closed forms, round trips, brute force and equal-area sampling."""
import math

import numpy as np
import pytest


@pytest.mark.parametrize("nside", [1, 2, 3, 4, 8, 16, 64, 256])
def test_roundtrip_all_pixels(oracle, nside):
    """pix2ang then ang2pix returns every pixel (the semantics of the
    reference test tests/cudaHealpix/test_pix2ang.cu:80-112)."""
    npix = 12 * nside * nside
    pix = np.arange(npix, dtype=np.int32)
    th, ph = oracle.pix2ang(nside, pix)
    assert np.all((th > 0) & (th < math.pi)) and np.all((ph >= 0) & (ph < 2 * math.pi))
    assert np.array_equal(oracle.ang2pix(nside, th, ph), pix)


@pytest.mark.parametrize("nside", [4096, 8192])
def test_roundtrip_large_nside_sampled(oracle, nside):
    rng = np.random.default_rng(nside)
    pix = rng.integers(0, 12 * nside * nside, 300000).astype(np.int32)
    th, ph = oracle.pix2ang(nside, pix)
    assert np.array_equal(oracle.ang2pix(nside, th, ph), pix)


@pytest.mark.parametrize("nside", [1, 2, 4, 16, 128, 1000])
def test_ring_closed_forms(oracle, nside):
    """Ring z: cap 1 - i^2/(3N^2), belt (2N - i) 2/(3N); 4i pixels per cap
    ring, 4N per belt ring; phi of the first pixel pi/(4i) in cap rings and
    pi/(4N) * (1 or 1/2) in the belt (Gorski et al. 2005, eqs. 4-9)."""
    N = nside
    rings = np.arange(1, 4 * N)
    starts = np.zeros(rings.size + 1, np.int64)
    counts = np.where(rings < N, 4 * rings, np.where(rings > 3 * N, 4 * (4 * N - rings), 4 * N))
    starts[1:] = np.cumsum(counts)
    assert starts[-1] == 12 * N * N
    pix = starts[:-1].astype(np.int32)
    th, ph = oracle.pix2ang(N, pix)
    i = rings
    z_expect = np.where(i < N, 1 - i * i / (3.0 * N * N),
                        np.where(i > 3 * N, -(1 - (4 * N - i) ** 2 / (3.0 * N * N)),
                                 (2 * N - i) * 2.0 / (3 * N)))
    assert np.allclose(np.cos(th), z_expect, rtol=0, atol=4e-15)
    k = np.where(i < N, i, np.where(i > 3 * N, 4 * N - i, N))
    shifted = (i < N) | (i > 3 * N) | (((i - N) & 1) == 0)
    ph_expect = np.where(shifted, math.pi / (4 * k), 0.0)
    ph_expect = np.where((i >= N) & (i <= 3 * N) & ~shifted, 0.0, ph_expect)
    assert np.allclose(ph, ph_expect, rtol=0, atol=1e-15)
    # every pixel maps to its own ring's z
    allpix = np.arange(12 * N * N, dtype=np.int32) if N <= 128 else pix
    th2, _ = oracle.pix2ang(N, allpix)
    assert np.unique(np.round(np.cos(th2), 12)).size == 4 * N - 1


@pytest.mark.parametrize("nside", [1, 2, 4])
def test_equal_area_sampling(oracle, nside):
    """Uniform points on the sphere fall into pixels in equal proportions
    (a wrong boundary formula breaks the equal-area property), and each
    pixel centre lies in its own pixel."""
    rng = np.random.default_rng(0)
    n = 3_000_000
    th = np.arccos(rng.uniform(-1, 1, n))
    ph = rng.uniform(0, 2 * math.pi, n)
    pix = oracle.ang2pix(nside, th, ph)
    npix = 12 * nside * nside
    counts = np.bincount(pix, minlength=npix)
    expect = n / npix
    assert np.all(np.abs(counts - expect) < 6 * math.sqrt(expect))


def test_known_values(oracle):
    """Published NSIDE 1 geometry: pixel 0 at (acos(2/3), pi/4), pixel 4 at
    (pi/2, 0); the north pole lies in pixel 0..3 by phi quadrant."""
    th, ph = oracle.pix2ang(1, np.array([0, 4, 11], np.int32))
    assert th[0] == pytest.approx(math.acos(2 / 3), abs=1e-15)
    assert ph[0] == pytest.approx(math.pi / 4, abs=1e-15)
    assert th[1] == pytest.approx(math.pi / 2, abs=1e-15) and ph[1] == 0.0
    assert th[2] == pytest.approx(math.pi - math.acos(2 / 3), abs=1e-15)
    q = oracle.ang2pix(1, np.full(4, 1e-3), np.array([0.1, 1.7, 3.3, 5.0]))
    assert list(q) == [0, 1, 2, 3]
    # negative and > 2 pi phi wrap (non-negative fmodulo of healpix_cxx)
    a = oracle.ang2pix(64, np.array([1.0, 1.0, 1.0]), np.array([0.3, 0.3 - 2 * math.pi, 0.3 + 4 * math.pi]))
    assert a[0] == a[1] == a[2]


def _brute_disc(pixvec, theta, phi, radius):
    v0 = np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi),
                   math.cos(theta)])
    d = np.arccos(np.clip(pixvec @ v0, -1, 1))
    return d, set(np.nonzero(d <= radius)[0].tolist())


@pytest.mark.parametrize("nside", [1, 4, 16, 64])
def test_query_disc_brute_force(oracle, nside):
    """Non-inclusive query_disc == all pixels whose centre is within the
    radius (ties within 1e-9 rad of the boundary are ignored); ranges are
    sorted and disjoint."""
    npix = 12 * nside * nside
    th, ph = oracle.pix2ang(nside, np.arange(npix, dtype=np.int32))
    vec = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    rng = np.random.default_rng(nside)
    cases = [(0.0, 0.0, 0.3), (math.pi, 1.0, 0.4), (1e-9, 2.0, 0.5), (math.pi / 2, 0.0, 0.2),
             (math.pi / 2, 2 * math.pi - 1e-3, 0.2), (1.0, -0.5, 1.2), (2.0, 7.0, 2.9),
             (0.5, 0.5, 3.2), (0.7, 0.1, 0.0)]
    cases += [(math.acos(rng.uniform(-1, 1)), rng.uniform(-1, 7), rng.uniform(0.01, 1.5))
              for _ in range(150)]
    for t0, p0, r in cases:
        rr = oracle.query_disc(nside, t0, p0, r)
        got = set()
        for a, b in rr:
            got |= set(range(a, b))
        assert np.all(rr[:, 0] < rr[:, 1])
        assert np.all(rr[1:, 0] > rr[:-1, 1]) if len(rr) > 1 else True
        d, brute = _brute_disc(vec, t0, p0, r)
        for p in got ^ brute:
            assert abs(d[p] - r) < 1e-9, (t0, p0, r, p)


def test_query_disc_area(oracle):
    """|D| tracks the cap area (1 - cos r)/2 * 12 N^2 at NSIDE 512, 5 deg."""
    r = math.radians(5.0)
    rr = oracle.query_disc(512, 1.0, 2.0, r)
    n = int(np.sum(rr[:, 1] - rr[:, 0]))
    expect = (1 - math.cos(r)) / 2 * 12 * 512 ** 2
    assert abs(n - expect) / expect < 0.02


@pytest.mark.parametrize("nside", [1, 2, 16, 256])
def test_get_interpol_properties(oracle, nside):
    """Bilinear weights sum to one, are non-negative, reproduce the pixel
    centre exactly (weight 1 on the pixel) and interpolate a function of z
    linearly within a ring pair."""
    rng = np.random.default_rng(1)
    th = np.arccos(rng.uniform(-1, 1, 20000))
    ph = rng.uniform(0, 2 * math.pi, 20000)
    pix, w = oracle.get_interpol(nside, th, ph)
    assert np.allclose(w.sum(1), 1.0, atol=1e-12)
    assert np.all(w >= -1e-12)
    assert np.all((pix >= 0) & (pix < 12 * nside * nside))
    cpix = np.arange(0, 12 * nside * nside, max(1, nside * nside // 7), dtype=np.int32)
    cth, cph = oracle.pix2ang(nside, cpix)
    p2, w2 = oracle.get_interpol(nside, cth, cph)
    for i in range(cpix.size):
        wsum = sum(w2[i, k] for k in range(4) if p2[i, k] == cpix[i])
        assert wsum == pytest.approx(1.0, abs=1e-9)


def test_cap_division_identity(oracle):
    """The kernels' division-free phi in the polar caps equals the IEEE
    quotient for every cap pixel up to NSIDE 8192 (exhaustive)."""
    assert oracle.L.or_check_cap_division(8192) == 0
