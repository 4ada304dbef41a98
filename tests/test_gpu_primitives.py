"""Device HEALPix, query_disc, the sky->beam matching and the FP64 library:
bit-exact against the CPU oracle (the semantics of tests/cudaHealpix/
test_pix2ang.cu:80-114 and tests/cudaConvolution/
test_sky_to_beam_matching_kernel.cu:97-283, with exact equality instead of
1e-10 / 1e-6 tolerances)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_libm_bit_exact(gpu, oracle):
    rng = np.random.default_rng(11)
    n = 400000
    x = np.concatenate([rng.uniform(-13, 13, n), rng.normal(0, 1e-3, n // 4),
                        (rng.integers(-8, 9, n // 4) * (math.pi / 4)) + rng.normal(0, 1e-12, n // 4)])
    for op, name in ((0, "sin"), (1, "cos")):
        d = gpu.libm(op, x)
        h = oracle.vec(name, x)
        assert np.array_equal(d.view(np.int64), h.view(np.int64)), name
    xa = np.concatenate([rng.uniform(-1, 1, n), 1 - 10 ** rng.uniform(-17, 0, n // 2),
                         -1 + 10 ** rng.uniform(-17, 0, n // 4), [1.0, -1.0, 0.5, -0.5, 0.0]])
    d = gpu.libm(2, xa)
    h = oracle.vec("acos", xa)
    assert np.array_equal(d.view(np.int64), h.view(np.int64))
    y = rng.normal(size=n) * 10 ** rng.uniform(-8, 3, n)
    x2 = rng.normal(size=n) * 10 ** rng.uniform(-8, 3, n)
    y[:10] = 0.0
    x2[10:20] = 0.0
    d = gpu.libm(3, y, x2)
    h = oracle.vec("atan2", y, x2)
    assert np.array_equal(d.view(np.int64), h.view(np.int64))


def test_sqrt_div_fmod_correctly_rounded(gpu):
    """The index path relies on IEEE sqrt, division and an exact fmod on the
    device; compare with the host's (IEEE) results."""
    rng = np.random.default_rng(12)
    n = 400000
    x = 10 ** rng.uniform(-20, 20, n)
    assert np.array_equal(gpu.libm(4, x).view(np.int64), np.sqrt(x).view(np.int64))
    y = 10 ** rng.uniform(-5, 5, n)
    assert np.array_equal(gpu.libm(5, x, y).view(np.int64), (x / y).view(np.int64))
    v = rng.uniform(-100, 100, n)
    for m in (4.0, 2 * math.pi):
        got = gpu.libm(6, v, np.full(n, m))
        ref = np.fmod(v, m)
        ref = np.where(v >= 0, np.where(v < m, v, ref), np.where(ref + m == m, 0.0, ref + m))
        assert np.array_equal(got.view(np.int64), ref.view(np.int64)), m


def test_fixed_point_conversion_is_rint(gpu):
    """The integer sums' conversion (one fma with 1.5 * 2^52 and an integer
    subtract) equals rint(v * s) as int64, ties to even, for float v and
    power-of-two scales s with |v s| < 2^50 (the bound the scales keep)."""
    rng = np.random.default_rng(13)
    n = 200000
    k = rng.integers(-20, 60, n)
    s = np.ldexp(1.0, k)
    v = (rng.normal(size=n) * np.ldexp(1.0, rng.integers(-30, 10, n))).astype(np.float32)
    # keep |v s| < 2^50
    v = np.where(np.abs(v.astype(np.float64) * s) < 2.0 ** 50, v, np.float32(0.75))
    s = np.where(np.abs(v.astype(np.float64) * s) < 2.0 ** 50, s, 1.0)
    # exact ties of both signs, zeros and the ends of the range
    m = rng.integers(-2 ** 20, 2 ** 20, 2000)
    tv = ((m + 0.5) / 2.0).astype(np.float32)
    v = np.concatenate([v, tv, -tv, np.float32([0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 1.5, -1.5])])
    s = np.concatenate([s, np.full(2 * m.size, 2.0), np.ones(8)])
    big = np.float32([2.0 ** 23 - 1, -(2.0 ** 23 - 1)])
    v = np.concatenate([v, big])
    s = np.concatenate([s, np.full(2, 2.0 ** 26)])
    got = gpu.libm(8, v.astype(np.float64), s).view(np.int64)
    ref = np.rint(v.astype(np.float64) * s).astype(np.int64)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16, 64, 128, 256, 512, 1024, 2048])
def test_pix2ang_ang2pix_all_pixels(gpu, oracle, nside):
    npix = 12 * nside * nside
    pix = np.arange(npix, dtype=np.int32)
    th_d, ph_d = gpu.pix2ang(nside, pix)
    th_h, ph_h = oracle.pix2ang(nside, pix)
    assert np.array_equal(th_d.view(np.int64), th_h.view(np.int64))
    assert np.array_equal(ph_d.view(np.int64), ph_h.view(np.int64))
    back = gpu.ang2pix(nside, th_h, ph_h)
    assert np.array_equal(back, pix)


@pytest.mark.parametrize("nside", [4096, 8192])
def test_pix2ang_ang2pix_large_nside_sampled(gpu, oracle, nside):
    rng = np.random.default_rng(nside)
    pix = np.sort(rng.integers(0, 12 * nside * nside, 2_000_000)).astype(np.int32)
    pix[:4] = [0, 1, 12 * nside * nside - 2, 12 * nside * nside - 1]
    th_d, ph_d = gpu.pix2ang(nside, pix)
    th_h, ph_h = oracle.pix2ang(nside, pix)
    assert np.array_equal(th_d, th_h) and np.array_equal(ph_d, ph_h)
    assert np.array_equal(gpu.ang2pix(nside, th_h, ph_h), pix)


@pytest.mark.parametrize("nside", [64, 1024, 4096])
def test_ang2pix_random_angles(gpu, oracle, nside):
    rng = np.random.default_rng(7 + nside)
    n = 2_000_000
    th = np.arccos(rng.uniform(-1, 1, n))
    th[: n // 10] = rng.uniform(0, 0.02, n // 10)          # the theta < 0.01 rule
    th[n // 10: n // 5] = math.pi - rng.uniform(0, 0.02, n // 10)
    ph = rng.uniform(-7, 14, n)
    assert np.array_equal(gpu.ang2pix(nside, th, ph), oracle.ang2pix(nside, th, ph))


@pytest.mark.parametrize("nside,radius_deg", [(8, 30.0), (64, 5.0), (128, 2.5), (256, 1.5),
                                              (512, 5.0), (2048, 1.0), (16, 200.0)])
def test_query_disc_lists(gpu, oracle, nside, radius_deg):
    rng = np.random.default_rng(nside)
    n = 3000 if nside <= 512 else 500
    ra = rng.uniform(0, 2 * math.pi, n).astype(np.float32)
    dec = np.arcsin(rng.uniform(-1, 1, n)).astype(np.float32)
    dec[:6] = [math.pi / 2, -math.pi / 2, 1.5, -1.5, 0.0, 1.5707]
    ra[6:10] = [0.0, -0.1, 6.283, 7.0]
    r = math.radians(radius_deg)
    off_h, pix_h = oracle.query_disc_many(nside, ra, dec, r)
    off_d, pix_d = gpu.query_disc(nside, ra, dec, r)
    assert np.array_equal(off_h, off_d)
    assert np.array_equal(pix_h, pix_d)


@pytest.mark.parametrize("nside_sky,nside_beam,cap_deg", [(256, 256, 5.0), (128, 1024, 2.5),
                                                          (512, 4096, 5.0), (256, 2048, 1.5)])
def test_match_bit_exact(gpu, oracle, nside_sky, nside_beam, cap_deg):
    rng = np.random.default_rng(3)
    n = 1000 if nside_sky < 512 else 200
    ra = rng.uniform(0, 2 * math.pi, n).astype(np.float32)
    dec = (math.pi / 2 - rng.uniform(0, math.pi, n)).astype(np.float32)
    pa = rng.uniform(0, math.pi, n).astype(np.float32)
    off, pix = oracle.query_disc_many(nside_sky, ra, dec, math.radians(cap_deg))
    b_h, chi_h = oracle.match(nside_sky, nside_beam, ra, dec, pa, off, pix)
    b_d, chi_d = gpu.match(nside_sky, nside_beam, ra, dec, pa, off, pix)
    assert np.array_equal(b_h, b_d)
    assert np.array_equal(chi_h.view(np.int64), chi_d.view(np.int64))


@pytest.mark.parametrize("nside", [1, 2, 16, 512])
def test_get_interpol(gpu, oracle, nside):
    rng = np.random.default_rng(5)
    n = 200000
    th = np.arccos(rng.uniform(-1, 1, n))
    th[:100] = rng.uniform(0, 0.05, 100)
    ph = rng.uniform(0, 2 * math.pi, n)
    p_d, w_d = gpu.get_interpol(nside, th, ph)
    p_h, w_h = oracle.get_interpol(nside, th, ph)
    assert np.array_equal(p_d, p_h)
    assert np.array_equal(w_d.view(np.int64), w_h.view(np.int64))


@pytest.mark.parametrize("nside_sky,nside_beam,cap_deg", [(512, 4096, 5.0), (256, 2048, 1.5),
                                                          (128, 1024, 2.5), (2048, 8192, 0.3),
                                                          (64, 64, 40.0)])
def test_match_fast_path_bit_exact(gpu, oracle, nside_sky, nside_beam, cap_deg):
    """The mapping the convolution kernels use (single precision, certified
    by error margins, exact FP64 where a floor could flip) returns exactly
    the oracle's beam pixels; cos/sin(2 chi) agree with the oracle's to the
    single-precision error bound (9.2e-7 measured in float32 emulation;
    tolerance 2e-6).  The single-precision form alone declines (defers to
    the exact path) on a small fraction of visits, and every pixel it does
    certify is the oracle's.  The same for the forward's double-precision
    form with a float atan2 for alpha (floors certified against its 1.5e-6
    rad bound, declined visits mapped by the FP64-alpha form)."""
    rng = np.random.default_rng(17)
    n = 600 if nside_sky <= 512 else 3000
    ra = rng.uniform(0, 2 * math.pi, n).astype(np.float32)
    dec = (math.pi / 2 - np.arccos(rng.uniform(-1, 1, n))).astype(np.float32)
    dec[:3] = [1.5707, -1.5707, 1.5]
    pa = rng.uniform(-math.pi, math.pi, n).astype(np.float32)
    off, pix = oracle.query_disc_many(nside_sky, ra, dec, math.radians(cap_deg))
    b_h, chi_h = oracle.match(nside_sky, nside_beam, ra, dec, pa, off, pix)
    b_f, c2, s2 = gpu.match_fast(nside_sky, nside_beam, ra, dec, pa, off, pix)
    assert np.array_equal(b_h, b_f)
    c2_ref = np.cos(2 * chi_h).astype(np.float32)
    s2_ref = np.sin(2 * chi_h).astype(np.float32)
    assert np.max(np.abs(c2 - c2_ref)) <= 2e-6
    assert np.max(np.abs(s2 - s2_ref)) <= 2e-6
    # the forward kernel's certified double-precision form: the same pixels,
    # cos/sin(2 chi) to float rounding
    b_3, c3, s3 = gpu.match_fast(nside_sky, nside_beam, ra, dec, pa, off, pix, mode=3)
    assert np.array_equal(b_h, b_3)
    assert np.max(np.abs(c3 - c2_ref)) <= 2.5e-7
    assert np.max(np.abs(s3 - s2_ref)) <= 2.5e-7
    # the one-wave-per-sample forward's form: float alpha, mode 3 where it
    # declines; and that form alone, which must decline rarely
    b_4, c4, s4 = gpu.match_fast(nside_sky, nside_beam, ra, dec, pa, off, pix, mode=4)
    assert np.array_equal(b_h, b_4)
    assert np.max(np.abs(c4 - c2_ref)) <= 2.5e-7
    assert np.max(np.abs(s4 - s2_ref)) <= 2.5e-7
    b_5, _, _ = gpu.match_fast(nside_sky, nside_beam, ra, dec, pa, off, pix, mode=5)
    cert5 = b_5 != -2
    assert np.array_equal(b_5[cert5], b_h[cert5])
    print("float-alpha form declined %.2e" % (1.0 - cert5.mean()))
    assert 1.0 - cert5.mean() < 0.02
    b_1, _, _ = gpu.match_fast(nside_sky, nside_beam, ra, dec, pa, off, pix, mode=2)
    cert = b_1 != -2
    assert np.array_equal(b_1[cert], b_h[cert])
    declined = 1.0 - cert.mean()
    print("declined fraction %.2e of %d visits" % (declined, b_1.size))
    assert declined < 0.01
