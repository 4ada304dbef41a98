"""TOD and adjoint parity of the HIP path against the CPU oracle.

Every family runs on co-polar Gaussian beams (aBeams[2] = bBeams[2] = 0)
and on cross-polar beams (inputs.crosspolar_polbeam: all beam maps
non-zero), so the A2/B2 terms of convolver.cpp:157-164 are compared too.

Tolerances: TOD within 1e-6 relative RMS of the reference-faithful oracle
(float products, FP64 accumulation, convolver.cpp:157-164); the error against
the full-FP64 oracle is asserted too.  Adjoints against the oracle's FP64
adjoints (relative RMS 1e-6); dot-product test <Ax,y> = <x,A^T y> to 1e-6.
"""
import math

import numpy as np
import pytest

from pisco4cmb_amd import inputs, pypisco

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def run_gpu(nside_sky, sky, beam, ra, dec, pa, polflag="p", ptg=100000):
    conv = pypisco.GPUConvolver(pypisco.default_run_config(ptgPerConv=ptg))
    conv.configure_execution()
    conv.update_sky(pypisco.Sky(nside_sky, *sky))
    conv.update_beam(beam)
    conv.update_scan(pypisco.Scan(ra.size, ra, dec, pa))
    a = np.zeros(ra.size, np.float32)
    b = np.zeros(ra.size, np.float32)
    conv.exec_convolution(a, b, polflag)
    return conv, a, b


def beams_of(b):
    return list(b.aBeams) + list(b.bBeams)


KINDS = ["copolar", "crosspolar"]


@pytest.fixture(scope="module", params=KINDS)
def c1(request):
    """Config 1: NSIDE 128 iid sky, beam NSIDE 1024, cap 2.5 deg, FWHM 1 deg,
    ellipticity 0.05, leakage 1%, 1e4 raster samples, pa cycling."""
    sky = inputs.iid_sky(128, seed=1)
    beam = inputs.make_polbeam(request.param, 1024, 2.5, 1.0, 0.05, 0.01, nside_sky=128)
    ra, dec, pa = inputs.raster_scan(10000)
    return sky, beam, ra, dec, pa


def test_forward_config1_matches_oracle(c1, oracle):
    sky, beam, ra, dec, pa = c1
    _, a, b = run_gpu(128, sky, beam, ra, dec, pa)
    ra_, rb_ = oracle.convolve(128, sky[0], sky[1], sky[2], 1024, beams_of(beam), ra, dec, pa)
    assert rel_rms(a, ra_) < 1e-6 and rel_rms(b, rb_) < 1e-6
    # per sample: the kernels form cos/sin(2 chi) in float (error <= 2e-6),
    # which moves a sample by far less than 1e-5 of the TOD's scale
    scale = np.abs(ra_).max()
    assert np.max(np.abs(a - ra_)) < 1e-5 * scale
    assert np.max(np.abs(b - rb_)) < 1e-5 * np.abs(rb_).max()
    da, db = oracle.convolve(128, sky[0], sky[1], sky[2], 1024, beams_of(beam), ra, dec, pa,
                             precision64=True)
    assert rel_rms(a, da) < 1e-6 and rel_rms(b, db) < 1e-6
    print("config1 rel-rms vs oracle %.3g/%.3g, vs FP64 oracle %.3g/%.3g"
          % (rel_rms(a, ra_), rel_rms(b, rb_), rel_rms(a, da), rel_rms(b, db)))


def test_polflags(c1, oracle):
    sky, beam, ra, dec, pa = c1
    n = 2000
    _, a, b = run_gpu(128, sky, beam, ra[:n], dec[:n], pa[:n], "p")
    _, a2, b2 = run_gpu(128, sky, beam, ra[:n], dec[:n], pa[:n], "a")
    _, a3, b3 = run_gpu(128, sky, beam, ra[:n], dec[:n], pa[:n], "b")
    assert np.array_equal(a, a2) and np.all(b2 == 0)
    assert np.array_equal(b, b3) and np.all(a3 == 0)
    with pytest.raises(ValueError):
        run_gpu(128, sky, beam, ra[:n], dec[:n], pa[:n], "x")


def test_forward_chunking_and_tail(c1, oracle):
    """ptgPerConv does not change results and the nsamples % ptgPerConv tail
    is computed (the reference drops it, gpuconvolver.cu:386)."""
    sky, beam, ra, dec, pa = c1
    n = 3001
    _, a1, b1 = run_gpu(128, sky, beam, ra[:n], dec[:n], pa[:n], ptg=1000)
    _, a2, b2 = run_gpu(128, sky, beam, ra[:n], dec[:n], pa[:n], ptg=7)
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
    ra_, rb_ = oracle.convolve(128, sky[0], sky[1], sky[2], 1024, beams_of(beam), ra[:n],
                               dec[:n], pa[:n])
    assert rel_rms(a1[-1:], ra_[-1:]) < 1e-6


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("nside_sky,nside_beam,cap,fwhm,n", [(256, 2048, 1.5, 0.5, 4000),
                                                             (512, 4096, 5.0, 0.33, 300),
                                                             (64, 64, 10.0, 3.0, 3000)])
def test_forward_other_configs(oracle, nside_sky, nside_beam, cap, fwhm, n, kind):
    sky = inputs.iid_sky(nside_sky, seed=nside_sky)
    beam = inputs.make_polbeam(kind, nside_beam, cap, fwhm, 0.05, 0.01, nside_sky=nside_sky)
    ra, dec, pa = inputs.ces_scan(n, fs=2.0)
    _, a, b = run_gpu(nside_sky, sky, beam, ra, dec, pa)
    ra_, rb_ = oracle.convolve(nside_sky, *sky[:3], nside_beam, beams_of(beam), ra, dec, pa)
    assert rel_rms(a, ra_) < 1e-6 and rel_rms(b, rb_) < 1e-6


def test_adjoint_map_matches_oracle_and_dot_product(c1, oracle):
    sky, beam, ra, dec, pa = c1
    n = 3000
    ra, dec, pa = ra[:n], dec[:n], pa[:n]
    rng = np.random.default_rng(9)
    conv, a, b = run_gpu(128, sky, beam, ra, dec, pa)
    # y = A x + noise: <A x, y> of the order of |A x| |y|, no cancellation
    ga = (a + a.std() * rng.normal(size=n)).astype(np.float32)
    gb = (b + b.std() * rng.normal(size=n)).astype(np.float32)
    gI, gQ, gU = conv.exec_adjoint_map(ga, gb)
    oI, oQ, oU = oracle.adjoint_map(128, 1024, beams_of(beam), ra, dec, pa, ga, gb)
    for g, o in ((gI, oI), (gQ, oQ), (gU, oU)):
        assert rel_rms(g, o) < 1e-6
    # <A x, y> == <x, A^T y>
    lhs = float(np.dot(a.astype(np.float64), ga) + np.dot(b.astype(np.float64), gb))
    rhs = float(np.dot(sky[0].astype(np.float64), gI) + np.dot(sky[1].astype(np.float64), gQ)
                + np.dot(sky[2].astype(np.float64), gU))
    assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), 1e-30)


def test_adjoint_beam_matches_oracle_and_dot_product(c1, oracle):
    sky, beam, ra, dec, pa = c1
    n = 2000
    ra, dec, pa = ra[:n], dec[:n], pa[:n]
    rng = np.random.default_rng(10)
    ga = rng.normal(size=n).astype(np.float32)
    gb = rng.normal(size=n).astype(np.float32)
    conv, a, b = run_gpu(128, sky, beam, ra, dec, pa)
    oa, ob = conv.exec_adjoint_beam(ga, gb)
    og = oracle.adjoint_beam(128, *sky[:3], 1024, beams_of(beam), ra, dec, pa, ga, gb)
    for k in range(3):
        assert rel_rms(oa[k], og[:, k]) < 1e-6
        assert rel_rms(ob[k], og[:, 3 + k]) < 1e-6
    assert np.all(oa[3] == 0) and np.all(ob[3] == 0)
    lhs = float(np.dot(a.astype(np.float64), ga) + np.dot(b.astype(np.float64), gb))
    rhs = sum(float(np.dot(beam.aBeams[k].astype(np.float64), oa[k]))
              + float(np.dot(beam.bBeams[k].astype(np.float64), ob[k])) for k in range(3))
    assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), 1e-30)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("polflag", ["p", "a", "b"])
def test_host_api_split_copies(oracle, polflag, kind):
    """From 2^18 samples the host-API forward runs in 4 launches whose TOD
    copies overlap the next launch: 2^18 + 7 samples (a ragged last part)
    give the bits of the device-resident forward (one launch) for every
    polFlag, and the last part's samples match the oracle."""
    import torch

    from pisco4cmb_amd import torch_ops
    ns, nb, n = 64, 512, (1 << 18) + 7
    sky = inputs.iid_sky(ns, seed=71)
    beam = inputs.make_polbeam(kind, nb, 3.0, 1.2, 0.05, 0.01, nside_sky=ns)
    ra, dec, pa = inputs.ces_scan(n, t0=300.0)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to("cuda")
    da, db = (x.cpu().numpy() for x in torch_ops.beam_convolve(
        *(t(x) for x in sky[:3]), t(np.stack(beam.aBeams)), t(np.stack(beam.bBeams)), t(ra),
        t(dec), t(pa), ns, nb))
    _, a, b = run_gpu(ns, sky, beam, ra, dec, pa, polflag, ptg=n)
    assert np.array_equal(a, da if polflag != "b" else np.zeros_like(a))
    assert np.array_equal(b, db if polflag != "a" else np.zeros_like(b))
    k = slice(n - 500, n)
    oa, ob = oracle.convolve(ns, *sky[:3], nb, beams_of(beam), ra[k], dec[k], pa[k])
    assert rel_rms(da[k], oa) < 1e-6 and rel_rms(db[k], ob) < 1e-6


@pytest.mark.parametrize("kind", KINDS)
def test_host_api_split_copies_interpolation(oracle, kind):
    """The same 4-launch host-API forward in beam interpolation mode: 2^18 + 7
    samples, whose last 3,000 equal a host-API call over those samples alone
    (one launch) bit for bit, and whose last 300 match the oracle's
    interpolated forward."""
    ns, nb, n = 64, 512, (1 << 18) + 7
    sky = inputs.iid_sky(ns, seed=73)
    beam = inputs.make_polbeam(kind, nb, 4.0, 1.5, 0.08, 0.02, nside_sky=ns)
    ra, dec, pa = inputs.ces_scan(n, t0=900.0)

    def run(r, d, q):
        conv = pypisco.GPUConvolver(pypisco.default_run_config(ptgPerConv=r.size))
        conv.configure_execution()
        conv.set_tuning("beam_interpolation", 1)
        conv.update_sky(pypisco.Sky(ns, *sky[:3]))
        conv.update_beam(beam)
        conv.update_scan(pypisco.Scan(r.size, r, d, q))
        a = np.zeros(r.size, np.float32)
        b = np.zeros(r.size, np.float32)
        conv.exec_convolution(a, b)
        return a, b

    fa, fb = run(ra, dec, pa)
    k = slice(n - 3000, n)
    ta, tb = run(ra[k], dec[k], pa[k])
    assert np.array_equal(fa[k], ta) and np.array_equal(fb[k], tb)
    m = slice(n - 300, n)
    oracle.set_beam_interpolation(True)
    try:
        oa, ob = oracle.convolve(ns, *sky[:3], nb, beams_of(beam), ra[m], dec[m], pa[m])
    finally:
        oracle.set_beam_interpolation(False)
    assert rel_rms(fa[m], oa) < 1e-6 and rel_rms(fb[m], ob) < 1e-6
