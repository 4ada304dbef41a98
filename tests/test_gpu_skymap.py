"""The GPU map-maker (pisco_skymap_*) against the CPU oracle's SkyMap
restatement (skymap.cpp:68-231): the normal equations bit for bit (the
device adds each pixel's samples in the reference's order), the solved
I, Q, U and hits bit for bit, over several accumulation passes and calls,
with shared pointing, singular and under-covered pixels, invalid pointing,
and at the size of a 64-detector NSIDE 512 scan.  A constant sky convolved
on the GPU and mapped on the GPU comes back as that sky."""
import math

import numpy as np
import pytest
import torch

from pisco4cmb_amd import inputs, pypisco, torch_ops
from pisco4cmb_amd._lib import check, lib
from pisco4cmb_amd.mapmaker import DeviceSkyMap

pytestmark = pytest.mark.gpu
DEG = math.pi / 180


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def pointing(rng, ndet, n):
    ra = rng.uniform(0, 2 * math.pi, (ndet, n)).astype(np.float32)
    dec = np.arcsin(rng.uniform(-1, 1, (ndet, n))).astype(np.float32)
    pa = rng.uniform(-math.pi, math.pi, (ndet, n)).astype(np.float32)
    return ra, dec, pa


def assert_same_map(oracle, nside, m_normal, m_solved, ora_AtA, ora_AtD):
    AtA, AtD = m_normal
    assert np.array_equal(AtA, ora_AtA)
    assert np.array_equal(AtD, ora_AtD)
    I, Q, U, hits = m_solved
    oI, oQ, oU, ohits, _ = oracle.skymap_solve(nside, ora_AtA, ora_AtD)
    assert np.array_equal(hits, ohits)
    for a, b in ((I, oI), (Q, oQ), (U, oU)):
        assert np.array_equal(a.view(np.int32), b.view(np.int32))


@pytest.mark.parametrize("chunk", [1 << 24, 7777])
def test_host_api_bit_exact_over_calls_and_passes(oracle, chunk):
    nside = 64
    rng = np.random.default_rng(11)
    ndet, n = 4, 60000
    ra, dec, pa = pointing(rng, ndet, n)
    # concentrate half the samples on a small patch so pixels get many hits
    ra[:, ::2] = rng.uniform(1.0, 1.05, (ndet, n // 2)).astype(np.float32)
    dec[:, ::2] = rng.uniform(0.3, 0.35, (ndet, n // 2)).astype(np.float32)
    pol = np.array([0.0, 0.25 * math.pi, 0.5 * math.pi, 0.75 * math.pi], np.float32)
    data = rng.normal(0, 50, (ndet, n)).astype(np.float32)
    m = pypisco.SkyMap(nside)
    m.set_chunk(chunk)
    m.accumulate_data(n, ra[:3], dec[:3], pa[:3], 3, pol[:3], data[:3])
    m.accumulate_data(n, ra[3:], dec[3:], pa[3:], 1, pol[3:], data[3:])
    m.solve_map()
    AtA, AtD, _ = oracle.skymap_accumulate(nside, ra, dec, pa, pol, data)
    assert_same_map(oracle, nside, m.get_normal(),
                    (m.get_stokes_I(), m.get_stokes_Q(), m.get_stokes_U(), m.get_hitmap()),
                    AtA, AtD)
    assert m.get_hitmap().max() > 50
    m.close()


def test_device_shared_pointing_bit_exact(oracle):
    nside = 32
    rng = np.random.default_rng(12)
    ra, dec, pa = pointing(rng, 1, 30000)
    pol = np.array([0.1, 1.2, -0.4], np.float32)
    data = rng.normal(0, 1, (3, 30000)).astype(np.float32)
    m = DeviceSkyMap(nside)
    m.accumulate(t(ra[0]), t(dec[0]), t(pa[0]), t(pol), t(data), shared_pointing=True)
    I, Q, U, hits = (x.cpu().numpy() for x in m.solve())
    torch.cuda.synchronize()
    AtA, AtD, _ = oracle.skymap_accumulate(nside, ra[0], dec[0], pa[0], pol, data,
                                           shared_pointing=True)
    AtA_d = np.empty((12 * nside * nside, 9))
    AtD_d = np.empty((12 * nside * nside, 3))
    check(lib.pisco_skymap_get_normal(m.handle, AtA_d.ctypes.data, AtD_d.ctypes.data))
    assert_same_map(oracle, nside, (AtA_d, AtD_d), (I, Q, U, hits), AtA, AtD)
    assert m.skipped() == 0
    m.close()


def test_singular_and_undercovered_pixels(oracle):
    """5 samples at one angle (singular: I, Q, U = AtD, hits 0) and a pixel
    with 2 hits (INSUFFCOV: zeros, hits kept), as skymap.cpp:184-223."""
    nside = 4
    ra = np.array([[1.0] * 5 + [3.0] * 2], np.float32)
    dec = np.array([[0.2] * 5 + [-0.5] * 2], np.float32)
    pa = np.zeros((1, 7), np.float32)
    data = np.arange(1, 8, dtype=np.float32)[None]
    m = pypisco.SkyMap(nside)
    m.accumulate_data(7, ra, dec, pa, 1, np.zeros(1, np.float32), data)
    m.solve_map()
    AtA, AtD, _ = oracle.skymap_accumulate(nside, ra, dec, pa, np.zeros(1, np.float32), data)
    assert_same_map(oracle, nside, m.get_normal(),
                    (m.get_stokes_I(), m.get_stokes_Q(), m.get_stokes_U(), m.get_hitmap()),
                    AtA, AtD)
    assert m.singular == 1
    assert sorted(m.get_hitmap()[m.get_hitmap() > 0]) == [2]
    m.close()


def test_invalid_pointing_raises_after_accumulating_the_rest(oracle):
    nside = 8
    rng = np.random.default_rng(13)
    ra, dec, pa = pointing(rng, 2, 1000)
    dec[0, 10] = 2.0
    ra[1, 20] = np.nan
    pol = np.zeros(2, np.float32)
    data = rng.normal(size=(2, 1000)).astype(np.float32)
    m = pypisco.SkyMap(nside)
    with pytest.raises(ValueError):
        m.accumulate_data(1000, ra, dec, pa, 2, pol, data)
    AtA, AtD, skipped = oracle.skymap_accumulate(nside, ra, dec, pa, pol, data)
    assert skipped == 2
    A, D = m.get_normal()
    assert np.array_equal(A, AtA) and np.array_equal(D, AtD)
    m.close()


def test_full_size_64_detectors_nside512(oracle):
    """64 detectors x 250000 samples of one CES pointing into an NSIDE 512
    map: sums and solution bit for bit."""
    nside = 512
    n, ndet = 250000, 64
    ra, dec, pa = inputs.ces_scan(n, t0=3600.0)
    rng = np.random.default_rng(14)
    pol = rng.uniform(-math.pi, math.pi, ndet).astype(np.float32)
    data = rng.normal(0, 30, (ndet, n)).astype(np.float32)
    m = DeviceSkyMap(nside)
    m.accumulate(t(ra), t(dec), t(pa), t(pol), t(data), shared_pointing=True)
    I, Q, U, hits = (x.cpu().numpy() for x in m.solve())
    AtA, AtD, _ = oracle.skymap_accumulate(nside, ra, dec, pa, pol, data, shared_pointing=True)
    oI, oQ, oU, ohits, _ = oracle.skymap_solve(nside, AtA, AtD)
    assert np.array_equal(hits, ohits)
    for a, b in ((I, oI), (Q, oQ), (U, oU)):
        assert np.array_equal(a.view(np.int32), b.view(np.int32))
    assert (hits >= 3).sum() > 10000
    m.close()


def test_gpu_convolve_then_map_round_trip():
    """A constant I, Q, U sky through a 1-degree pair on the GPU, mapped on
    the GPU with detector angles 0 and 90 degrees, comes back as that sky."""
    ns, nb = 64, 512
    beam = inputs.gaussian_polbeam(nb, 3.0, 1.0, 0.0, 0.0, nside_sky=ns)
    npix = 12 * ns * ns
    I0, Q0, U0 = 100.0, 10.0, -7.0
    I, Q, U = (torch.full((npix,), v, dtype=torch.float32, device="cuda") for v in (I0, Q0, U0))
    rng = np.random.default_rng(15)
    n = 400000
    ra = t(rng.uniform(2.0, 2.6, n).astype(np.float32))
    dec = t(rng.uniform(-0.3, 0.3, n).astype(np.float32))
    pa = t(rng.uniform(-math.pi, math.pi, n).astype(np.float32))
    ta, tb = torch_ops.beam_convolve(I, Q, U, t(np.stack(beam.aBeams)), t(np.stack(beam.bBeams)),
                                     ra, dec, pa, ns, nb)
    m = DeviceSkyMap(ns)
    m.accumulate(ra, dec, pa, t(np.array([0.0, math.pi / 2], np.float32)),
                 torch.stack([ta, tb]), shared_pointing=True)
    Im, Qm, Um, hits = (x.cpu().numpy() for x in m.solve())
    sel = hits >= 20
    assert sel.sum() > 1000
    assert abs(np.median(Im[sel]) - I0) < 0.1
    assert abs(np.median(Qm[sel]) - Q0) < 0.05
    assert abs(np.median(Um[sel]) - U0) < 0.05
    m.close()
