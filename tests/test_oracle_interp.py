"""CPU: the oracle's beam interpolation mode (the interpolating form of
convolver.cpp:124-147, commented out in the reference) pinned by closed
forms: a beam linear in theta is reproduced exactly at any (rho, sigma)
between the first and the last full ring, which the nearest-pixel mode
cannot do; a uniform beam stays normalised up to the cap edge; and the
adjoints satisfy the dot-product identity."""
import math

import numpy as np
import pytest

from pisco4cmb_amd import pypisco

DEG = math.pi / 180


@pytest.fixture
def interp(oracle):
    oracle.set_beam_interpolation(True)
    yield oracle
    oracle.set_beam_interpolation(False)


def point_source_setup(oracle, ns, nb, cap_deg, beam_fn, n=400, seed=0):
    npb = pypisco.beam_npix_for_radius(nb, cap_deg * DEG)
    th, _ = oracle.pix2ang(nb, np.arange(npb, dtype=np.int32))
    a0 = beam_fn(th).astype(np.float32)
    z = np.zeros(npb, np.float32)
    beams = [a0, z, z, z, a0, z, z, z]
    npix = 12 * ns * ns
    p0 = 12 * ns * ns // 2 + 3 * ns
    t0, f0 = oracle.pix2ang(ns, np.array([p0], np.int32))
    I = np.zeros(npix, np.float32)
    I[p0] = 1.0
    rng = np.random.default_rng(seed)
    # samples around the source, up to the cap
    r = np.sqrt(rng.uniform(0, 1, n)) * cap_deg * DEG
    b = rng.uniform(0, 2 * math.pi, n)
    dec0 = math.pi / 2 - t0[0]
    dec = (dec0 + r * np.cos(b)).astype(np.float32)
    ra = (f0[0] + r * np.sin(b) / math.cos(dec0)).astype(np.float32)
    pa = rng.uniform(-math.pi, math.pi, n).astype(np.float32)
    # angular distance of the source from each sample (3-vectors)
    v0 = np.array([math.sin(t0[0]) * math.cos(f0[0]), math.sin(t0[0]) * math.sin(f0[0]),
                   math.cos(t0[0])])
    d64, r64 = dec.astype(np.float64), ra.astype(np.float64)
    vs = np.stack([np.cos(d64) * np.cos(r64), np.cos(d64) * np.sin(r64), np.sin(d64)], 1)
    rho = np.arccos(np.clip(vs @ v0, -1, 1))
    return I, None, beams, ra, dec, pa, rho, npb, th


def test_linear_beam_is_reproduced(oracle, interp):
    ns, nb, cap = 64, 256, 6.0
    I, _, beams, ra, dec, pa, rho, npb, th = point_source_setup(
        oracle, ns, nb, cap, lambda t: 1.0 + 5.0 * t)
    npix = 12 * ns * ns
    zero = np.zeros(npix, np.float32)
    ta, tb = interp.convolve(ns, I, zero, zero, nb, beams, ra, dec, pa)
    th1 = th[0]   # the first ring; the last rings are avoided by 0.3 deg
    sel = (rho > 2 * th1) & (rho < cap * DEG - 0.3 * DEG) & (ta != 0)
    assert sel.sum() > 100
    expect = 1.0 + 5.0 * rho[sel]
    assert np.abs(ta[sel] - expect).max() < 2e-6 * expect.max()
    assert np.array_equal(ta, tb)
    # the nearest-pixel mode gives 1 + 5 theta(nearest pixel), off by up to
    # half a ring spacing in theta
    oracle.set_beam_interpolation(False)
    na, _ = oracle.convolve(ns, I, zero, zero, nb, beams, ra, dec, pa)
    oracle.set_beam_interpolation(True)
    assert np.abs(na[sel] - expect).max() > 1e-4


def test_uniform_beam_stays_normalised_at_the_edge(oracle, interp):
    ns, nb, cap = 64, 256, 4.0
    I, _, beams, ra, dec, pa, rho, npb, th = point_source_setup(
        oracle, ns, nb, cap, lambda t: np.ones_like(t), n=600, seed=1)
    zero = np.zeros(12 * ns * ns, np.float32)
    ta, _ = interp.convolve(ns, I, zero, zero, nb, beams, ra, dec, pa)
    inside = rho < cap * DEG - 0.2 * DEG
    assert np.abs(ta[inside] - 1.0).max() < 1e-6
    # at the edge a sample is either outside (0) or normalised (1)
    edge = ~inside
    assert np.all((ta[edge] == 0) | (np.abs(ta[edge] - 1.0) < 1e-6))


def test_adjoints_dot_product(oracle, interp):
    ns, nb = 32, 256
    _, _, beams, ra, dec, pa, _, npb, _ = point_source_setup(
        oracle, ns, nb, 5.0, lambda t: np.exp(-t / 0.02), n=200, seed=2)
    rng = np.random.default_rng(3)
    beams = [b + rng.normal(0, 0.1, b.size).astype(np.float32) for b in beams]
    npix = 12 * ns * ns
    I, Q, U = (rng.normal(size=npix).astype(np.float32) for _ in range(3))
    ga, gb = (rng.normal(size=ra.size).astype(np.float32) for _ in range(2))
    da, db = interp.convolve(ns, I, Q, U, nb, beams, ra, dec, pa, precision64=True)
    lhs = float(da @ ga + db @ gb)
    gI, gQ, gU = interp.adjoint_map(ns, nb, beams, ra, dec, pa, ga, gb)
    rhs = float(gI @ I + gQ @ Q + gU @ U)
    assert abs(lhs - rhs) < 1e-6 * abs(lhs)
    g = interp.adjoint_beam(ns, I, Q, U, nb, beams, ra, dec, pa, ga, gb)
    rhs_b = sum(float(g[:, k] @ beams[k].astype(np.float64)) for k in range(3)) + \
        sum(float(g[:, 3 + k] @ beams[4 + k].astype(np.float64)) for k in range(3))
    assert abs(lhs - rhs_b) < 1e-5 * abs(lhs)
