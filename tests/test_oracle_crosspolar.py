"""CPU: the oracle's cross-polar terms pinned without retyping them.

GRASP-like beams carry a cross-polar E field, so tilde U and tilde V of
PolBeam::beam_from_fields (polbeam.cpp:170-183) are non-zero, and the TOD
reads them through the A2/B2 terms of convolver.cpp:157-164.  Pins:

  beam_from_fields   closed forms: real fields (cos a, sin a) give
                     (I, Q, U, V) = (1, cos 2a, sin 2a, 0); E_cx = i E_co
                     gives U = 0, V = 2|E_co|^2; a rotation of the field
                     basis by b turns Q + iU by exp(-2ib) and leaves I, V;
                     the product's PolBeam equals the oracle bit for bit
  forward            the visit term of detector a is
                     Re[(Q - iU)(A1 + iA2) exp(2i chi)] and that of b its
                     negative with B, so turning a beam's (A1, A2) into
                     (-A2, A1) equals feeding the sky (U, -Q): checked per
                     detector, in FP64 (1e-12) and in the reference's float
                     arithmetic (1e-6)
  adjoints           dot products <F x, g> = <x, F^T g> with every beam
                     column non-zero, and per beam column: the beam
                     gradient's column k dotted with column k equals the
                     forward with only that column kept
A dropped A2/B2 term, a wrong sign in one of them, or a slot swap between the
a and b maps fails one of these (the module's last test checks that claim on
a planted mistake in a numpy copy of the visit term).
"""
import math

import numpy as np
import pytest

from pisco4cmb_amd import inputs, pypisco

DEG = math.pi / 180
NS, NB, CAP, FWHM = 64, 512, 6.0, 2.0


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def xbeam(oracle):
    npix = pypisco.beam_npix_for_radius(NB, CAP * DEG)
    fa = inputs.crosspolar_fields(NB, npix, FWHM, 0.05, 0.0, 0.35, 0.6, 0.3)
    fb = inputs.crosspolar_fields(NB, npix, FWHM, 0.05, 30.0, 0.25, -0.9, 1.2)
    beams = oracle.polbeam_from_fields(npix, NB, NS, 0.01, fa, fb)
    return npix, fa, fb, beams


@pytest.fixture(scope="module")
def scan():
    rng = np.random.default_rng(21)
    n = 300
    ra = rng.uniform(0, 2 * math.pi, n).astype(np.float32)
    dec = np.arcsin(rng.uniform(-0.9, 0.9, n)).astype(np.float32)
    pa = rng.uniform(-math.pi, math.pi, n).astype(np.float32)
    return ra, dec, pa


def test_crosspolar_beam_is_crosspolar(xbeam):
    """The fixture exercises the terms: |A2|, |B2| >= 0.1 max|A1| (and
    A3, B3 non-zero), with both signs, and differ between a and b."""
    _, _, _, b = xbeam
    for k1, k2 in ((1, 2), (5, 6)):
        assert np.abs(b[k2]).max() >= 0.1 * np.abs(b[k1]).max()
        assert (b[k2] > 0).any() and (b[k2] < 0).any()
    assert np.abs(b[3]).max() > 0 and np.abs(b[7]).max() > 0
    assert not np.allclose(b[2], b[6])


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_beam_from_fields_closed_forms(oracle, which):
    rng = np.random.default_rng(5)
    n = 4000
    assert pypisco.beam_npix_for_radius(64, 40 * DEG) >= n   # a valid PolBeam size

    def tilde(f):
        if which == "oracle":
            return oracle.beam_from_fields(n, *f)
        bm = pypisco.PolBeam(64, n, 0.0, "p")
        bm.beam_from_fields("a", *f)
        return bm.Ia, bm.Qa, bm.Ua, bm.Va

    # real fields E_co = cos a, E_cx = sin a
    a = rng.uniform(-math.pi, math.pi, n)
    I, Q, U, V = tilde((np.abs(np.cos(a)), np.where(np.cos(a) < 0, math.pi, 0.0),
                        np.abs(np.sin(a)), np.where(np.sin(a) < 0, math.pi, 0.0)))
    assert np.allclose(I, 1.0, atol=2e-7)
    assert np.allclose(Q, np.cos(2 * a), atol=2e-7)
    assert np.allclose(U, np.sin(2 * a), atol=2e-7)
    assert np.all(np.abs(V) < 2e-7)
    # E_cx = i E_co
    m = rng.uniform(0.1, 2.0, n)
    ph = rng.uniform(-math.pi, math.pi, n)
    I, Q, U, V = tilde((m, ph, m, ph + math.pi / 2))
    assert np.allclose(I, 2 * m * m, rtol=2e-7)
    assert np.all(np.abs(Q) <= 2e-7 * m * m) and np.all(np.abs(U) <= 2e-7 * m * m)
    assert np.allclose(V, 2 * m * m, rtol=2e-7)
    # rotation of the field basis by b: Q + iU -> exp(-2ib)(Q + iU)
    co = rng.normal(size=n) + 1j * rng.normal(size=n)
    cx = rng.normal(size=n) + 1j * rng.normal(size=n)
    f0 = (np.abs(co), np.angle(co), np.abs(cx), np.angle(cx))
    for b in (0.3, math.pi / 4, -1.1):
        co2 = co * math.cos(b) + cx * math.sin(b)
        cx2 = -co * math.sin(b) + cx * math.cos(b)
        I0, Q0, U0, V0 = (np.asarray(x, np.float64) for x in tilde(f0))
        I1, Q1, U1, V1 = (np.asarray(x, np.float64) for x in
                          tilde((np.abs(co2), np.angle(co2), np.abs(cx2), np.angle(cx2))))
        P = (Q0 + 1j * U0) * np.exp(-2j * b)
        scale = I0 + 1e-30
        assert np.all(np.abs(I1 - I0) <= 1e-6 * scale)
        assert np.all(np.abs(V1 - V0) <= 1e-6 * scale)
        assert np.all(np.abs(Q1 - P.real) <= 1e-6 * scale)
        assert np.all(np.abs(U1 - P.imag) <= 1e-6 * scale)


def test_product_crosspolar_polbeam_matches_oracle_bit_exact(oracle, xbeam):
    npix, fa, fb, beams = xbeam
    bm = pypisco.PolBeam(NB, npix, 0.01, "p")
    bm.beam_from_fields("a", *fa)
    bm.beam_from_fields("b", *fb)
    bm.build(NS)
    for k in range(4):
        assert np.array_equal(bm.aBeams[k], beams[k])
        assert np.array_equal(bm.bBeams[k], beams[4 + k])
    # inputs.crosspolar_polbeam is this construction
    b2 = inputs.crosspolar_polbeam(NB, CAP, FWHM, 0.05, 0.01, nside_sky=NS)
    for k in range(4):
        assert np.array_equal(b2.aBeams[k], beams[k])
        assert np.array_equal(b2.bBeams[k], beams[4 + k])


def rotate(beams, det):
    """(A1, A2) -> (-A2, A1) for detector det ('a', 'b' or 'p')."""
    out = [b.copy() for b in beams]
    for d, o in (("a", 0), ("b", 4)):
        if det in (d, "p"):
            out[o + 1] = -beams[o + 2]
            out[o + 2] = beams[o + 1].copy()
    return out


@pytest.mark.parametrize("det", ["a", "b", "p"])
def test_polarisation_rotation_identity(oracle, xbeam, scan, det):
    """TOD(I, Q, U; beam with (A1, A2) -> (-A2, A1)) = TOD(I, U, -Q; beam),
    for the rotated detector; the other detector's TOD is unchanged."""
    _, _, _, beams = xbeam
    ra, dec, pa = scan
    I, Q, U, _ = inputs.iid_sky(NS, 31)
    rb = rotate(beams, det)
    mQ = (-Q).astype(np.float32)
    for p64, tol in ((True, 1e-12), (False, 1e-6)):
        a1, b1 = oracle.convolve(NS, I, Q, U, NB, rb, ra, dec, pa, precision64=p64)
        a0, b0 = oracle.convolve(NS, I, Q, U, NB, beams, ra, dec, pa, precision64=p64)
        a2, b2 = oracle.convolve(NS, I, U, mQ, NB, beams, ra, dec, pa, precision64=p64)
        assert rel(a1, a2 if det in "ap" else a0) < tol
        assert rel(b1, b2 if det in "bp" else b0) < tol
        assert rel(a2, a0) > 1e-2 and rel(b2, b0) > 1e-2   # the identity is not trivial


def test_adjoint_dot_products_per_beam_column(oracle, xbeam, scan):
    """<F x, g> = <x, F^T g> for the sky, and for each used beam column k
    separately: <G_k, A_k> = <F(beam with only column k), g>.  FP64 forward;
    the adjoints round cos/sin(2 chi) to float like the reference, so the
    tolerance is 1e-6 relative."""
    _, _, _, beams = xbeam
    ra, dec, pa = scan
    rng = np.random.default_rng(41)
    npix = 12 * NS * NS
    I, Q, U = (rng.normal(size=npix).astype(np.float32) for _ in range(3))
    ga, gb = (rng.normal(size=ra.size).astype(np.float32) for _ in range(2))
    da, db = oracle.convolve(NS, I, Q, U, NB, beams, ra, dec, pa, precision64=True)
    lhs = float(np.dot(da, ga) + np.dot(db, gb))
    gI, gQ, gU = oracle.adjoint_map(NS, NB, beams, ra, dec, pa, ga, gb)
    rhs = float(np.dot(gI, I) + np.dot(gQ, Q) + np.dot(gU, U))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
    G = oracle.adjoint_beam(NS, I, Q, U, NB, beams, ra, dec, pa, ga, gb)
    zero = np.zeros_like(beams[0])
    for col, slot in enumerate((0, 1, 2, 4, 5, 6)):
        only = [beams[k] if k == slot else zero for k in range(8)]
        fa, fb = oracle.convolve(NS, I, Q, U, NB, only, ra, dec, pa, precision64=True)
        want = float(np.dot(fa, ga) + np.dot(fb, gb))
        got = float(np.dot(G[:, col], beams[slot].astype(np.float64)))
        assert abs(got - want) <= 1e-6 * max(abs(want), 1e-3 * abs(lhs)), (slot, got, want)


def test_identity_catches_planted_mistakes(oracle, xbeam, scan):
    """The rotation identity used above fails for the plausible mistakes in
    the visit term: planted in a numpy copy of the detector-a term, evaluated
    on the oracle's own beam pixels and chi."""
    _, _, _, beams = xbeam
    ra, dec, pa = scan
    I, Q, U, _ = inputs.iid_sky(NS, 31)
    off, pix = oracle.query_disc_many(NS, ra, dec, oracle.beam_rho_max(NB, beams[0].size))
    bpix, chi = oracle.match(NS, NB, ra, dec, pa, off, pix)
    keep = bpix < beams[0].size
    smp = np.repeat(np.arange(ra.size), np.diff(off))[keep]
    p, b, c2 = pix[keep], bpix[keep], 2 * chi[keep]
    c, s = np.cos(c2), np.sin(c2)

    def tod(q, u, A1, A2, form):
        q, u, a1, a2 = q[p].astype(np.float64), u[p].astype(np.float64), A1[b], A2[b]
        t = {"right": q * (a1 * c - a2 * s) + u * (a2 * c + a1 * s),
             "sign": q * (a1 * c + a2 * s) + u * (a2 * c + a1 * s),
             "dropped": q * (a1 * c) + u * (a2 * c + a1 * s),
             "swap": q * (a1 * c - a2 * s) + u * (a1 * c + a2 * s)}[form]
        return np.bincount(smp, t, minlength=ra.size)

    A1, A2 = beams[1].astype(np.float64), beams[2].astype(np.float64)
    for form in ("right", "sign", "dropped", "swap"):
        lhs = tod(Q, U, -A2, A1, form)
        rhs = tod(U, -Q, A1, A2, form)
        if form == "right":
            assert rel(lhs, rhs) < 1e-12
            assert rel(tod(Q, U, A1, A2, form), oracle.convolve(
                NS, 0 * I, Q, U, NB, beams, ra, dec, pa, precision64=True)[0]) < 1e-9
        else:
            assert rel(lhs, rhs) > 1e-2, form


# ---- V-term mode (the reference GPU kernel's V A3, V B3) --------------------

def test_v_term_is_the_I_term_of_the_V_columns(oracle, xbeam, scan):
    """With I = Q = U = 0, the V-term TOD is sum V_p A3[b] (B3 for b): bit for
    bit the CPU Convolver's TOD of the sky I := V with a beam whose A0, B0
    are A3, B3.  With V = None the mode is off (the CPU Convolver's bits)."""
    _, _, _, beams = xbeam
    ra, dec, pa = scan
    rng = np.random.default_rng(51)
    npix = 12 * NS * NS
    V = rng.normal(size=npix).astype(np.float32)
    z = np.zeros(npix, np.float32)
    vbeam = [beams[3], beams[1], beams[2], beams[3], beams[7], beams[5], beams[6], beams[7]]
    ref = oracle.convolve(NS, V, z, z, NB, vbeam, ra, dec, pa)
    oracle.set_v_term(V)
    try:
        got = oracle.convolve(NS, z, z, z, NB, beams, ra, dec, pa)
    finally:
        oracle.set_v_term(None)
    for x, y in zip(got, ref):
        assert np.array_equal(x, y)
    I, Q, U, _ = inputs.iid_sky(NS, 31)
    a0 = oracle.convolve(NS, I, Q, U, NB, beams, ra, dec, pa)
    a1 = oracle.convolve(NS, I, Q, U, NB, beams, ra, dec, pa)
    assert all(np.array_equal(x, y) for x, y in zip(a0, a1))


def test_v_term_linearity_and_circular_beam(oracle, scan):
    """FP64: TOD_V(I, Q, U, V) = TOD(I, Q, U) + TOD_V(0, 0, 0, V) (1e-12).
    A circularly polarised beam (E_cx = i E_co, no leakage) has tilde
    V = tilde I and tilde Q = tilde U = 0, so A3 = A0, A1 = A2 = 0, and a sky
    with V = I sees twice its I TOD (1e-12)."""
    ra, dec, pa = scan
    npix = pypisco.beam_npix_for_radius(NB, CAP * DEG)
    f = inputs.crosspolar_fields(NB, npix, FWHM, 0.05)
    circ = (f[0], f[1], f[0], f[1] + math.pi / 2)
    beams = oracle.polbeam_from_fields(npix, NB, NS, 0.0, circ, circ)
    assert np.allclose(beams[3], beams[0], rtol=1e-6, atol=0)
    assert np.abs(beams[1]).max() <= 1e-6 * beams[0].max()
    assert np.abs(beams[2]).max() <= 1e-6 * beams[0].max()
    I, Q, U, _ = inputs.iid_sky(NS, 33)
    V = np.random.default_rng(52).normal(size=I.size).astype(np.float32)
    z = np.zeros_like(I)
    xb = inputs.crosspolar_polbeam(NB, CAP, FWHM, 0.05, 0.01, nside_sky=NS)
    xbl = list(xb.aBeams) + list(xb.bBeams)
    base = oracle.convolve(NS, I, Q, U, NB, xbl, ra, dec, pa, precision64=True)
    oracle.set_v_term(V)
    try:
        full = oracle.convolve(NS, I, Q, U, NB, xbl, ra, dec, pa, precision64=True)
        vonly = oracle.convolve(NS, z, z, z, NB, xbl, ra, dec, pa, precision64=True)
    finally:
        oracle.set_v_term(None)
    for k in range(2):
        assert rel(full[k], base[k] + vonly[k]) < 1e-12
        assert np.abs(vonly[k]).max() > 1e-3 * np.abs(base[k]).max()   # not empty
    # circular beam, V = I
    twice = oracle.convolve(NS, I, z, z, NB, beams, ra, dec, pa, precision64=True)
    oracle.set_v_term(I)
    try:
        both = oracle.convolve(NS, I, z, z, NB, beams, ra, dec, pa, precision64=True)
    finally:
        oracle.set_v_term(None)
    for k in range(2):
        assert rel(both[k], 2 * twice[k]) < 1e-6   # A3 = A0 to float rounding
