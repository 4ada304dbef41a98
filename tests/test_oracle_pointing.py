"""CPU: the oracle's BPoint::focal_plane_to_sky restatement
(src/Bpoint/bpoint.cpp:553-674) pinned without the reference build (it needs
SOFA): the zero offset returns the array centre; an independent 3-vector
construction of the focal plane gives every detector's position and
position angle; az/el and source-relative outputs relate to ra/dec as
documented; and a glibc build of the same restatement agrees to rounding."""
import math

import numpy as np

DEG = math.pi / 180


def unit(ra, dec):
    return np.stack([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)], -1)


def frame(ra, dec):
    """North and East unit tangents at (ra, dec)."""
    north = np.stack([-np.sin(dec) * np.cos(ra), -np.sin(dec) * np.sin(ra), np.cos(dec)], -1)
    east = np.stack([-np.sin(ra), np.cos(ra), np.zeros_like(ra)], -1)
    return north, east


def wrap(a):
    return (a + math.pi) % (2 * math.pi) - math.pi


def scan(rng, n):
    x = rng.uniform(0, 2 * math.pi, n)
    y = rng.uniform(-80, 80, n) * DEG
    rot = rng.uniform(-math.pi, math.pi, n)
    return x, y, rot


def test_zero_offset_is_the_array_centre(oracle):
    rng = np.random.default_rng(1)
    x, y, rot = scan(rng, 5000)
    X, Y, R = oracle.focal_plane_to_sky([0.0], [0.0], [0.0], x, y, rot, 1)
    assert np.array_equal(X[0], x)
    assert np.abs(Y[0] - y).max() < 4e-16
    assert np.abs(wrap(R[0] - rot)).max() < 1e-15


def test_three_vector_focal_plane(oracle):
    """Detector direction d = cos dx cos dy b + sin dx cos dy r + sin dy u,
    with b the boresight, u the receiver's up axis at position angle rot
    (North through East) and r its right axis (position angle rot - 90 deg,
    West for rot = 0: the sky seen looking out).  The detector's position
    angle is that of the great circle from d towards u, plus its
    polarisation angle."""
    rng = np.random.default_rng(2)
    nd, n = 40, 400
    dx = rng.uniform(-3, 3, nd) * DEG
    dy = rng.uniform(-3, 3, nd) * DEG
    pol = rng.uniform(-math.pi, math.pi, nd)
    x, y, rot = scan(rng, n)
    X, Y, R = oracle.focal_plane_to_sky(dx, dy, pol, x, y, rot, 1)
    b = unit(x, y)
    north, east = frame(x, y)
    u = np.cos(rot)[:, None] * north + np.sin(rot)[:, None] * east
    r = np.sin(rot)[:, None] * north - np.cos(rot)[:, None] * east
    for k in range(nd):
        d = (math.cos(dx[k]) * math.cos(dy[k]) * b + math.sin(dx[k]) * math.cos(dy[k]) * r
             + math.sin(dy[k]) * u)
        ra_d = np.arctan2(d[:, 1], d[:, 0])
        dec_d = np.arcsin(np.clip(d[:, 2], -1, 1))
        assert np.abs(wrap(X[k] - ra_d)).max() < 1e-12
        assert np.abs(Y[k] - dec_d).max() < 1e-12
        nd_, ed_ = frame(ra_d, dec_d)
        t = u - np.sum(u * d, -1, keepdims=True) * d
        pa = np.arctan2(np.sum(t * ed_, -1), np.sum(t * nd_, -1))
        assert np.abs(wrap(R[k] - pol[k] - pa)).max() < 1e-11
        # the offset radius: cos r = cos dx cos dy
        cosr = np.sum(unit(X[k], Y[k]) * b, -1)
        assert np.abs(cosr - math.cos(dx[k]) * math.cos(dy[k])).max() < 1e-13


def test_azel_and_source_relative_forms(oracle):
    rng = np.random.default_rng(3)
    dx = rng.uniform(-2, 2, 8) * DEG
    dy = rng.uniform(-2, 2, 8) * DEG
    pol = rng.uniform(-1, 1, 8)
    x, y, rot = scan(rng, 300)
    X1, Y1, R1 = oracle.focal_plane_to_sky(dx, dy, pol, x, y, rot, 1)
    X0, Y0, R0 = oracle.focal_plane_to_sky(dx, dy, pol, x, y, rot, 0)
    X2, Y2, R2 = oracle.focal_plane_to_sky(dx, dy, pol, x, y, rot, 2)
    assert np.array_equal(Y0, Y1) and np.array_equal(Y2, Y1)
    assert np.array_equal(R0, R1) and np.array_equal(R2, R1)
    assert np.abs((X0 - x) + (X1 - x)).max() < 1e-15
    assert np.abs(X2 + (R1 - pol[:, None])).max() < 1e-15   # x = -pa_det, rot = pa_det + pol


def test_glibc_build_agrees(oracle, oracle_glibc):
    rng = np.random.default_rng(4)
    dx = rng.uniform(-3, 3, 16) * DEG
    dy = rng.uniform(-3, 3, 16) * DEG
    pol = rng.uniform(-1, 1, 16)
    x, y, rot = scan(rng, 2000)
    A = oracle.focal_plane_to_sky(dx, dy, pol, x, y, rot, 1)
    B = oracle_glibc.focal_plane_to_sky(dx, dy, pol, x, y, rot, 1)
    for a, b in zip(A, B):
        assert np.abs(a - b).max() < 1e-14
