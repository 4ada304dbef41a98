"""Seeded synthetic inputs for the convolution workloads.

Skies (RING HEALPix, float32):
  iid_sky            I ~ N(0, sigma_I^2), Q,U ~ N(0, sigma_P^2) per pixel
  powerlaw_sky       Gaussian random field with C_l proportional to l^-2
                     (l >= 2), synthesised ring by ring: an associated-Legendre
                     recursion in l for every m (vectorised over rings) and an
                     inverse FFT per ring; I, Q, U are independent scalar fields
  add_point_sources  point sources of log-uniform amplitude
Scans (float32 radians, pa counter-clockwise):
  raster_scan        rows x columns over a square patch, pa cycling a list
  ces_scan           constant-elevation scan at a fixed site: azimuth spin,
                     sidereal drift (ra, dec, parallactic angle)
Beams:
  gaussian_polbeam   co-polar elliptical Gaussian pair, built like
                     polbeam.cpp:290-372 + build() (Q/U leakage only through
                     epsilon: aBeams[2] = bBeams[2] = 0)
  crosspolar_polbeam a pair whose E fields carry a cross-polar component with
                     its own phase (PolBeam::beam_from_fields,
                     polbeam.cpp:117-184), as GRASP beams do: tilde U, tilde V
                     and so aBeams[2], bBeams[2], aBeams[3], bBeams[3] are
                     non-zero, and the a and b fields differ

Everything here is set-up code; the convolution never depends on it.
"""
import math

import numpy as np

from . import pypisco

DEG = math.pi / 180.0


# ---------------------------------------------------------------- skies ----

def iid_sky(nside, seed, sigma_i=10.0, sigma_p=math.sqrt(10.0)):
    """I ~ N(0, sigma_i^2), Q, U ~ N(0, sigma_p^2), V = 0 (defaults: 100 uK^2
    and 10 uK^2 variances)."""
    rng = np.random.default_rng(seed)
    n = 12 * nside * nside
    I = rng.normal(0.0, sigma_i, n).astype(np.float32)
    Q = rng.normal(0.0, sigma_p, n).astype(np.float32)
    U = rng.normal(0.0, sigma_p, n).astype(np.float32)
    return I, Q, U, np.zeros(n, np.float32)


def _ring_geometry(nside):
    """theta, first pixel, pixel count and phi offset of every ring."""
    rings = np.arange(1, 4 * nside)
    npix = 12 * nside * nside
    ncap = 2 * nside * (nside - 1)
    z = np.empty(rings.size)
    start = np.empty(rings.size, np.int64)
    count = np.empty(rings.size, np.int64)
    shifted = np.empty(rings.size, bool)
    north = rings < nside
    south = rings > 3 * nside
    equ = ~north & ~south
    r = rings[north]
    z[north] = 1.0 - r * r * 4.0 / npix
    start[north] = 2 * r * (r - 1)
    count[north] = 4 * r
    shifted[north] = True
    r = rings[equ]
    z[equ] = (2 * nside - r) * (8.0 * nside / npix)
    start[equ] = ncap + (r - nside) * 4 * nside
    count[equ] = 4 * nside
    shifted[equ] = ((r - nside) & 1) == 0
    r = 4 * nside - rings[south]
    z[south] = r * r * 4.0 / npix - 1.0
    start[south] = npix - 2 * r * (r + 1)
    count[south] = 4 * r
    shifted[south] = True
    theta = np.arccos(np.clip(z, -1, 1))
    phi0 = np.where(shifted, np.pi / count, 0.0)
    return theta, start, count, phi0


def _synthesize_rings(nside, alm_fn, lmax):
    """Real map from a_lm (alm_fn(m) -> complex array over l = m..lmax)."""
    theta, start, count, phi0 = _ring_geometry(nside)
    x = np.cos(theta)
    s = np.sin(theta)
    logs = np.log(np.maximum(s, 1e-300))
    nr = theta.size
    F = np.zeros((lmax + 1, nr), np.complex128)     # F_m(theta_r)
    logfac = 0.0                                     # sum_{k<=m} log((2k-1)/(2k))
    for m in range(lmax + 1):
        if m > 0:
            logfac += math.log((2 * m - 1) / (2 * m))
        alm = alm_fn(m)
        # lambda_mm with the Condon-Shortley phase
        lmm = np.exp(0.5 * math.log((2 * m + 1) / (4 * math.pi)) + 0.5 * logfac + m * logs)
        if m % 2:
            lmm = -lmm
        acc = alm[0] * lmm
        if m < lmax:
            lp = lmm
            lc = x * math.sqrt(2 * m + 3) * lmm
            acc = acc + alm[1] * lc
            for ell in range(m + 2, lmax + 1):
                a = math.sqrt((4.0 * ell * ell - 1) / (ell * ell - m * m))
                b = math.sqrt(((ell - 1) ** 2 - m * m) / (4.0 * (ell - 1) ** 2 - 1))
                ln = a * (x * lc - b * lp)
                acc = acc + alm[ell - m] * ln
                lp, lc = lc, ln
        F[m] = acc
    out = np.empty(12 * nside * nside, np.float64)
    for k in range(nr):
        n = int(count[k])
        buf = np.zeros(n, np.complex128)
        ms = np.arange(lmax + 1)
        coef = F[:, k] * np.exp(1j * ms * phi0[k])
        coef[1:] *= 2.0                      # m > 0 and m < 0 together
        np.add.at(buf, ms % n, coef)
        vals = np.fft.ifft(buf) * n
        out[start[k]:start[k] + n] = vals.real
    return out


def powerlaw_field(nside, seed, rms, lmax=None, slope=-2.0):
    """Gaussian random field, C_l = A l^slope for l >= 2, rescaled to the
    requested pixel rms."""
    if lmax is None:
        lmax = min(3 * nside - 1, 1024)
    rng = np.random.default_rng(seed)
    ell = np.arange(lmax + 1, dtype=np.float64)
    cl = np.zeros(lmax + 1)
    cl[2:] = ell[2:] ** slope
    amp = {}

    def alm_fn(m):
        if m not in amp:
            n = lmax + 1 - m
            sd = np.sqrt(cl[m:])
            if m == 0:
                amp[m] = rng.normal(size=n) * sd + 0j
            else:
                amp[m] = (rng.normal(size=n) + 1j * rng.normal(size=n)) * sd / math.sqrt(2)
        return amp[m]

    f = _synthesize_rings(nside, alm_fn, lmax)
    f -= f.mean()
    f *= rms / max(f.std(), 1e-30)
    return f


def powerlaw_sky(nside, seed, rms_i=100.0, rms_p=10.0, lmax=None):
    """I, Q, U independent power-law fields (uK), V = 0."""
    I = powerlaw_field(nside, seed, rms_i, lmax).astype(np.float32)
    Q = powerlaw_field(nside, seed + 1, rms_p, lmax).astype(np.float32)
    U = powerlaw_field(nside, seed + 2, rms_p, lmax).astype(np.float32)
    return I, Q, U, np.zeros_like(I)


def add_point_sources(I, nside, seed, n=200, amp_min=1e2, amp_max=1e4):
    """Adds n single-pixel sources of log-uniform amplitude to I (in place)."""
    rng = np.random.default_rng(seed)
    npix = 12 * nside * nside
    pix = rng.choice(npix, size=n, replace=False)
    amp = np.exp(rng.uniform(math.log(amp_min), math.log(amp_max), n))
    I[pix] += amp.astype(np.float32)
    return pix, amp


# ---------------------------------------------------------------- scans ----

def raster_scan(nsamples=10000, center_ra=math.pi, center_dec=0.0, size_deg=20.0,
                pa_cycle_deg=(0.0, 45.0, 90.0, 135.0)):
    """Raster over a size x size degree patch: rows of constant dec, ra
    sweeping; pa cycles through pa_cycle_deg sample by sample."""
    nrow = int(round(math.sqrt(nsamples)))
    ncol = int(math.ceil(nsamples / nrow))
    half = 0.5 * size_deg * DEG
    dec_rows = center_dec + np.linspace(-half, half, nrow)
    ra_cols = center_ra + np.linspace(-half, half, ncol)
    k = np.arange(nsamples)
    row = k // ncol
    col = k % ncol
    col = np.where(row % 2 == 0, col, ncol - 1 - col)     # boustrophedon
    dec = dec_rows[row]
    ra = ra_cols[col]
    pa = np.asarray(pa_cycle_deg)[k % len(pa_cycle_deg)] * DEG
    return (ra.astype(np.float32), dec.astype(np.float32), pa.astype(np.float32))


def ces_scan(nsamples, fs=100.0, lat_deg=-23.0, el_deg=45.0, rpm=1.0, t0=0.0,
             lst0=0.0):
    """Constant-elevation scan: azimuth spins at rpm, the sky drifts at the
    sidereal rate; returns ra in [0, 2pi), dec, and the parallactic angle as
    pa.  At lat -23 deg, el 45 deg, dec spans about [-68, +22] deg."""
    t = t0 + np.arange(nsamples, dtype=np.float64) / fs
    az = 2 * np.pi * rpm / 60.0 * t
    el = el_deg * DEG
    lat = lat_deg * DEG
    sd = np.sin(el) * np.sin(lat) + np.cos(el) * np.cos(lat) * np.cos(az)
    dec = np.arcsin(np.clip(sd, -1, 1))
    ha = np.arctan2(-np.sin(az) * np.cos(el),
                    np.sin(el) * np.cos(lat) - np.cos(el) * np.sin(lat) * np.cos(az))
    sidereal = 2 * np.pi / 86164.0905
    lst = lst0 + sidereal * t
    ra = np.mod(lst - ha, 2 * np.pi)
    pa = np.arctan2(np.sin(ha), np.tan(lat) * np.cos(dec) - np.sin(dec) * np.cos(ha))
    ra32 = ra.astype(np.float32)
    ra32[ra32 >= np.float32(2 * np.pi)] = 0.0
    return ra32, dec.astype(np.float32), pa.astype(np.float32)


# ---------------------------------------------------------------- beams ----

def gaussian_polbeam(nside_beam, cap_deg, fwhm_deg, ellipticity=0.0, epsilon=0.0,
                     nside_sky=None, dets="p", phi0_deg=0.0):
    """PolBeam of an elliptical Gaussian pair: fwhm_x = fwhm, fwhm_y =
    fwhm * (1 - ellipticity); nPixels = ang2pix(nside_beam, cap, 0)."""
    npix = pypisco.beam_npix_for_radius(nside_beam, cap_deg * DEG)
    b = pypisco.PolBeam(nside_beam, npix, epsilon, dets)
    b.make_unpol_gaussian_elliptical_beams(fwhm_deg, fwhm_deg * (1.0 - ellipticity),
                                           phi0_deg)
    b.build(nside_sky if nside_sky is not None else nside_beam)
    return b


def beam_pixel_angles(nside_beam, npix):
    """(theta, phi) in float64 of the RING pixels [0, npix) of a map of NSIDE
    nside_beam (vectorised; for building field models, not for indexing)."""
    theta_r, start, count, phi0 = _ring_geometry(nside_beam)
    p = np.arange(int(npix), dtype=np.int64)
    r = np.searchsorted(start, p, side="right") - 1
    theta = theta_r[r]
    phi = phi0[r] + 2.0 * np.pi * (p - start[r]) / count[r]
    return theta, phi


def crosspolar_fields(nside_beam, npix, fwhm_deg, ellipticity=0.0, phi0_deg=0.0, xpol=0.35,
                      xphase=0.6, twist=0.3, curvature=0.3):
    """Co- and cross-polar E fields (magnitude, phase) of a feed over the beam
    pixels, in the form PolBeam::beam_from_fields takes (polbeam.cpp:117-143).

    The co-polar magnitude is the square root of the elliptical Gaussian of
    make_unpol_gaussian_elliptical_beams (polbeam.cpp:332-357) with a phase
    that grows as (rho/sigma_x)^2.  The cross-polar magnitude is xpol times
    the co-polar one, modulated as 0.4 + 0.6 cos^2(sigma - twist), with phase
    xphase + 0.5 sin(sigma) + curvature/2 (rho/sigma_x)^2.  Returns four
    float64 arrays magEco, phaseEco, magEcx, phaseEcx."""
    theta, phi = beam_pixel_angles(nside_beam, npix)
    sx = fwhm_deg * DEG / 2.35482
    sy = fwhm_deg * (1.0 - ellipticity) * DEG / 2.35482
    p0 = phi0_deg * DEG
    a = math.cos(p0) ** 2 / (2 * sx * sx) + math.sin(p0) ** 2 / (2 * sy * sy)
    b = -math.sin(2 * p0) / (4 * sx * sx) + math.sin(2 * p0) / (4 * sy * sy)
    c = math.sin(p0) ** 2 / (2 * sx * sx) + math.cos(p0) ** 2 / (2 * sy * sy)
    cs, sn = np.cos(phi), np.sin(phi)
    g = np.exp(-(a * cs * cs + 2.0 * b * cs * sn + c * sn * sn) * theta * theta)
    u = (theta / sx) ** 2
    mag_co = np.sqrt(g)
    phs_co = curvature * u
    mag_cx = xpol * mag_co * (0.4 + 0.6 * np.cos(phi - twist) ** 2)
    phs_cx = xphase + 0.5 * sn + 0.5 * curvature * u
    return mag_co, phs_co, mag_cx, phs_cx


def crosspolar_polbeam(nside_beam, cap_deg, fwhm_deg, ellipticity=0.0, epsilon=0.0,
                       nside_sky=None, xpol=(0.35, 0.25), xphase=(0.6, -0.9), twist=(0.3, 1.2),
                       phi0_deg=(0.0, 30.0)):
    """PolBeam pair with cross-polar fields: detector a and b each from
    crosspolar_fields with their own (xpol, xphase, twist, phi0) through
    PolBeam.beam_from_fields, then build(); nPixels = ang2pix(nside_beam,
    cap, 0).  Every beam map of both detectors is non-zero, including
    aBeams[2], bBeams[2] (tilde U) and aBeams[3], bBeams[3] (tilde V)."""
    npix = pypisco.beam_npix_for_radius(nside_beam, cap_deg * DEG)
    bm = pypisco.PolBeam(nside_beam, npix, epsilon, "p")
    for k, det in enumerate("ab"):
        f = crosspolar_fields(nside_beam, npix, fwhm_deg, ellipticity, phi0_deg[k], xpol[k],
                              xphase[k], twist[k])
        bm.beam_from_fields(det, *f)
    bm.build(nside_sky if nside_sky is not None else nside_beam)
    return bm


def make_polbeam(kind, nside_beam, cap_deg, fwhm_deg, ellipticity=0.0, epsilon=0.0,
                 nside_sky=None):
    """kind "copolar" (gaussian_polbeam) or "crosspolar" (crosspolar_polbeam)."""
    if kind == "copolar":
        return gaussian_polbeam(nside_beam, cap_deg, fwhm_deg, ellipticity, epsilon,
                                nside_sky=nside_sky)
    if kind == "crosspolar":
        return crosspolar_polbeam(nside_beam, cap_deg, fwhm_deg, ellipticity, epsilon,
                                  nside_sky=nside_sky)
    raise ValueError("beam kind must be 'copolar' or 'crosspolar'")
