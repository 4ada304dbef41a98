"""Writes tests/golden/oracle_small.npz: small seeded input/output vectors of
the CPU oracle (oracle/pisco_oracle.c), inputs included, so that the GPU
paths can be compared with fixed vectors and any change of the oracle shows
as a mismatch:
  - Convolver forward (nearest beam pixel and interpolation modes) of a
    detector pair, sky NSIDE 16, beam NSIDE 128 (cap 10 deg, FWHM 4 deg),
    300 raster samples, and the sky adjoint for fixed cotangents;
  - the same with cross-polar beams (PolBeam::beam_from_fields of fields with
    a cross-polar component, so every beam map is non-zero): forward, sky
    adjoint and beam adjoint;
  - SkyMap accumulate + solve of 4 detectors x 2000 samples into NSIDE 8;
  - BPoint::focal_plane_to_sky of 5 detectors x 100 samples (ra/dec form).
The oracle's own pins (tests/test_oracle_*.py) check these functions; the
fixtures freeze their outputs.

    python tests/golden/make_oracle_fixtures.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle_bind  # noqa: E402

from pisco4cmb_amd import inputs, pypisco  # noqa: E402

DEG = math.pi / 180


def convolver_case():
    ns, nb = 16, 128
    I, Q, U, _ = inputs.iid_sky(ns, seed=101)
    npb = pypisco.beam_npix_for_radius(nb, 10.0 * DEG)
    b = pypisco.PolBeam(nb, npb, 0.01, "p")
    b.make_unpol_gaussian_elliptical_beams(4.0, 3.6, 15.0)
    b.build(ns)
    beams = np.stack(list(b.aBeams) + list(b.bBeams))
    ra, dec, pa = inputs.raster_scan(300, size_deg=40.0)
    g = np.random.default_rng(102).normal(size=(2, 300)).astype(np.float32)
    return dict(c_nside=ns, c_nside_beam=nb, c_I=I, c_Q=Q, c_U=U, c_beams=beams, c_ra=ra,
                c_dec=dec, c_pa=pa, c_g=g)


def xconvolver_case(o):
    """Cross-polar pair: fields from inputs.crosspolar_fields, beams built by
    the oracle's beam_from_fields + build (polbeam.cpp:117-288)."""
    ns, nb = 16, 128
    I, Q, U, _ = inputs.iid_sky(ns, seed=111)
    npb = pypisco.beam_npix_for_radius(nb, 10.0 * DEG)
    fa = inputs.crosspolar_fields(nb, npb, 4.0, 0.1, 0.0, 0.35, 0.6, 0.3)
    fb = inputs.crosspolar_fields(nb, npb, 4.0, 0.1, 30.0, 0.25, -0.9, 1.2)
    beams = np.stack(o.polbeam_from_fields(npb, nb, ns, 0.01, fa, fb))
    ra, dec, pa = inputs.raster_scan(300, size_deg=40.0, center_dec=0.4)
    g = np.random.default_rng(112).normal(size=(2, 300)).astype(np.float32)
    return dict(x_nside=ns, x_nside_beam=nb, x_I=I, x_Q=Q, x_U=U, x_beams=beams, x_ra=ra,
                x_dec=dec, x_pa=pa, x_g=g)


def skymap_case():
    rng = np.random.default_rng(103)
    ndet, n = 4, 2000
    ra = rng.uniform(0, 2 * math.pi, (ndet, n)).astype(np.float32)
    dec = np.arcsin(rng.uniform(-1, 1, (ndet, n))).astype(np.float32)
    pa = rng.uniform(-math.pi, math.pi, (ndet, n)).astype(np.float32)
    pol = np.array([0.0, 0.7, 1.4, 2.1], np.float32)
    data = rng.normal(0, 10, (ndet, n)).astype(np.float32)
    return dict(m_nside=8, m_ra=ra, m_dec=dec, m_pa=pa, m_pol=pol, m_data=data)


def focal_case():
    rng = np.random.default_rng(104)
    return dict(f_dx=rng.uniform(-2, 2, 5) * DEG, f_dy=rng.uniform(-2, 2, 5) * DEG,
                f_pol=rng.uniform(-1, 1, 5), f_x=rng.uniform(0, 2 * math.pi, 100),
                f_y=rng.uniform(-1.2, 1.2, 100), f_rot=rng.uniform(-math.pi, math.pi, 100))


def outputs(o, f):
    out = {}
    beams = list(f["c_beams"])
    args = (int(f["c_nside"]), f["c_I"], f["c_Q"], f["c_U"], int(f["c_nside_beam"]), beams,
            f["c_ra"], f["c_dec"], f["c_pa"])
    out["c_tod_a"], out["c_tod_b"] = o.convolve(*args, nthreads=1)
    gI, gQ, gU = o.adjoint_map(int(f["c_nside"]), int(f["c_nside_beam"]), beams, f["c_ra"],
                               f["c_dec"], f["c_pa"], f["c_g"][0], f["c_g"][1])
    out["c_gI"], out["c_gQ"], out["c_gU"] = gI, gQ, gU
    xb = list(f["x_beams"])
    xargs = (int(f["x_nside"]), f["x_I"], f["x_Q"], f["x_U"], int(f["x_nside_beam"]), xb,
             f["x_ra"], f["x_dec"], f["x_pa"])
    out["x_tod_a"], out["x_tod_b"] = o.convolve(*xargs, nthreads=1)
    gI, gQ, gU = o.adjoint_map(int(f["x_nside"]), int(f["x_nside_beam"]), xb, f["x_ra"],
                               f["x_dec"], f["x_pa"], f["x_g"][0], f["x_g"][1])
    out["x_gI"], out["x_gQ"], out["x_gU"] = gI, gQ, gU
    out["x_gbeam"] = o.adjoint_beam(*xargs, f["x_g"][0], f["x_g"][1])
    o.set_beam_interpolation(True)
    try:
        out["c_tod_a_interp"], out["c_tod_b_interp"] = o.convolve(*args, nthreads=1)
    finally:
        o.set_beam_interpolation(False)
    AtA, AtD, _ = o.skymap_accumulate(int(f["m_nside"]), f["m_ra"], f["m_dec"], f["m_pa"],
                                      f["m_pol"], f["m_data"])
    I, Q, U, hits, _ = o.skymap_solve(int(f["m_nside"]), AtA, AtD)
    out.update(m_AtA=AtA, m_AtD=AtD, m_I=I, m_Q=Q, m_U=U, m_hits=hits)
    X, Y, R = o.focal_plane_to_sky(f["f_dx"], f["f_dy"], f["f_pol"], f["f_x"], f["f_y"],
                                   f["f_rot"], 1)
    out.update(f_xdet=X, f_ydet=Y, f_rotdet=R)
    return out


def main():
    o = oracle_bind.Oracle()
    f = {}
    f.update(convolver_case())
    f.update(xconvolver_case(o))
    f.update(skymap_case())
    f.update(focal_case())
    f.update(outputs(o, f))
    path = os.path.join(HERE, "oracle_small.npz")
    np.savez_compressed(path, **f)
    print("wrote %s (%d arrays, %d bytes)" % (path, len(f), os.path.getsize(path)))


if __name__ == "__main__":
    main()
