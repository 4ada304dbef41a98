"""The reference's text formats through a GPU convolution: a sky file of
I Q U V rows per RING pixel (scripts/create_cmb.py:96-97) and tilde-beam
files of tilde I, Q, U, V rows per beam pixel for detectors a and b
(scripts/fields2beams.py:175-176), both written with numpy.savetxt, are
loaded by the product's native readers (pypisco.read_sky_txt,
PolBeam.read_tilde_beams), built (PolBeam.build) and convolved on the GPU.
The oracle gets the same files parsed by numpy and the beams built by its own
PolBeam restatement.  The tilde beams come from cross-polar fields, so every
beam map is non-zero.  Checks: forward TOD with and without the V term and
the sky adjoint, relative RMS 1e-6."""
import math

import numpy as np
import pytest

from pisco4cmb_amd import inputs, pypisco

pytestmark = pytest.mark.gpu
DEG = math.pi / 180


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def tilde_from_fields(f):
    """tilde I, Q, U, V in float64 from (magEco, phaseEco, magEcx, phaseEcx),
    as fields2beams.py forms them (fields2beams.py:63-66)."""
    co = f[0] * np.exp(1j * f[1])
    cx = f[2] * np.exp(1j * f[3])
    p = co * np.conj(cx)
    return np.abs(co) ** 2 + np.abs(cx) ** 2, np.abs(co) ** 2 - np.abs(cx) ** 2, \
        2 * p.real, -2 * p.imag


def test_reference_text_files_through_gpu(tmp_path, oracle):
    ns, nb, cap, eps = 128, 1024, 3.0, 0.01
    I, Q, U, _ = inputs.powerlaw_sky(ns, seed=91, lmax=128)
    V = np.random.default_rng(92).normal(0, 2.0, I.size)
    sky_path = tmp_path / "cmb.txt"
    np.savetxt(sky_path, np.vstack([I, Q, U, V]).T.astype(np.float64))
    npix = pypisco.beam_npix_for_radius(nb, cap * DEG)
    paths = {}
    for det, (xp, xph, tw, phi0) in (("a", (0.35, 0.6, 0.3, 0.0)), ("b", (0.25, -0.9, 1.2, 30.0))):
        f = inputs.crosspolar_fields(nb, npix, 1.0, 0.05, phi0, xp, xph, tw)
        paths[det] = tmp_path / ("hpx_tilde_beams_%s.txt" % det)
        np.savetxt(paths[det], np.vstack(tilde_from_fields(f)).T)

    # product: native readers, PolBeam.build, GPU
    sky = pypisco.read_sky_txt(sky_path)
    assert sky.get_nside() == ns
    beam = pypisco.PolBeam(nb, npix, eps, "p")
    beam.read_tilde_beams(paths["a"], "a")
    beam.read_tilde_beams(paths["b"], "b")
    beam.build(ns)
    ra, dec, pa = inputs.ces_scan(3000, t0=55.0)

    # oracle: numpy's parse, the oracle's PolBeam restatement
    sk = np.loadtxt(sky_path).astype(np.float32)
    tilde = [np.loadtxt(paths[d]).astype(np.float32) for d in "ab"]
    obeams = oracle.polbeam_build(npix, nb, ns, eps, "p",
                                  [tilde[0][:, k] for k in range(4)] +
                                  [tilde[1][:, k] for k in range(4)])
    assert np.abs(obeams[2]).max() >= 0.1 * np.abs(obeams[1]).max()
    for k in range(4):   # the two builds of the same files agree closely
        assert rel_rms(beam.aBeams[k], obeams[k]) < 1e-6
        assert rel_rms(beam.bBeams[k], obeams[4 + k]) < 1e-6

    for vterm in (0, 1):
        conv = pypisco.GPUConvolver(pypisco.default_run_config())
        conv.configure_execution()
        conv.set_tuning("v_term", vterm)
        conv.update_sky(sky)
        conv.update_beam(beam)
        conv.update_scan(pypisco.Scan(ra.size, ra, dec, pa))
        a = np.zeros(ra.size, np.float32)
        b = np.zeros(ra.size, np.float32)
        conv.exec_convolution(a, b)
        oracle.set_v_term(sk[:, 3] if vterm else None)
        try:
            oa, ob = oracle.convolve(ns, sk[:, 0], sk[:, 1], sk[:, 2], nb, obeams, ra, dec, pa)
        finally:
            oracle.set_v_term(None)
        assert rel_rms(a, oa) < 1e-6 and rel_rms(b, ob) < 1e-6, vterm
        if vterm == 0:
            rng = np.random.default_rng(93)
            ga, gb = (rng.normal(size=ra.size).astype(np.float32) for _ in range(2))
            got = conv.exec_adjoint_map(ga, gb)
            ref = oracle.adjoint_map(ns, nb, obeams, ra, dec, pa, ga, gb)
            for x, y in zip(got, ref):
                assert rel_rms(x, y) < 1e-6
        conv.close()
