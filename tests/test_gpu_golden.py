"""The GPU paths against the committed oracle fixtures
(tests/golden/oracle_small.npz, inputs included): forward TOD in the
nearest-pixel and interpolation modes and the sky adjoint within relative
RMS 1e-6, the cross-polar case's forward and both adjoints likewise, SkyMap
sums and maps and the focal-plane pointing bit for bit."""
import os

import numpy as np
import pytest

from pisco4cmb_amd import pypisco

pytestmark = pytest.mark.gpu
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_small.npz")


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIX))


class _Beam:
    """A built PolBeam's maps as update_beam reads them."""

    def __init__(self, nside, beams):
        self.nside = nside
        self.aBeams = [np.ascontiguousarray(beams[k], np.float32) for k in range(4)]
        self.bBeams = [np.ascontiguousarray(beams[4 + k], np.float32) for k in range(4)]
        self.nPixels = self.aBeams[0].size

    def get_nside(self):
        return self.nside

    def size(self):
        return self.aBeams[0].size


@pytest.mark.parametrize("interp", [0, 1])
def test_convolver_fixture(fx, interp):
    ns, nb = int(fx["c_nside"]), int(fx["c_nside_beam"])
    n = fx["c_ra"].size
    conv = pypisco.GPUConvolver(pypisco.default_run_config())
    conv.configure_execution()
    conv.set_tuning("beam_interpolation", interp)
    conv.update_sky(pypisco.Sky(ns, fx["c_I"], fx["c_Q"], fx["c_U"]))
    conv.update_beam(_Beam(nb, fx["c_beams"]))
    conv.update_scan(pypisco.Scan(n, fx["c_ra"], fx["c_dec"], fx["c_pa"]))
    ta = np.zeros(n, np.float32)
    tb = np.zeros(n, np.float32)
    conv.exec_convolution(ta, tb)
    sfx = "_interp" if interp else ""
    assert rel_rms(ta, fx["c_tod_a" + sfx]) < 1e-6
    assert rel_rms(tb, fx["c_tod_b" + sfx]) < 1e-6
    if not interp:
        gI, gQ, gU = conv.exec_adjoint_map(fx["c_g"][0], fx["c_g"][1])
        for g, k in ((gI, "c_gI"), (gQ, "c_gQ"), (gU, "c_gU")):
            assert rel_rms(g, fx[k]) < 1e-6
    conv.close()


def test_crosspolar_convolver_fixture(fx):
    """Cross-polar beams (every beam map non-zero): forward TOD, sky adjoint
    and beam adjoint against the frozen oracle outputs (relative RMS 1e-6)."""
    ns, nb = int(fx["x_nside"]), int(fx["x_nside_beam"])
    n = fx["x_ra"].size
    conv = pypisco.GPUConvolver(pypisco.default_run_config())
    conv.configure_execution()
    conv.update_sky(pypisco.Sky(ns, fx["x_I"], fx["x_Q"], fx["x_U"]))
    conv.update_beam(_Beam(nb, fx["x_beams"]))
    conv.update_scan(pypisco.Scan(n, fx["x_ra"], fx["x_dec"], fx["x_pa"]))
    ta = np.zeros(n, np.float32)
    tb = np.zeros(n, np.float32)
    conv.exec_convolution(ta, tb)
    assert rel_rms(ta, fx["x_tod_a"]) < 1e-6 and rel_rms(tb, fx["x_tod_b"]) < 1e-6
    gI, gQ, gU = conv.exec_adjoint_map(fx["x_g"][0], fx["x_g"][1])
    for g, k in ((gI, "x_gI"), (gQ, "x_gQ"), (gU, "x_gU")):
        assert rel_rms(g, fx[k]) < 1e-6
    oa, ob = conv.exec_adjoint_beam(fx["x_g"][0], fx["x_g"][1])
    for k in range(3):
        assert rel_rms(oa[k], fx["x_gbeam"][:, k]) < 1e-6
        assert rel_rms(ob[k], fx["x_gbeam"][:, 3 + k]) < 1e-6
    conv.close()


def test_skymap_fixture(fx):
    ndet, n = fx["m_data"].shape
    m = pypisco.SkyMap(int(fx["m_nside"]))
    m.accumulate_data(n, fx["m_ra"], fx["m_dec"], fx["m_pa"], ndet, fx["m_pol"], fx["m_data"])
    m.solve_map()
    AtA, AtD = m.get_normal()
    assert np.array_equal(AtA, fx["m_AtA"]) and np.array_equal(AtD, fx["m_AtD"])
    assert np.array_equal(m.get_hitmap(), fx["m_hits"])
    for got, k in ((m.get_stokes_I(), "m_I"), (m.get_stokes_Q(), "m_Q"),
                   (m.get_stokes_U(), "m_U")):
        assert np.array_equal(got.view(np.int32), fx[k].view(np.int32))
    m.close()


def test_focal_plane_fixture(fx):
    nd, n = fx["f_dx"].size, fx["f_x"].size
    out = [np.empty(nd * n) for _ in range(3)]
    bp = pypisco.BPoint()
    bp.focal_plane_to_sky(nd, fx["f_dx"], fx["f_dy"], fx["f_pol"], n, fx["f_x"], fx["f_y"],
                          fx["f_rot"], 1, *out)
    for got, k in zip(out, ("f_xdet", "f_ydet", "f_rotdet")):
        assert np.array_equal(got, fx[k].ravel())
    bp.close()
