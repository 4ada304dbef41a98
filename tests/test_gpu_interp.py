"""Beam interpolation mode on the GPU (tuning key beam_interpolation)
against the CPU oracle's interpolation mode: forward TOD within relative
RMS 1e-6, sky and beam adjoints within relative RMS 1e-6 of the oracle's
FP64 adjoints, and the dot-product identity with the GPU forward to 1e-6.
The batched launches refuse the mode."""
import ctypes
import math

import numpy as np
import pytest

from pisco4cmb_amd import inputs, pypisco
from pisco4cmb_amd._lib import PiscoError

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


@pytest.fixture
def interp_oracle(oracle):
    oracle.set_beam_interpolation(True)
    yield oracle
    oracle.set_beam_interpolation(False)


@pytest.mark.parametrize("kind", ["copolar", "crosspolar"])
@pytest.mark.parametrize("nside,nside_beam,cap,fwhm,n", [(64, 512, 4.0, 1.5, 3000),
                                                        (128, 1024, 2.0, 0.7, 3000)])
def test_interpolated_forward_and_adjoints(interp_oracle, nside, nside_beam, cap, fwhm, n, kind):
    o = interp_oracle
    sky = inputs.iid_sky(nside, seed=31)
    beam = inputs.make_polbeam(kind, nside_beam, cap, fwhm, 0.08, 0.02, nside_sky=nside)
    ra, dec, pa = inputs.raster_scan(n, size_deg=15.0)
    conv = pypisco.GPUConvolver(pypisco.default_run_config(ptgPerConv=1000))
    conv.configure_execution()
    conv.set_tuning("beam_interpolation", 1)
    assert conv.get_tuning("beam_interpolation") == 1
    conv.update_sky(pypisco.Sky(nside, *sky[:3]))
    conv.update_beam(beam)
    conv.update_scan(pypisco.Scan(n, ra, dec, pa))
    ta = np.zeros(n, np.float32)
    tb = np.zeros(n, np.float32)
    conv.exec_convolution(ta, tb)
    beams = list(beam.aBeams) + list(beam.bBeams)
    oa, ob = o.convolve(nside, *sky[:3], nside_beam, beams, ra, dec, pa, nthreads=8)
    assert rel_rms(ta, oa) < 1e-6 and rel_rms(tb, ob) < 1e-6
    # the nearest-pixel TOD differs (the mode is in effect)
    o.set_beam_interpolation(False)
    na, _ = o.convolve(nside, *sky[:3], nside_beam, beams, ra, dec, pa, nthreads=8)
    o.set_beam_interpolation(True)
    assert rel_rms(ta, na) > 1e-4

    rng = np.random.default_rng(32)
    ga, gb = (rng.normal(size=n).astype(np.float32) for _ in range(2))
    gI, gQ, gU = conv.exec_adjoint_map(ga, gb)
    rI, rQ, rU = o.adjoint_map(nside, nside_beam, beams, ra, dec, pa, ga, gb)
    for g, r in ((gI, rI), (gQ, rQ), (gU, rU)):
        assert rel_rms(g, r) < 1e-6
    lhs = float(ta.astype(np.float64) @ ga + tb.astype(np.float64) @ gb)
    rhs = float(sum(g.astype(np.float64) @ x.astype(np.float64) for g, x in
                    ((gI, sky[0]), (gQ, sky[1]), (gU, sky[2]))))
    assert abs(lhs - rhs) < 1e-6 * abs(lhs)
    gab = conv.exec_adjoint_beam(ga, gb)
    rb = o.adjoint_beam(nside, *sky[:3], nside_beam, beams, ra, dec, pa, ga, gb)
    got = np.stack([gab[0][k] for k in range(3)] + [gab[1][k] for k in range(3)], 1)
    assert rel_rms(got, rb) < 1e-6
    conv.close()


def test_batched_launch_refuses_interpolation():
    import torch
    from pisco4cmb_amd.batched import DetectorBatch
    ns = 32
    beam = inputs.gaussian_polbeam(256, 4.0, 1.5, nside_sky=ns)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    batch = DetectorBatch([(t(np.stack(beam.aBeams)), t(np.stack(beam.bBeams)), 256, 0.0)])
    batch.set_tuning("beam_interpolation", 1)
    try:
        sky = [torch.zeros(12 * ns * ns, device="cuda") for _ in range(3)]
        ra, dec, pa = (t(x) for x in inputs.raster_scan(10))
        with pytest.raises(PiscoError):
            batch.forward(*sky, ra, dec, pa, ns)
    finally:
        batch.close()
