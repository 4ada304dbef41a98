"""The deterministic sky adjoint (tuning key adjmap_deterministic): every
contribution rounded once to a power-of-two fixed point and summed as 64-bit
integers, so the output bits do not depend on the schedule.  Checked: the
same bits for different launch shapes (persistent blocks, samples per
window) and repeated calls, for the single-pair, batched and per-pair-pointing
adjoints; within relative RMS 1e-6 of the floating-point adjoint and of the
CPU oracle; the dot-product identity; refused with beam interpolation; the
default (-1) gives the integer sums' bits for single pairs, and with beam
interpolation falls back to float sums instead of refusing."""
import math

import numpy as np
import pytest
import torch

from pisco4cmb_amd import inputs, torch_ops
from pisco4cmb_amd._lib import PiscoError, check, lib
from pisco4cmb_amd.batched import DetectorBatch

pytestmark = pytest.mark.gpu
DEG = math.pi / 180


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


@pytest.fixture
def ctx():
    c = torch_ops._ctx(torch.device("cuda:0"))
    yield c
    for k, v in ((b"adjmap_deterministic", -1), (b"adjmap_blocks", 0), (b"adjmap_chunk", -1)):
        check(lib.pisco_set_tuning(c, k, v))


def adjoint(ctx, beam8, nb, ra, dec, pa, ga, gb, ns):
    npix = 12 * ns * ns
    out = [torch.empty(npix, device="cuda") for _ in range(3)]
    check(lib.pisco_dev_adjoint_map(ctx, torch_ops._stream(torch.device("cuda:0")), ns, nb,
                                    beam8.shape[0], beam8.data_ptr(), ra.numel(), ra.data_ptr(),
                                    dec.data_ptr(), pa.data_ptr(), ga.data_ptr(), gb.data_ptr(),
                                    *[o.data_ptr() for o in out]))
    return [o.cpu().numpy() for o in out]


@pytest.mark.parametrize("kind", ["copolar", "crosspolar"])
def test_single_pair_bits_independent_of_launch_shape(ctx, oracle, kind):
    ns, nb = 128, 1024
    beam = inputs.make_polbeam(kind, nb, 3.0, 1.0, 0.05, 0.01, nside_sky=ns)
    beam8 = torch_ops.pack_beam(t(np.stack(beam.aBeams)), t(np.stack(beam.bBeams)))
    ra, dec, pa = inputs.ces_scan(6000, t0=200.0)
    rng = np.random.default_rng(41)
    ga, gb = (rng.normal(size=ra.size).astype(np.float32) for _ in range(2))
    args = (beam8, nb, t(ra), t(dec), t(pa), t(ga), t(gb), ns)
    check(lib.pisco_set_tuning(ctx, b"adjmap_deterministic", 0))
    ref = adjoint(ctx, *args)   # floating-point mode
    check(lib.pisco_set_tuning(ctx, b"adjmap_deterministic", -1))
    default = adjoint(ctx, *args)   # the default: integer sums for a single pair
    check(lib.pisco_set_tuning(ctx, b"adjmap_deterministic", 1))
    runs = []
    for blocks, chunk in ((0, 16), (97, 8), (300, 32), (0, 16)):
        check(lib.pisco_set_tuning(ctx, b"adjmap_blocks", blocks))
        check(lib.pisco_set_tuning(ctx, b"adjmap_chunk", chunk))
        runs.append(adjoint(ctx, *args))
    for r in runs[1:] + [default]:
        for a, b in zip(r, runs[0]):
            assert np.array_equal(a.view(np.int32), b.view(np.int32))
    for a, b in zip(runs[0], ref):
        assert rel_rms(a, b) < 1e-6
    oI, oQ, oU = oracle.adjoint_map(ns, nb, list(beam.aBeams) + list(beam.bBeams), ra, dec, pa,
                                    ga, gb)
    for a, b in zip(runs[0], (oI, oQ, oU)):
        assert rel_rms(a, b) < 1e-6
    # dot product with the forward
    sky = inputs.iid_sky(ns, seed=42)
    ta, tb = torch_ops.beam_convolve(*(t(x) for x in sky[:3]), t(np.stack(beam.aBeams)),
                                     t(np.stack(beam.bBeams)), t(ra), t(dec), t(pa), ns, nb)
    lhs = float(ta.double().cpu() @ torch.from_numpy(ga).double()
                + tb.double().cpu() @ torch.from_numpy(gb).double())
    rhs = sum(float(np.dot(g.astype(np.float64), x.astype(np.float64)))
              for g, x in zip(runs[0], sky[:3]))
    assert abs(lhs - rhs) < 1e-6 * abs(lhs)


@pytest.mark.parametrize("per_pair", [False, True])
def test_batched_bits_independent_of_launch_shape(per_pair):
    ns = 128
    rng = np.random.default_rng(43)
    specs = []
    for k in range(5):
        fwhm = float(rng.uniform(0.6, 1.5))
        b = inputs.make_polbeam("crosspolar" if k % 2 else "copolar", 512, 3 * fwhm, fwhm, 0.05,
                                0.01, nside_sky=ns)
        specs.append((t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), 512,
                      float(rng.uniform(-1, 1))))
    batch = DetectorBatch(specs)
    ra, dec, pa = inputs.ces_scan(3000, t0=700.0)
    if per_pair:
        ra, dec, pa = (np.stack([x + np.float32(0.01 * k) for k in range(5)]) for x in
                       (ra, dec, pa))
    g = torch.randn((5, 2, 3000), device="cuda")
    batch.set_tuning("adjmap_deterministic", 0)
    ref = [x.cpu().numpy() for x in batch.adjoint_map(g, t(ra), t(dec), t(pa), ns)]
    batch.set_tuning("adjmap_deterministic", 1)
    runs = []
    for blocks, chunk in ((0, 16), (131, 8), (0, 16)):
        batch.set_tuning("adjmap_blocks", blocks)
        batch.set_tuning("adjmap_chunk", chunk)
        runs.append([x.cpu().numpy() for x in batch.adjoint_map(g, t(ra), t(dec), t(pa), ns)])
    for r in runs[1:]:
        for a, b in zip(r, runs[0]):
            assert np.array_equal(a.view(np.int32), b.view(np.int32))
    for a, b in zip(runs[0], ref):
        assert rel_rms(a, b) < 1e-6
    batch.close()


def test_refused_with_interpolation(ctx):
    check(lib.pisco_set_tuning(ctx, b"adjmap_deterministic", 1))
    check(lib.pisco_set_tuning(ctx, b"beam_interpolation", 1))
    try:
        ns, nb = 32, 256
        beam = inputs.gaussian_polbeam(nb, 4.0, 1.5, nside_sky=ns)
        beam8 = torch_ops.pack_beam(t(np.stack(beam.aBeams)), t(np.stack(beam.bBeams)))
        ra, dec, pa = (t(x) for x in inputs.raster_scan(10))
        g = torch.zeros(10, device="cuda")
        with pytest.raises(PiscoError):
            adjoint(ctx, beam8, nb, ra, dec, pa, g, g, ns)
    finally:
        check(lib.pisco_set_tuning(ctx, b"beam_interpolation", 0))


def test_default_with_interpolation_uses_float_sums(ctx, oracle):
    """adjmap_deterministic -1 (the default) with beam interpolation: the
    float sums, against the oracle's interpolation mode."""
    ns, nb = 32, 256
    beam = inputs.gaussian_polbeam(nb, 4.0, 1.5, nside_sky=ns)
    beam8 = torch_ops.pack_beam(t(np.stack(beam.aBeams)), t(np.stack(beam.bBeams)))
    ra, dec, pa = inputs.raster_scan(400)
    rng = np.random.default_rng(43)
    ga, gb = (rng.normal(size=ra.size).astype(np.float32) for _ in range(2))
    check(lib.pisco_set_tuning(ctx, b"beam_interpolation", 1))
    try:
        got = adjoint(ctx, beam8, nb, t(ra), t(dec), t(pa), t(ga), t(gb), ns)
    finally:
        check(lib.pisco_set_tuning(ctx, b"beam_interpolation", 0))
    oracle.set_beam_interpolation(True)
    try:
        ref = oracle.adjoint_map(ns, nb, list(beam.aBeams) + list(beam.bBeams), ra, dec, pa, ga,
                                 gb)
    finally:
        oracle.set_beam_interpolation(False)
    for a, b in zip(got, ref):
        assert rel_rms(a, b) < 1e-6
