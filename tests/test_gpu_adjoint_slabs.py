"""The sky adjoint on discs larger than its LDS window: a chunk whose window
exceeds the window capacity (adjmap_cap) is split into slabs of consecutive
sky rings, each summed in LDS and flushed in turn, instead of taking global
atomics.  Cases: configuration 5's 1-degree family (sky NSIDE 2048, beam
NSIDE 2048, cap 3 deg: 34,489-pixel discs), and a 10-degree cap at NSIDE 512
(about 24,000 pixels), with cross-polar beams; window capacities from 512 to
6,144 pixels (one to dozens of slabs).  Checks: relative RMS 1e-6 against
the oracle's FP64 adjoint, single pair and batched; the deterministic mode
gives the same bits for every capacity (the slabs change only where the
integer sums are formed); the dot product with the forward."""
import math

import numpy as np
import pytest
import torch

from pisco4cmb_amd import inputs, torch_ops
from pisco4cmb_amd._lib import check, lib
from pisco4cmb_amd.batched import DetectorBatch

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda")


def adjoint(ctx, ns, nb, beam8, r, d, p, ga, gb):
    dev = torch.device("cuda:0")
    out = [torch.empty(12 * ns * ns, device=dev) for _ in range(3)]
    check(lib.pisco_dev_adjoint_map(ctx, torch_ops._stream(dev), ns, nb, beam8.shape[0],
                                    beam8.data_ptr(), r.numel(), r.data_ptr(), d.data_ptr(),
                                    p.data_ptr(), ga.data_ptr(), gb.data_ptr(),
                                    *[o.data_ptr() for o in out]))
    return [o.cpu().numpy() for o in out]


@pytest.mark.parametrize("ns,nb,cap_deg,fwhm,n", [(2048, 2048, 3.0, 1.0, 160),
                                                 (512, 1024, 10.0, 3.0, 300)])
def test_large_discs_in_slabs(oracle, ns, nb, cap_deg, fwhm, n):
    sky = inputs.iid_sky(ns, seed=ns + 3)
    beam = inputs.crosspolar_polbeam(nb, cap_deg, fwhm, 0.05, 0.01, nside_sky=ns)
    ra, dec, pa = inputs.ces_scan(n, t0=321.0)
    rng = np.random.default_rng(ns)
    ga, gb = (rng.normal(size=n).astype(np.float32) for _ in range(2))
    beams = list(beam.aBeams) + list(beam.bBeams)
    ref = oracle.adjoint_map(ns, nb, beams, ra, dec, pa, ga, gb)
    dev = torch.device("cuda:0")
    ctx = torch_ops.new_ctx(dev)
    try:
        beam8 = torch_ops.pack_beam(t(np.stack(beam.aBeams)), t(np.stack(beam.bBeams)))
        r, d, p, g1, g2 = (t(x) for x in (ra, dec, pa, ga, gb))
        check(lib.pisco_set_tuning(ctx, b"adjmap_lanes", 64))
        bits = []
        for cap in (6144, 2048, 512):
            check(lib.pisco_set_tuning(ctx, b"adjmap_cap", cap))
            check(lib.pisco_set_tuning(ctx, b"adjmap_deterministic", 0))
            got = adjoint(ctx, ns, nb, beam8, r, d, p, g1, g2)
            for x, y in zip(got, ref):
                assert rel_rms(x, y) < 1e-6, cap
            check(lib.pisco_set_tuning(ctx, b"adjmap_deterministic", 1))
            bits.append(adjoint(ctx, ns, nb, beam8, r, d, p, g1, g2))
        for b in bits[1:]:
            for x, y in zip(b, bits[0]):
                assert np.array_equal(x.view(np.int32), y.view(np.int32))
        for x, y in zip(bits[0], ref):
            assert rel_rms(x, y) < 1e-6
        # dot product with the forward (cotangents correlated with the TOD)
        check(lib.pisco_set_tuning(ctx, b"adjmap_deterministic", 0))
        check(lib.pisco_set_tuning(ctx, b"adjmap_cap", 6144))
        Is, Qs, Us = (t(x) for x in sky[:3])
        ta = torch.empty(n, device=dev)
        tb = torch.empty(n, device=dev)
        check(lib.pisco_dev_forward(ctx, torch_ops._stream(dev), ns, Is.data_ptr(),
                                    Qs.data_ptr(), Us.data_ptr(), nb, beam8.shape[0],
                                    beam8.data_ptr(), n, r.data_ptr(), d.data_ptr(),
                                    p.data_ptr(), ta.data_ptr(), tb.data_ptr()))
        a, b = ta.cpu().numpy(), tb.cpu().numpy()
        ya = (a + a.std() * rng.normal(size=n)).astype(np.float32)
        yb = (b + b.std() * rng.normal(size=n)).astype(np.float32)
        y1, y2 = t(ya), t(yb)
        gI, gQ, gU = adjoint(ctx, ns, nb, beam8, r, d, p, y1, y2)
        lhs = float(np.dot(a.astype(np.float64), ya) + np.dot(b.astype(np.float64), yb))
        rhs = sum(float(np.dot(g.astype(np.float64), x.astype(np.float64)))
                  for g, x in zip((gI, gQ, gU), sky[:3]))
        assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
    finally:
        torch.cuda.synchronize()
        lib.pisco_destroy(ctx)


def test_batched_large_discs_in_slabs(oracle):
    """Configuration 5's 1-degree family batched: two pairs (beam NSIDE 2048,
    cap 3 deg) with angle offsets; the batched sky adjoint (deterministic,
    the same bits for window capacities 6,144 and 1,024) equals the sum of the
    oracle's adjoints of the two pairs."""
    ns, nb, n = 2048, 2048, 120
    sky = inputs.iid_sky(ns, seed=77)
    beams = [inputs.crosspolar_polbeam(nb, 3.0, 1.0, 0.05, 0.01, nside_sky=ns),
             inputs.gaussian_polbeam(nb, 3.0, 1.0, 0.08, 0.01, nside_sky=ns)]
    offs = [0.3, -1.2]
    ra, dec, pa = inputs.ces_scan(n, t0=777.0)
    batch = DetectorBatch([(t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), nb, o)
                           for b, o in zip(beams, offs)])
    g = np.random.default_rng(78).normal(size=(2, 2, n)).astype(np.float32)
    gt = t(g)
    ref = [np.zeros(12 * ns * ns) for _ in range(3)]
    for k, (b, o) in enumerate(zip(beams, offs)):
        pk = (pa + np.float32(o)).astype(np.float32)
        for c, x in enumerate(oracle.adjoint_map(ns, nb, list(b.aBeams) + list(b.bBeams), ra,
                                                 dec, pk, g[k, 0], g[k, 1])):
            ref[c] += x
    batch.set_tuning("batch_pair_major", 0)
    got = [x.cpu().numpy() for x in batch.adjoint_map(gt, t(ra), t(dec), t(pa), ns)]
    for x, y in zip(got, ref):
        assert rel_rms(x, y) < 1e-6
    batch.set_tuning("adjmap_deterministic", 1)
    runs = []
    for cap in (6144, 1024):
        batch.set_tuning("adjmap_cap", cap)
        runs.append([x.cpu().numpy() for x in batch.adjoint_map(gt, t(ra), t(dec), t(pa), ns)])
    for x, y in zip(runs[1], runs[0]):
        assert np.array_equal(x.view(np.int32), y.view(np.int32))
    for x, y in zip(runs[0], ref):
        assert rel_rms(x, y) < 1e-6
    batch.close()
