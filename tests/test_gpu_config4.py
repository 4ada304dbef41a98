"""The batched configuration at its own shape: 64 detectors (32 pairs), each
pair with its own seeded elliptical beam (FWHM U[0.2, 1.0] deg, ellipticity
U[0, 0.1], cap 3 x FWHM, beam NSIDE 2048), its own position-angle offset,
on sky NSIDE 512, with the default launch shapes (auto pair ranges),
forward plus sky adjoint captured in a hipGraph (BatchedStepGraph) and
replayed.  The reference's GPUConvolver holds one pair per object
(gpuconvolver.hpp:17-82), so the oracle is the CPU Convolver run pair by
pair.  Half of the pairs have cross-polar beams (all beam maps non-zero).

Checks on a 1,200-sample prefix of the CES scan:
  - the TOD of 6 pairs spanning the range of disc radii against the oracle
    (relative RMS 1e-6 per detector) and the FP64 oracle;
  - the batched sky adjoint against the sum over all 32 pairs of the
    oracle's FP64 adjoint (relative RMS 1e-6 per map);
  - replays equal the eager batched calls bit for bit."""
import math

import numpy as np
import pytest
import torch

from pisco4cmb_amd import inputs
from pisco4cmb_amd.batched import BatchedStepGraph, DetectorBatch

pytestmark = pytest.mark.gpu
NS, NB, NPAIRS, N = 512, 2048, 32, 1200


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.fixture(scope="module")
def case():
    rng = np.random.default_rng(404)
    sky = inputs.powerlaw_sky(NS, seed=405, lmax=256)
    pairs = []
    for k in range(NPAIRS):
        fwhm = float(rng.uniform(0.2, 1.0))
        ell = float(rng.uniform(0.0, 0.1))
        kind = "crosspolar" if k % 2 else "copolar"
        beam = inputs.make_polbeam(kind, NB, 3 * fwhm, fwhm, ell, 0.01, nside_sky=NS)
        pairs.append((beam, float(rng.uniform(-math.pi, math.pi))))
    ra, dec, pa = inputs.ces_scan(N, t0=2000.0)
    g = rng.normal(size=(NPAIRS, 2, N)).astype(np.float32)
    return sky, pairs, ra, dec, pa, g


def test_configuration4_graph_against_oracle(case, oracle):
    sky, pairs, ra, dec, pa, g = case
    I, Q, U = (t(x) for x in sky[:3])
    rat, dect, pat, gt = t(ra), t(dec), t(pa), t(g)
    specs = [(t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), NB, off) for b, off in pairs]
    batch = DetectorBatch(specs)
    assert batch.get_tuning("batch_pair_range") == 0 and batch.get_tuning("batch_pair_major") == -1
    step = BatchedStepGraph(batch, I, Q, U, rat, dect, pat, gt, NS)
    tod, grad = step.replay()
    torch.cuda.synchronize()
    tod = tod.cpu().numpy()
    grad = [x.cpu().numpy() for x in grad]
    # replays equal eager calls
    eager_tod = batch.forward(I, Q, U, rat, dect, pat, NS).cpu().numpy()
    eager_grad = [x.cpu().numpy() for x in batch.adjoint_map(gt, rat, dect, pat, NS)]
    step.replay()
    torch.cuda.synchronize()
    assert np.array_equal(step.tod.cpu().numpy().view(np.int32), tod.view(np.int32))
    assert np.array_equal(eager_tod.view(np.int32), tod.view(np.int32))
    for a, b in zip(eager_grad, grad):
        assert rel_rms(a, b) < 1e-6

    radii = np.array([b.get_rho_max() for b, _ in pairs])
    order = np.argsort(radii)
    picks = order[np.linspace(0, NPAIRS - 1, 6).round().astype(int)]
    for k in picks:
        b, off = pairs[k]
        pa_k = (pa + np.float32(off)).astype(np.float32)
        beams = list(b.aBeams) + list(b.bBeams)
        oa, ob = oracle.convolve(NS, *sky[:3], NB, beams, ra, dec, pa_k, nthreads=0)
        assert rel_rms(tod[k, 0], oa) < 1e-6 and rel_rms(tod[k, 1], ob) < 1e-6, k
        da, db = oracle.convolve(NS, *sky[:3], NB, beams, ra, dec, pa_k, precision64=True,
                                 nthreads=0)
        assert rel_rms(tod[k, 0], da) < 1e-6 and rel_rms(tod[k, 1], db) < 1e-6, k

    ref = [np.zeros(12 * NS * NS) for _ in range(3)]
    for k, (b, off) in enumerate(pairs):
        pa_k = (pa + np.float32(off)).astype(np.float32)
        o = oracle.adjoint_map(NS, NB, list(b.aBeams) + list(b.bBeams), ra, dec, pa_k,
                               g[k, 0], g[k, 1])
        for c in range(3):
            ref[c] += o[c]
    for c in range(3):
        assert rel_rms(grad[c], ref[c]) < 1e-6, c
    del step
    batch.close()
