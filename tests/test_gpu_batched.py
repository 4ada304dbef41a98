"""Batched detector pairs (pisco4cmb_amd.batched) against the single-pair
forward and the CPU oracle: per pair, the batched TOD is the TOD of that pair
with its own pa (pa + pa_offset in float).  Pairs have distinct beams, disc
radii (cap 3 x FWHM), beam NSIDEs and position-angle offsets; 7 pairs
exercise the padding of the table to a multiple of the group size; pairs
alternate between co-polar and cross-polar beams (all beam maps non-zero).
Each DetectorBatch has a context of its own, so launch shapes are set on the
batch.  Tolerance: relative RMS 1e-6 per detector."""
import math

import numpy as np
import pytest
import torch

from pisco4cmb_amd import inputs, torch_ops
from pisco4cmb_amd.batched import DetectorBatch
from pisco4cmb_amd._lib import PISCO_ERR_STATE, PiscoError

pytestmark = pytest.mark.gpu
NS = 128


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


@pytest.fixture(scope="module")
def case():
    rng = np.random.default_rng(33)
    sky = inputs.iid_sky(NS, seed=34)
    pairs = []
    for k in range(7):
        fwhm = float(rng.uniform(0.6, 2.0))
        nb = 512 if k % 3 else 1024
        beam = inputs.make_polbeam("crosspolar" if k % 2 == 0 else "copolar", nb, 3 * fwhm,
                                   fwhm, float(rng.uniform(0, 0.1)), 0.01, nside_sky=NS)
        pairs.append((beam, nb, float(rng.uniform(-math.pi, math.pi))))
    ra, dec, pa = inputs.raster_scan(1500)
    return sky, pairs, ra, dec, pa


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.parametrize("group", [1, 2, 4])
def test_batched_forward_matches_single_pair_and_oracle(case, oracle, group):
    sky, pairs, ra, dec, pa = case
    I, Q, U = (t(x) for x in sky[:3])
    rat, dect, pat = t(ra), t(dec), t(pa)
    specs = [(t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), nb, off) for b, nb, off in pairs]
    batch = DetectorBatch(specs)
    batch.set_tuning("batch_group", group)
    tod = batch.forward(I, Q, U, rat, dect, pat, NS).cpu().numpy()
    for k, (b, nb, off) in enumerate(pairs):
        pa_k = (pa + np.float32(off)).astype(np.float32)
        ta, tb = torch_ops.beam_convolve(I, Q, U, specs[k][0], specs[k][1], rat, dect,
                                         t(pa_k), NS, nb)
        ta, tb = ta.cpu().numpy(), tb.cpu().numpy()
        assert rel_rms(tod[k, 0], ta) < 1e-6 and rel_rms(tod[k, 1], tb) < 1e-6
        if k < 3:   # the CPU oracle on a few pairs
            ra_, rb_ = oracle.convolve(NS, *sky[:3], nb, list(b.aBeams) + list(b.bBeams),
                                       ra, dec, pa_k)
            assert rel_rms(tod[k, 0], ra_) < 1e-6 and rel_rms(tod[k, 1], rb_) < 1e-6
    batch.close()


@pytest.mark.parametrize("adj_group", [1, 2, 4])
def test_batched_adjoint_sum_of_pairs_and_dot_product(case, adj_group):
    """Batched sky adjoint = sum over pairs of the single-pair adjoints
    (relative RMS 1e-6); <tod, g> = <sky, A^T g> to 1e-6 through the
    autograd binding."""
    from pisco4cmb_amd.batched import batched_convolve
    sky, pairs, ra, dec, pa = case
    I, Q, U = (t(x).requires_grad_(True) for x in sky[:3])
    rat, dect, pat = t(ra), t(dec), t(pa)
    specs = [(t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), nb, off) for b, nb, off in pairs]
    batch = DetectorBatch(specs)
    batch.set_tuning("batch_adj_group", adj_group)
    rng = np.random.default_rng(8)
    g = t(rng.normal(size=(len(pairs), 2, ra.size)).astype(np.float32))
    tod = batched_convolve(I, Q, U, batch, rat, dect, pat, NS)
    L = (tod.double() * g.double()).sum()
    L.backward()
    got = [x.grad.cpu().numpy() for x in (I, Q, U)]
    ref = [np.zeros(12 * NS * NS) for _ in range(3)]
    for k, (b, nb, off) in enumerate(pairs):
        pa_k = t((pa + np.float32(off)).astype(np.float32))
        Ik, Qk, Uk = (t(x).requires_grad_(True) for x in sky[:3])
        ta, tb = torch_ops.beam_convolve(Ik, Qk, Uk, specs[k][0], specs[k][1], rat, dect, pa_k,
                                         NS, nb)
        ((ta.double() * g[k, 0].double()).sum() + (tb.double() * g[k, 1].double()).sum()).backward()
        for c, x in enumerate((Ik, Qk, Uk)):
            ref[c] += x.grad.cpu().numpy()
    for c in range(3):
        assert rel_rms(got[c], ref[c]) < 1e-6
    lhs = float(L.item())
    rhs = sum(float(np.dot(got[c].astype(np.float64), sky[c])) for c in range(3))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
    batch.close()


def test_graph_replays_match_eager(case):
    """A captured batched forward + adjoint step replays to the eager results
    (three replays with eager work and synchronisations in between, the sky
    changed before the last one)."""
    from pisco4cmb_amd.batched import BatchedStepGraph
    sky, pairs, ra, dec, pa = case
    I, Q, U = (t(x) for x in sky[:3])
    rat, dect, pat = t(ra), t(dec), t(pa)
    specs = [(t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), nb, off) for b, nb, off in pairs]
    batch = DetectorBatch(specs)
    g = torch.randn((len(pairs), 2, ra.size), device="cuda")
    step = BatchedStepGraph(batch, I, Q, U, rat, dect, pat, g, NS)
    for rep in range(3):
        if rep == 2:
            I.mul_(0.5)
        tod, grad = step.replay()
        torch.cuda.synchronize()
        got_tod = tod.cpu().numpy()
        got_grad = [x.cpu().numpy() for x in grad]
        ref_tod = batch.forward(I, Q, U, rat, dect, pat, NS).cpu().numpy()
        ref_grad = [x.cpu().numpy() for x in batch.adjoint_map(g, rat, dect, pat, NS)]
        torch.cuda.synchronize()
        assert rel_rms(got_tod, ref_tod) < 1e-6
        for a, b in zip(got_grad, ref_grad):
            assert rel_rms(a, b) < 1e-6
    batch.close()


def test_autotuner_picks_a_valid_shape(case):
    from pisco4cmb_amd.autotune import autotune_batch, tuning
    sky, pairs, ra, dec, pa = case
    I, Q, U = (t(x) for x in sky[:3])
    specs = [(t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), nb, off) for b, nb, off in pairs]
    batch = DetectorBatch(specs)
    before = tuning()
    res = autotune_batch(batch, I, Q, U, t(ra), t(dec), t(pa), NS, sample=1024, reps=1)
    cur = tuning(batch=batch)
    for key, (best, times) in res.items():
        assert best in times and cur[key] == best
    assert tuning() == before   # single-pair calls keep their shapes
    batch.close()


@pytest.mark.parametrize("per_pair", [False, True])
def test_pair_major_matches_interleaved(case, per_pair):
    """batch_pair_major = 1 (the pairs one after another through the
    single-pair kernels) against 0 (all pairs per chunk of samples): forward
    TOD and sky adjoint within relative RMS 1e-6, with a shared pointing and
    with per-pair pointing."""
    sky, pairs, ra, dec, pa = case
    I, Q, U = (t(x) for x in sky[:3])
    if per_pair:
        ra, dec, pa = (np.stack([x + np.float32(0.003 * k) for k in range(len(pairs))])
                       for x in (ra, dec, pa))
    rat, dect, pat = t(ra), t(dec), t(pa)
    specs = [(t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), nb, off) for b, nb, off in pairs]
    batch = DetectorBatch(specs)
    g = torch.randn((len(pairs), 2, ra.shape[-1]), device="cuda")
    res = {}
    for mode in (0, 1):
        batch.set_tuning("batch_pair_major", mode)
        tod = batch.forward(I, Q, U, rat, dect, pat, NS).cpu().numpy()
        adj = [x.cpu().numpy() for x in batch.adjoint_map(g, rat, dect, pat, NS)]
        res[mode] = (tod, adj)
    for k in range(len(pairs)):
        for d in range(2):
            assert rel_rms(res[1][0][k, d], res[0][0][k, d]) < 1e-6
    for a, b in zip(res[1][1], res[0][1]):
        assert rel_rms(a, b) < 1e-6
    batch.close()


def test_pair_major_auto_for_large_beams(oracle):
    """Two pairs with the headline beam shape (cap 5 deg, beam NSIDE 4096:
    12.2 MB packed beams) run pair by pair by default; the batched TOD equals
    the single-pair forward and the oracle, the batched sky adjoint the sum
    of the single-pair adjoints (relative RMS 1e-6)."""
    ns = 512
    sky = inputs.iid_sky(ns, seed=71)
    I, Q, U = (t(x) for x in sky[:3])
    rng = np.random.default_rng(72)
    beams = [inputs.make_polbeam(kind, 4096, 5.0, float(rng.uniform(0.3, 0.36)),
                                 float(rng.uniform(0, 0.1)), 0.01, nside_sky=ns)
             for kind in ("crosspolar", "copolar")]
    offs = [0.4, -1.1]
    ra, dec, pa = inputs.ces_scan(1200, t0=900.0)
    rat, dect, pat = t(ra), t(dec), t(pa)
    specs = [(t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), 4096, o) for b, o in zip(beams, offs)]
    batch = DetectorBatch(specs)
    tod = batch.forward(I, Q, U, rat, dect, pat, ns).cpu().numpy()
    g = torch.randn((2, 2, ra.size), device="cuda")
    got = [x.cpu().numpy() for x in batch.adjoint_map(g, rat, dect, pat, ns)]
    ref = [np.zeros(12 * ns * ns) for _ in range(3)]
    for k, (b, o) in enumerate(zip(beams, offs)):
        pa_k = (pa + np.float32(o)).astype(np.float32)
        Ik, Qk, Uk = (t(x).requires_grad_(True) for x in sky[:3])
        ta, tb = torch_ops.beam_convolve(Ik, Qk, Uk, specs[k][0], specs[k][1], rat, dect,
                                         t(pa_k), ns, 4096)
        assert rel_rms(tod[k, 0], ta.detach().cpu().numpy()) < 1e-6
        assert rel_rms(tod[k, 1], tb.detach().cpu().numpy()) < 1e-6
        ((ta * g[k, 0]).sum() + (tb * g[k, 1]).sum()).backward()
        for c, x in enumerate((Ik, Qk, Uk)):
            ref[c] += x.grad.cpu().numpy()
        if k == 0:
            n = 100
            oa, ob = oracle.convolve(ns, *sky[:3], 4096, list(b.aBeams) + list(b.bBeams),
                                     ra[:n], dec[:n], pa_k[:n], nthreads=8)
            assert rel_rms(tod[k, 0, :n], oa) < 1e-6 and rel_rms(tod[k, 1, :n], ob) < 1e-6
    for a, r in zip(got, ref):
        assert rel_rms(a, r) < 1e-6
    batch.close()


def test_pair_ranges(case):
    """Batched launches over ranges of the pair table (batch_pair_range 4 and
    auto) against one range of all pairs (-1): forward and sky adjoint within
    relative RMS 1e-6; the deterministic sky adjoint gives the same bits."""
    sky, pairs, ra, dec, pa = case
    I, Q, U = (t(x) for x in sky[:3])
    rat, dect, pat = t(ra), t(dec), t(pa)
    specs = [(t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), nb, off) for b, nb, off in pairs]
    batch = DetectorBatch(specs)
    g = torch.randn((len(pairs), 2, ra.size), device="cuda")
    res, det = {}, {}
    batch.set_tuning("batch_pair_major", 0)
    for rng_ in (-1, 4, 0):
        batch.set_tuning("batch_pair_range", rng_)
        tod = batch.forward(I, Q, U, rat, dect, pat, NS).cpu().numpy()
        adj = [x.cpu().numpy() for x in batch.adjoint_map(g, rat, dect, pat, NS)]
        res[rng_] = (tod, adj)
        batch.set_tuning("adjmap_deterministic", 1)
        det[rng_] = [x.cpu().numpy() for x in batch.adjoint_map(g, rat, dect, pat, NS)]
        batch.set_tuning("adjmap_deterministic", 0)
    for rng_ in (4, 0):
        assert rel_rms(res[rng_][0], res[-1][0]) < 1e-6
        for a, b in zip(res[rng_][1], res[-1][1]):
            assert rel_rms(a, b) < 1e-6
        for a, b in zip(det[rng_], det[-1]):
            assert np.array_equal(a.view(np.int32), b.view(np.int32))
    batch.close()


def test_graph_survives_larger_calls(case):
    """A captured step keeps replaying to the same bits when, after capture,
    (1) a single-pair call on the same device needs more workspace than the
    batch's calls did (it runs on another context), and (2) the batch itself
    is called on a larger sky (NSIDE 256 against the captured 128): that
    call is refused (PISCO_ERR_STATE), as its workspace is frozen for the
    graph, and works again once unfrozen.  The sky adjoint runs in its
    deterministic mode, whose bits do not depend on the order of the LDS
    sums, so replays are compared bit for bit."""
    from pisco4cmb_amd.batched import BatchedStepGraph
    sky, pairs, ra, dec, pa = case
    n = 300
    I, Q, U = (t(x) for x in sky[:3])
    rat, dect, pat = t(ra[:n]), t(dec[:n]), t(pa[:n])
    specs = [(t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), nb, off) for b, nb, off in pairs]
    batch = DetectorBatch(specs)
    batch.set_tuning("adjmap_deterministic", 1)
    g = torch.randn((len(pairs), 2, n), device="cuda")
    step = BatchedStepGraph(batch, I, Q, U, rat, dect, pat, g, NS)
    tod, grad = step.replay()
    torch.cuda.synchronize()
    ref = [tod.cpu().numpy()] + [x.cpu().numpy() for x in grad]
    # a larger single-pair call on the same device: sky NSIDE 512, beam NSIDE 4096
    ns2 = 512
    big = inputs.gaussian_polbeam(4096, 5.0, 0.33, 0.0, 0.0, nside_sky=ns2)
    sky2 = [t(x) for x in inputs.iid_sky(ns2, seed=9)[:3]]
    r2, d2, p2 = (t(x) for x in inputs.ces_scan(4000))
    for x in sky2:
        x.requires_grad_(True)
    ta, tb = torch_ops.beam_convolve(*sky2, t(np.stack(big.aBeams)), t(np.stack(big.bBeams)),
                                     r2, d2, p2, ns2, 4096)
    (ta.sum() + tb.sum()).backward()
    # the batch on a larger sky than captured
    with pytest.raises(PiscoError) as e:
        batch.adjoint_map(g, rat, dect, pat, 2 * NS)
    assert e.value.code == PISCO_ERR_STATE
    for _ in range(3):
        tod, grad = step.replay()
        torch.cuda.synchronize()
        got = [tod.cpu().numpy()] + [x.cpu().numpy() for x in grad]
        for a, b in zip(got, ref):
            assert np.array_equal(a.view(np.int32), b.view(np.int32))
    del step
    batch.freeze(False)
    out = batch.adjoint_map(g, rat, dect, pat, 2 * NS)
    assert all(bool(torch.isfinite(x).all()) for x in out)
    assert float(out[0].abs().sum()) > 0
    batch.close()
