"""Small discs: the forward and the sky adjoint with several samples per wave
(tuning keys forward_lanes and adjmap_lanes: 8, 16 or 32 lanes per sample)
against the one-sample-per-wave kernels and the CPU oracle.

Workload: sky NSIDE 256, beam NSIDE 2048, cap 1.5 deg (about 135 pixels per
disc), a CES scan with a sample count that is not a multiple of any group
size, plus samples whose discs hold the north pole.  Tolerances: TOD within
relative RMS 1e-6 of the oracle; sky adjoints within 1e-6 of the oracle's
FP64 adjoint; a sample's TOD does not depend on the other samples of the
launch, so launches split at 2^20 samples give the same bits as a launch of
the tail alone.  Every test runs on a co-polar Gaussian beam and on a
cross-polar beam (all beam maps non-zero)."""
import math

import numpy as np
import pytest

from pisco4cmb_amd import inputs, pypisco

pytestmark = pytest.mark.gpu
DEG = math.pi / 180
NS, NB = 256, 2048


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


@pytest.fixture(scope="module", params=["copolar", "crosspolar"])
def case(request):
    sky = inputs.iid_sky(NS, seed=61)
    beam = inputs.make_polbeam(request.param, NB, 1.5, 0.5, 0.05, 0.01, nside_sky=NS)
    ra, dec, pa = inputs.ces_scan(2989, t0=100.0)
    # 12 samples circling the north pole at 0.4 deg: discs that hold it
    k = np.arange(12)
    ra = np.concatenate([ra, (k * 0.5).astype(np.float32)])
    dec = np.concatenate([dec, np.full(12, (90 - 0.4) * DEG, np.float32)])
    pa = np.concatenate([pa, (k * 0.3 - 1.5).astype(np.float32)])
    return sky, beam, ra.astype(np.float32), dec.astype(np.float32), pa.astype(np.float32)


def conv_for(sky, beam, ra, dec, pa):
    conv = pypisco.GPUConvolver(pypisco.default_run_config())
    conv.configure_execution()
    conv.update_sky(pypisco.Sky(NS, *sky))
    conv.update_beam(beam)
    conv.update_scan(pypisco.Scan(ra.size, ra, dec, pa))
    return conv


def forward(conv, n):
    a = np.zeros(n, np.float32)
    b = np.zeros(n, np.float32)
    conv.exec_convolution(a, b)
    return a, b


def test_auto_defaults(case):
    conv = conv_for(*case)
    assert conv.get_tuning("forward_lanes") == 0
    assert conv.get_tuning("adjmap_lanes") == 0
    with pytest.raises(ValueError):
        conv.set_tuning("forward_lanes", 4)
    with pytest.raises(ValueError):
        conv.set_tuning("adjmap_lanes", 32)


def test_forward_lanes_vs_oracle(case, oracle):
    sky, beam, ra, dec, pa = case
    conv = conv_for(*case)
    oa, ob = oracle.convolve(NS, *sky[:3], NB, list(beam.aBeams) + list(beam.bBeams), ra, dec,
                             pa, nthreads=8)
    res = {}
    for lanes in (64, 32, 16, 8):
        conv.set_tuning("forward_lanes", lanes)
        res[lanes] = forward(conv, ra.size)
        assert rel_rms(res[lanes][0], oa) < 1e-6 and rel_rms(res[lanes][1], ob) < 1e-6
        # the polar samples too
        assert rel_rms(res[lanes][0][-12:], oa[-12:]) < 1e-6
    for lanes in (32, 16, 8):
        for x, y in zip(res[lanes], res[64]):
            assert rel_rms(x, y) < 1e-7


def test_forward_split_launches_same_bits(case):
    """2^20 + 3001 samples: the launch split at 2^20 (s_begin > 0) gives the
    tail the same bits as a launch of the tail alone."""
    sky, beam, ra, dec, pa = case
    n0 = 1 << 20
    r0, d0, p0 = inputs.ces_scan(n0, t0=5000.0)
    RA, DEC, PA = (np.concatenate([x, y]) for x, y in ((r0, ra), (d0, dec), (p0, pa)))
    for lanes in (8, 16):
        full = conv_for(sky, beam, RA, DEC, PA)
        full.set_tuning("forward_lanes", lanes)
        fa, fb = forward(full, RA.size)
        tail = conv_for(sky, beam, ra, dec, pa)
        tail.set_tuning("forward_lanes", lanes)
        ta, tb = forward(tail, ra.size)
        assert np.array_equal(fa[n0:], ta) and np.array_equal(fb[n0:], tb)


def test_adjoint_lanes_vs_oracle(case, oracle):
    sky, beam, ra, dec, pa = case
    rng = np.random.default_rng(62)
    ga, gb = (rng.normal(size=ra.size).astype(np.float32) for _ in range(2))
    conv = conv_for(*case)
    ref = oracle.adjoint_map(NS, NB, list(beam.aBeams) + list(beam.bBeams), ra, dec, pa, ga, gb)
    for lanes in (64, 16, 8):
        conv.set_tuning("adjmap_lanes", lanes)
        for chunk in (16, 128):
            conv.set_tuning("adjmap_chunk", chunk)
            got = conv.exec_adjoint_map(ga, gb)
            for k in range(3):
                assert rel_rms(got[k], ref[k]) < 1e-6


def test_adjoint_lanes_dot_product(case):
    """<A x, y> = <x, A^T y> with the several-samples-per-wave kernels, to
    1e-6 of <A x, y>.  y = A x plus noise of the same size keeps <A x, y> of
    the order of |A x| |y|: for an uncorrelated y the inner product cancels
    to ~1e-3 of its terms, and the float rounding of the LDS window sums
    (which depends on the order of the atomics) would decide the test."""
    sky, beam, ra, dec, pa = case
    conv = conv_for(*case)
    conv.set_tuning("forward_lanes", 8)
    conv.set_tuning("adjmap_lanes", 8)
    a, b = forward(conv, ra.size)
    rng = np.random.default_rng(63)
    ga, gb = ((x + x.std() * rng.normal(size=ra.size)).astype(np.float32) for x in (a, b))
    gI, gQ, gU = conv.exec_adjoint_map(ga, gb)
    lhs = float(np.dot(a.astype(np.float64), ga) + np.dot(b.astype(np.float64), gb))
    rhs = sum(float(np.dot(g.astype(np.float64), x.astype(np.float64)))
              for g, x in zip((gI, gQ, gU), sky[:3]))
    assert abs(lhs - rhs) < 1e-6 * abs(lhs)


def test_deterministic_adjoint_lanes_bits(case):
    """The deterministic sky adjoint with 8 lanes per sample: the same bits
    for any chunk size and block count."""
    sky, beam, ra, dec, pa = case
    rng = np.random.default_rng(64)
    ga, gb = (rng.normal(size=ra.size).astype(np.float32) for _ in range(2))
    conv = conv_for(*case)
    conv.set_tuning("adjmap_deterministic", 1)
    conv.set_tuning("adjmap_lanes", 8)
    runs = []
    for chunk, blocks in ((16, 0), (128, 0), (64, 37)):
        conv.set_tuning("adjmap_chunk", chunk)
        conv.set_tuning("adjmap_blocks", blocks)
        runs.append(conv.exec_adjoint_map(ga, gb))
    for r in runs[1:]:
        for x, y in zip(r, runs[0]):
            assert np.array_equal(x.view(np.int32), y.view(np.int32))


def test_autotune_pair(case, oracle):
    """The single-pair autotuner picks one of its candidates for each key,
    leaves it set, and the forward under the picked shape still equals the
    oracle."""
    import torch

    from pisco4cmb_amd import torch_ops
    from pisco4cmb_amd.autotune import autotune_pair, tuning
    from pisco4cmb_amd._lib import check, lib
    sky, beam, ra, dec, pa = case
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    I, Q, U = (t(x) for x in sky[:3])
    ba, bb = t(np.stack(beam.aBeams)), t(np.stack(beam.bBeams))
    ctx = torch_ops._ctx(torch.device("cuda:0"))
    try:
        res = autotune_pair(I, Q, U, ba, bb, t(ra), t(dec), t(pa), NS, NB, sample=2048, reps=1)
        cur = tuning()
        for key, (best, times) in res.items():
            assert best in times and cur[key] == best
        ta, tb = torch_ops.beam_convolve(I, Q, U, ba, bb, t(ra), t(dec), t(pa), NS, NB)
        oa, ob = oracle.convolve(NS, *sky[:3], NB, list(beam.aBeams) + list(beam.bBeams), ra,
                                 dec, pa, nthreads=8)
        assert rel_rms(ta.cpu().numpy(), oa) < 1e-6 and rel_rms(tb.cpu().numpy(), ob) < 1e-6
    finally:
        for k, v in ((b"forward_lanes", 0), (b"adjmap_lanes", 0), (b"adjmap_chunk", -1)):
            check(lib.pisco_set_tuning(ctx, k, v))
