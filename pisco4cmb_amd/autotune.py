"""Launch-shape autotuner: times the candidate tile shapes of a workload on
the device and keeps the fastest (pisco_set_tuning).

    best = autotune_batch(batch, I, Q, U, ra, dec, pa, nside_sky)
    best = autotune_pair(I, Q, U, beam_a, beam_b, ra, dec, pa, nside_sky, nside_beam)

Batched candidates: pair by pair or all pairs per chunk of samples
(batch_pair_major 0, 1), pairs per launch range (batch_pair_range 0 = by beam
bytes, 4, 8, 16, -1 = all), pairs per wave of the batched forward
(batch_group 1, 2, 4) and of the batched sky adjoint (batch_adj_group 1, 2,
4), and samples per LDS window of the sky adjoint (adjmap_chunk 8, 16, 32).
Keys are tuned one after the other, each with the best value of the ones
before.  Single-pair
candidates: lanes per sample of the forward (forward_lanes 64, 32, 16, 8:
one to eight samples per wave), lanes per sample of the sky adjoint
(adjmap_lanes 64, 16, 8) and samples per LDS window (adjmap_chunk 16, 32, 64,
128).  The timing runs on the first `sample` samples of the caller's own
pointing (HIP events on torch's current stream); up to the rounding of float
sums, results do not depend on the choice, only the time does.
"""
import ctypes

import torch

from ._lib import check, lib
from .torch_ops import _ctx


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps / 1e3


def _set(ctx, key, value):
    check(lib.pisco_set_tuning(ctx, key.encode(), int(value)))


def _get(ctx, key):
    v = ctypes.c_long()
    check(lib.pisco_get_tuning(ctx, key.encode(), ctypes.byref(v)))
    return v.value


def autotune_batch(batch, I, Q, U, ra, dec, pa, nside_sky, sample=65536, reps=2):
    """Picks batch_pair_major, batch_pair_range, batch_group, batch_adj_group
    and adjmap_chunk for a DetectorBatch on this pointing; returns {key:
    (best, {value: seconds})}.  The keys are set on the batch's own context."""
    ctx = batch.ctx
    n = min(sample, ra.numel())
    r, d, p = ra[:n], dec[:n], pa[:n]
    g = torch.randn((batch.npairs, 2, n), device=batch.device)
    result = {}
    step = lambda: (batch.forward(I, Q, U, r, d, p, nside_sky),
                    batch.adjoint_map(g, r, d, p, nside_sky))
    for key, values, fn in (
            ("batch_pair_major", (0, 1), step),
            ("batch_pair_range", (0, 4, 8, 16, -1), step),
            ("batch_group", (1, 2, 4), lambda: batch.forward(I, Q, U, r, d, p, nside_sky)),
            ("batch_adj_group", (1, 2, 4), lambda: batch.adjoint_map(g, r, d, p, nside_sky)),
            ("adjmap_chunk", (8, 16, 32), lambda: batch.adjoint_map(g, r, d, p, nside_sky))):
        times = {}
        for v in values:
            _set(ctx, key, v)
            times[v] = _time(fn, reps)
        best = min(times, key=times.get)
        _set(ctx, key, best)
        result[key] = (best, times)
    return result


def autotune_pair(I, Q, U, beam_a, beam_b, ra, dec, pa, nside_sky, nside_beam, sample=65536,
                  reps=2):
    """Picks forward_lanes, adjmap_lanes and adjmap_chunk for single-pair
    launches of this beam and pointing; returns {key: (best, {value:
    seconds})}.  The adjoint keys are tuned one after the other, each with
    the best value of the ones before."""
    from .torch_ops import _stream, pack_beam
    dev = I.device
    ctx = _ctx(dev)
    n = min(sample, ra.numel())
    r, d, p = ra[:n].contiguous(), dec[:n].contiguous(), pa[:n].contiguous()
    beam8 = pack_beam(beam_a, beam_b)
    npb = beam8.shape[0]
    npix = 12 * nside_sky * nside_sky
    if I.numel() != npix or Q.numel() != npix or U.numel() != npix:
        raise ValueError("I, Q, U must hold 12*nside_sky^2 pixels")
    ta = torch.empty(n, dtype=torch.float32, device=dev)
    tb = torch.empty_like(ta)
    ga = torch.randn(n, device=dev)
    gb = torch.randn(n, device=dev)
    out = [torch.empty(npix, dtype=torch.float32, device=dev) for _ in range(3)]

    def fwd():
        check(lib.pisco_dev_forward(ctx, _stream(dev), nside_sky, I.data_ptr(), Q.data_ptr(),
                                    U.data_ptr(), nside_beam, npb, beam8.data_ptr(), n,
                                    r.data_ptr(), d.data_ptr(), p.data_ptr(), ta.data_ptr(),
                                    tb.data_ptr()))

    def adj():
        check(lib.pisco_dev_adjoint_map(ctx, _stream(dev), nside_sky, nside_beam, npb,
                                        beam8.data_ptr(), n, r.data_ptr(), d.data_ptr(),
                                        p.data_ptr(), ga.data_ptr(), gb.data_ptr(),
                                        *[o.data_ptr() for o in out]))
    result = {}
    for key, values, fn in (("forward_lanes", (64, 32, 16, 8), fwd),
                            ("adjmap_lanes", (64, 16, 8), adj),
                            ("adjmap_chunk", (16, 32, 64, 128), adj)):
        times = {}
        for v in values:
            _set(ctx, key, v)
            times[v] = _time(fn, reps)
        best = min(times, key=times.get)
        _set(ctx, key, best)
        result[key] = (best, times)
    return result


def tuning(device=None, batch=None):
    """The current launch shapes of a DetectorBatch's context, or else of the
    context that single-pair calls on the device's current stream use."""
    ctx = batch.ctx if batch is not None else _ctx(
        device or torch.device("cuda", torch.cuda.current_device()))
    return {k: _get(ctx, k) for k in ("batch_group", "batch_adj_group", "adjmap_chunk",
                                      "adjmap_cap", "adjmap_rw", "adjmap_support",
                                      "forward_lanes", "adjmap_lanes", "batch_pair_major",
                                      "batch_pair_range")}
