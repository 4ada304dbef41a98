"""Batched detector pairs on the device: many PolBeam pairs that share one
boresight pointing (each with its own beam, disc radius and position-angle
offset) convolved in one launch (pisco_batch_create /
pisco_dev_batched_forward, include/pisco.h).

    batch = DetectorBatch([(beam_a, beam_b, nside_beam, pa_offset), ...])
    tod = batch.forward(I, Q, U, ra, dec, pa, nside_sky)   # [npairs, 2, n]
    gI, gQ, gU = batch.adjoint_map(g, ra, dec, pa, nside_sky)
    tod = batched_convolve(I, Q, U, batch, ra, dec, pa, nside_sky)  # autograd

beam_a, beam_b: float32 [4, npix] cuda tensors (PolBeam aBeams / bBeams);
pair k's TOD is what the single-pair forward gives with pa + pa_offset (in
float).  torch supplies device memory and the current stream only.

Each DetectorBatch owns a library context (its device workspace), shared
with no other object; BatchedStepGraph freezes that workspace once captured,
so the graph's kernel arguments stay valid whatever else runs on the device.
"""
import ctypes

import torch

from . import _lib
from ._lib import check, lib
from .torch_ops import _req, _stream, new_ctx, pack_beam


class DetectorBatch:
    def __init__(self, pairs, device=None):
        if not pairs:
            raise ValueError("a batch needs at least one pair")
        self.device = device or pairs[0][0].device
        self.beam8 = []
        table = (_lib.BatchPair * len(pairs))()
        for k, (ba, bb, nside_beam, pa_off) in enumerate(pairs):
            b8 = pack_beam(ba, bb)
            self.beam8.append(b8)   # the batch references the packed rows
            table[k].nside_beam = int(nside_beam)
            table[k].npix_beam = int(b8.shape[0])
            table[k].beam8 = b8.data_ptr()
            table[k].pa_offset = float(pa_off)
        self.npairs = len(pairs)
        self.nside_beams = [int(p[2]) for p in pairs]
        self.handle = None
        self.ctx = new_ctx(self.device)
        torch.cuda.synchronize(self.device)   # packed rows complete before the table refers to them
        h = ctypes.c_void_p()
        check(lib.pisco_batch_create(self.ctx, self.npairs, table, ctypes.byref(h)))
        self.handle = h

    def _pointing(self, ra, dec, pa):
        """(ra, dec, pa, nsamples, per_pair): [n] shared by every pair, or
        [npairs, n] with a pointing per pair (caller's pair order)."""
        ra, dec, pa = (_req(x, n) for x, n in ((ra, "ra"), (dec, "dec"), (pa, "pa")))
        if ra.shape != dec.shape or ra.shape != pa.shape:
            raise ValueError("ra, dec and pa must have one shape")
        if ra.dim() == 2:
            if ra.shape[0] != self.npairs:
                raise ValueError("a per-pair pointing must be [npairs, nsamples]")
            return ra, dec, pa, ra.shape[1], True
        return ra, dec, pa, ra.numel(), False

    def forward(self, I, Q, U, ra, dec, pa, nside_sky):
        """TOD of every pair: float32 [npairs, 2, nsamples] on the device.
        ra, dec, pa: [nsamples] shared by the pairs, or [npairs, nsamples]
        (e.g. from pointing.focal_plane_to_sky)."""
        I, Q, U = (_req(x, n) for x, n in ((I, "I"), (Q, "Q"), (U, "U")))
        ra, dec, pa, n, per_pair = self._pointing(ra, dec, pa)
        if I.numel() != 12 * nside_sky * nside_sky:
            raise ValueError("sky maps must have 12*nside_sky^2 pixels")
        tod = torch.empty((self.npairs, 2, n), dtype=torch.float32, device=self.device)
        fn = lib.pisco_dev_batched_forward_pp if per_pair else lib.pisco_dev_batched_forward
        check(fn(self.ctx, _stream(self.device), self.handle, nside_sky, I.data_ptr(),
                 Q.data_ptr(), U.data_ptr(), n, ra.data_ptr(), dec.data_ptr(), pa.data_ptr(),
                 tod.data_ptr()))
        return tod

    def adjoint_map(self, g, ra, dec, pa, nside_sky):
        """Sky gradient of <g, tod> summed over pairs: (gI, gQ, gU) float32
        maps; g is [npairs, 2, nsamples] like forward()'s output."""
        g = _req(g, "g")
        ra, dec, pa, n, per_pair = self._pointing(ra, dec, pa)
        if g.numel() != self.npairs * 2 * n:
            raise ValueError("g must be [npairs, 2, nsamples]")
        npix = 12 * nside_sky * nside_sky
        out = [torch.empty(npix, dtype=torch.float32, device=self.device) for _ in range(3)]
        fn = (lib.pisco_dev_batched_adjoint_map_pp if per_pair
              else lib.pisco_dev_batched_adjoint_map)
        check(fn(self.ctx, _stream(self.device), self.handle, nside_sky, n, ra.data_ptr(),
                 dec.data_ptr(), pa.data_ptr(), g.data_ptr(), *[o.data_ptr() for o in out]))
        return tuple(out)

    def set_tuning(self, key, value):
        """pisco_set_tuning on the batch's own context (launch shapes of its
        batched calls)."""
        check(lib.pisco_set_tuning(self.ctx, key.encode(), int(value)))

    def get_tuning(self, key):
        v = ctypes.c_long()
        check(lib.pisco_get_tuning(self.ctx, key.encode(), ctypes.byref(v)))
        return v.value

    def freeze(self, frozen=True):
        """Pin the batch's device workspace (pisco_freeze_buffers): while
        frozen, a call that needs more workspace raises PiscoError
        (PISCO_ERR_STATE) instead of re-allocating memory that a captured
        graph refers to."""
        check(lib.pisco_freeze_buffers(self.ctx, 1 if frozen else 0))

    def close(self):
        if getattr(self, "handle", None):
            lib.pisco_batch_destroy(self.handle)
            self.handle = None
        if getattr(self, "ctx", None):
            torch.cuda.synchronize(self.device)
            lib.pisco_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchedConvolution(torch.autograd.Function):
    """tod = batch.forward(I, Q, U, ...), differentiable in the sky; the
    backward runs the batched LDS adjoint (pixel lists recomputed)."""

    @staticmethod
    def forward(ctx, I, Q, U, batch, ra, dec, pa, nside_sky):
        ctx.batch = batch
        ctx.nside_sky = nside_sky
        ctx.save_for_backward(ra, dec, pa)
        return batch.forward(I, Q, U, ra, dec, pa, nside_sky)

    @staticmethod
    def backward(ctx, g):
        ra, dec, pa = ctx.saved_tensors
        gI, gQ, gU = ctx.batch.adjoint_map(g.contiguous().float(), ra, dec, pa, ctx.nside_sky)
        return gI, gQ, gU, None, None, None, None, None


def batched_convolve(I, Q, U, batch, ra, dec, pa, nside_sky):
    """Differentiable TOD [npairs, 2, nsamples] of a DetectorBatch."""
    return BatchedConvolution.apply(I, Q, U, batch, ra, dec, pa, nside_sky)


class BatchedStepGraph:
    """One batched forward + sky-adjoint step (every scan chunk) captured in a
    hipGraph (torch.cuda.CUDAGraph) and replayed: tod = forward(sky) and
    (gI, gQ, gU) = adjoint(g) on fixed input tensors, whose contents may
    change between replays.  A warm-up call runs first, so no allocation or
    first-use setup happens during capture; the batch's workspace is then
    frozen for as long as the graph lives (a larger call on the same batch
    raises instead of moving memory the graph uses)."""

    def __init__(self, batch, I, Q, U, ra, dec, pa, g, nside_sky):
        self.inputs = (I, Q, U, ra, dec, pa, g)
        self.batch = batch
        batch.forward(I, Q, U, ra, dec, pa, nside_sky)
        batch.adjoint_map(g, ra, dec, pa, nside_sky)
        torch.cuda.synchronize()
        batch.freeze(True)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.tod = batch.forward(I, Q, U, ra, dec, pa, nside_sky)
            self.grad = batch.adjoint_map(g, ra, dec, pa, nside_sky)

    def replay(self):
        self.graph.replay()
        return self.tod, self.grad
