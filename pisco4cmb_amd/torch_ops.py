"""torch.autograd binding of the convolution (device tensors, torch's stream).

    tod_a, tod_b = beam_convolve(I, Q, U, beam_a, beam_b, ra, dec, pa,
                                 nside_sky, nside_beam, V=None)

I, Q, U: float32 [12*nside_sky^2] cuda tensors; beam_a, beam_b: float32
[4, npix_beam] (PolBeam aBeams / bBeams); ra, dec, pa: float32 [nsamples].
The forward is the CPU Convolver's map (convolver.cpp:74-183) computed by the
HIP kernels.  backward() runs the hand-written adjoint kernels: the sky
gradient (TOD -> I,Q,U scatter) and, when requested, the beam gradient; the
pixel lists and beam-frame indices are recomputed in backward, never stored.
Pointing gets no gradient (the map is not differentiable in it).

torch only supplies device memory and the current stream; every computation
is a libpisco kernel.
"""
import ctypes

import torch

from . import _lib
from ._lib import check, lib, ptr_array

_CTX = {}


def new_ctx(device):
    """A library context (device workspace) of its own on `device`; the
    caller releases it with lib.pisco_destroy."""
    idx = device.index if device.index is not None else torch.cuda.current_device()
    c = ctypes.c_void_p()
    cfg = _lib.RunConfig(nStreams=1, maxMemUsage=1 << 48, deviceId=idx,
                         ptgPerConv=1 << 22, pixelsPerDisc=0)
    check(lib.pisco_create(ctypes.byref(cfg), ctypes.byref(c)))
    return c


def _ctx(device):
    """The context of (device, current stream).  A context's workspace is
    used by one stream at a time, so work that torch issues on two streams
    gets two contexts and never shares scratch memory."""
    idx = device.index if device.index is not None else torch.cuda.current_device()
    key = (idx, torch.cuda.current_stream(idx).cuda_stream)
    if key not in _CTX:
        _CTX[key] = new_ctx(torch.device("cuda", idx))
    return _CTX[key]


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _req(t, name, dtype=torch.float32):
    if not t.is_cuda or t.dtype != dtype:
        raise ValueError("%s must be a %s cuda tensor" % (name, dtype))
    return t.contiguous()


def pack_beam(beam_a, beam_b):
    """[4,npix] x 2 -> [npix, 8] rows A0 A1 A2 B0 B1 B2 A3 B3 (kernel layout)."""
    beam_a = _req(beam_a, "beam_a")
    beam_b = _req(beam_b, "beam_b")
    npix = beam_a.shape[1]
    out = torch.empty((npix, 8), dtype=torch.float32, device=beam_a.device)
    maps = ptr_array([beam_a[k].data_ptr() for k in range(4)] +
                     [beam_b[k].data_ptr() for k in range(4)])
    check(lib.pisco_dev_pack_beam(_ctx(beam_a.device), _stream(beam_a.device), npix, maps,
                                  out.data_ptr()))
    return out


def _adjoint_map(dev, nside_sky, nside_beam, beam8, ra, dec, pa, ga, gb):
    npix_sky = 12 * nside_sky * nside_sky
    out = [torch.empty(npix_sky, dtype=torch.float32, device=dev) for _ in range(3)]
    check(lib.pisco_dev_adjoint_map(
        _ctx(dev), _stream(dev), nside_sky, nside_beam, beam8.shape[0], beam8.data_ptr(),
        ra.numel(), ra.data_ptr(), dec.data_ptr(), pa.data_ptr(),
        ga.data_ptr() if ga is not None else None, gb.data_ptr() if gb is not None else None,
        *[o.data_ptr() for o in out]))
    return out


def _adjoint_beam(dev, nside_sky, I, Q, U, nside_beam, npix_beam, ra, dec, pa, ga, gb):
    g8 = torch.empty((npix_beam, 8), dtype=torch.float32, device=dev)
    check(lib.pisco_dev_adjoint_beam(
        _ctx(dev), _stream(dev), nside_sky, I.data_ptr(), Q.data_ptr(), U.data_ptr(),
        nside_beam, npix_beam, ra.numel(), ra.data_ptr(), dec.data_ptr(), pa.data_ptr(),
        ga.data_ptr() if ga is not None else None, gb.data_ptr() if gb is not None else None,
        g8.data_ptr()))
    return g8   # rows A0 A1 A2 B0 B1 B2 A3 B3


class BeamConvolution(torch.autograd.Function):
    @staticmethod
    def forward(ctx, I, Q, U, beam_a, beam_b, ra, dec, pa, nside_sky, nside_beam, V=None):
        dev = I.device
        I, Q, U = (_req(x, n) for x, n in ((I, "I"), (Q, "Q"), (U, "U")))
        ra, dec, pa = (_req(x, n) for x, n in ((ra, "ra"), (dec, "dec"), (pa, "pa")))
        npix_sky = 12 * nside_sky * nside_sky
        if I.numel() != npix_sky or Q.numel() != npix_sky or U.numel() != npix_sky:
            raise ValueError("sky maps must have 12*nside_sky^2 pixels")
        if V is not None:
            V = _req(V, "V")
            if V.numel() != npix_sky:
                raise ValueError("sky maps must have 12*nside_sky^2 pixels")
        n = ra.numel()
        if dec.numel() != n or pa.numel() != n:
            raise ValueError("ra, dec, pa must have the same length")
        beam8 = pack_beam(beam_a, beam_b)
        npix_beam = beam8.shape[0]
        tod_a = torch.empty(n, dtype=torch.float32, device=dev)
        tod_b = torch.empty(n, dtype=torch.float32, device=dev)
        if V is None:
            check(lib.pisco_dev_forward(_ctx(dev), _stream(dev), nside_sky, I.data_ptr(),
                                        Q.data_ptr(), U.data_ptr(), nside_beam, npix_beam,
                                        beam8.data_ptr(), n, ra.data_ptr(), dec.data_ptr(),
                                        pa.data_ptr(), tod_a.data_ptr(), tod_b.data_ptr()))
            ctx.save_for_backward(I, Q, U, beam8, ra, dec, pa)
        else:
            check(lib.pisco_dev_forward_v(_ctx(dev), _stream(dev), nside_sky, I.data_ptr(),
                                          Q.data_ptr(), U.data_ptr(), V.data_ptr(), nside_beam,
                                          npix_beam, beam8.data_ptr(), n, ra.data_ptr(),
                                          dec.data_ptr(), pa.data_ptr(), tod_a.data_ptr(),
                                          tod_b.data_ptr()))
            ctx.save_for_backward(I, Q, U, beam8, ra, dec, pa, V)
        ctx.has_v = V is not None
        ctx.nside_sky = nside_sky
        ctx.nside_beam = nside_beam
        return tod_a, tod_b

    @staticmethod
    def backward(ctx, g_a, g_b):
        saved = ctx.saved_tensors
        I, Q, U, beam8, ra, dec, pa = saved[:7]
        V = saved[7] if ctx.has_v else None
        dev = I.device
        npix_beam = beam8.shape[0]
        ns, nb = ctx.nside_sky, ctx.nside_beam
        ga = g_a.contiguous().float() if g_a is not None else None
        gb = g_b.contiguous().float() if g_b is not None else None
        gI = gQ = gU = gbeam_a = gbeam_b = gV = None
        if any(ctx.needs_input_grad[:3]):
            gI, gQ, gU = _adjoint_map(dev, ns, nb, beam8, ra, dec, pa, ga, gb)
        if ctx.has_v and ctx.needs_input_grad[10]:
            # the V term is the I term of a beam whose A0, B0 are A3, B3
            vbeam = torch.zeros_like(beam8)
            vbeam[:, 0] = beam8[:, 6]
            vbeam[:, 3] = beam8[:, 7]
            gV = _adjoint_map(dev, ns, nb, vbeam, ra, dec, pa, ga, gb)[0]
        if ctx.needs_input_grad[3] or ctx.needs_input_grad[4]:
            g8 = _adjoint_beam(dev, ns, I, Q, U, nb, npix_beam, ra, dec, pa, ga, gb)
            if ctx.has_v:
                # A3, B3 receive g V: the A0, B0 columns of the pass with V for I
                gv8 = _adjoint_beam(dev, ns, V, V, V, nb, npix_beam, ra, dec, pa, ga, gb)
                g8[:, 6] = gv8[:, 0]
                g8[:, 7] = gv8[:, 3]
            gbeam_a = g8[:, [0, 1, 2, 6]].t().contiguous()
            gbeam_b = g8[:, [3, 4, 5, 7]].t().contiguous()
        return gI, gQ, gU, gbeam_a, gbeam_b, None, None, None, None, None, gV


def beam_convolve(I, Q, U, beam_a, beam_b, ra, dec, pa, nside_sky, nside_beam, V=None):
    """Differentiable (tod_a, tod_b) of a detector pair.  With a sky V map,
    the TOD gains the V terms of the reference's GPU kernel (V A3, V B3;
    cudaConv2.cu:204,214,217-218) and V, A3 and B3 get gradients; without
    it, the CPU Convolver's map (no V)."""
    return BeamConvolution.apply(I, Q, U, beam_a, beam_b, ra, dec, pa, nside_sky,
                                 nside_beam, V)
