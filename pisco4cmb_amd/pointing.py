"""Detector pointing on the device: BPoint::focal_plane_to_sky
(src/Bpoint/bpoint.cpp:553-674) over pisco_dev_focal_plane_to_sky.

    ra, dec, pa = focal_plane_to_sky(det_dx, det_dy, det_pol, ra_ac, dec_ac, pa_ac)

det_dx, det_dy, det_pol: [ndets] float64 cuda tensors (radians; dx to the
right and dy up in the receiver frame seen looking out, polarisation angle
ccw from the receiver y axis); ra_ac, dec_ac, pa_ac: [nsamples] float64 (the
array centre and its position angle ccw from the pole).  Returns float32
[ndets, nsamples] tensors ready to be a Scan or a DetectorBatch pointing, or
the float64 values with float64=True.  flag selects az/el (0), ra/dec (1) or
source-relative (2) coordinates as the reference does.
"""
import torch

from ._lib import check, lib
from .torch_ops import _ctx, _stream


def _req64(t, name):
    if not t.is_cuda or t.dtype != torch.float64:
        raise ValueError("%s must be a float64 cuda tensor" % name)
    return t.contiguous()


def focal_plane_to_sky(det_dx, det_dy, det_pol, x_ac, y_ac, rot_ac, flag=1, float64=False):
    det_dx, det_dy, det_pol = (_req64(t, n) for t, n in ((det_dx, "det_dx"), (det_dy, "det_dy"),
                                                          (det_pol, "det_pol")))
    x_ac, y_ac, rot_ac = (_req64(t, n) for t, n in ((x_ac, "x_ac"), (y_ac, "y_ac"),
                                                     (rot_ac, "rot_ac")))
    nd, n = det_dx.numel(), x_ac.numel()
    if det_dy.numel() != nd or det_pol.numel() != nd:
        raise ValueError("detector arrays must have one length")
    if y_ac.numel() != n or rot_ac.numel() != n:
        raise ValueError("array-centre arrays must have one length")
    dev = x_ac.device
    dt = torch.float64 if float64 else torch.float32
    out = [torch.empty((nd, n), dtype=dt, device=dev) for _ in range(3)]
    ptr = [o.data_ptr() for o in out]
    d64 = ptr if float64 else [None] * 3
    f32 = [None] * 3 if float64 else ptr
    check(lib.pisco_dev_focal_plane_to_sky(_ctx(dev), _stream(dev), nd, det_dx.data_ptr(),
                                           det_dy.data_ptr(), det_pol.data_ptr(), n,
                                           x_ac.data_ptr(), y_ac.data_ptr(), rot_ac.data_ptr(),
                                           int(flag), *d64, *f32))
    return tuple(out)
