"""Binned map-making of device-resident TOD (pisco_skymap_*, include/pisco.h):
SkyMap::accumulate_data / solve_map (src/SkyMap/skymap.cpp:68-231) on torch
cuda tensors, on torch's current stream.

    m = DeviceSkyMap(nside)
    m.accumulate(ra, dec, pa, pol_angles, data)            # [ndet, n] float32
    m.accumulate(ra, dec, pa, pol_angles, data, shared_pointing=True)  # ra [n]
    I, Q, U, hits = m.solve()

The per-pixel sums equal the reference's sequential FP64 sums bit for bit
(samples are added detector-major, then by sample, as skymap.cpp:86-115).
"""
import ctypes

import torch

from ._lib import check, lib
from .torch_ops import _req, _stream


class DeviceSkyMap:
    def __init__(self, nside, device=None):
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        idx = self.device.index if self.device.index is not None else torch.cuda.current_device()
        self.nside = int(nside)
        self.npix = 12 * self.nside * self.nside
        h = ctypes.c_void_p()
        check(lib.pisco_skymap_create(idx, self.nside, ctypes.byref(h)))
        self.handle = h

    def accumulate(self, ra, dec, pa, pol_angles, data, shared_pointing=False):
        """data: [ndet, n]; ra, dec, pa: [ndet, n], or [n] with shared_pointing;
        pol_angles: [ndet] (radians, added to pa as skymap.cpp:95)."""
        data = _req(data, "data")
        if data.dim() != 2:
            raise ValueError("data must be [ndet, nsamples]")
        ndet, n = data.shape
        ra, dec, pa = (_req(x, nm) for x, nm in ((ra, "ra"), (dec, "dec"), (pa, "pa")))
        want = n if shared_pointing else ndet * n
        for x, nm in ((ra, "ra"), (dec, "dec"), (pa, "pa")):
            if x.numel() != want:
                raise ValueError("%s must hold %d values" % (nm, want))
        pol = _req(pol_angles, "pol_angles")
        if pol.numel() != ndet:
            raise ValueError("pol_angles must hold one angle per detector")
        check(lib.pisco_dev_skymap_accumulate(self.handle, _stream(self.device), n,
                                              ra.data_ptr(), dec.data_ptr(), pa.data_ptr(),
                                              int(bool(shared_pointing)), ndet, pol.data_ptr(),
                                              data.data_ptr()))

    def solve(self):
        """(I, Q, U float32, hits int32) cuda tensors of 12 nside^2."""
        I, Q, U = (torch.empty(self.npix, dtype=torch.float32, device=self.device)
                   for _ in range(3))
        hits = torch.empty(self.npix, dtype=torch.int32, device=self.device)
        check(lib.pisco_dev_skymap_solve(self.handle, _stream(self.device), I.data_ptr(),
                                         Q.data_ptr(), U.data_ptr(), hits.data_ptr(), None))
        return I, Q, U, hits

    def skipped(self):
        v = ctypes.c_long()
        check(lib.pisco_skymap_skipped(self.handle, ctypes.byref(v)))
        return v.value

    def reset(self):
        torch.cuda.current_stream(self.device).synchronize()
        check(lib.pisco_skymap_reset(self.handle))

    def close(self):
        if getattr(self, "handle", None) is not None:
            lib.pisco_skymap_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
