"""Device helpers for the GPU tests: torch allocates device memory, every
computation goes through libpisco's C-ABI (pisco4cmb_amd._lib)."""
import ctypes

import numpy as np
import torch

from pisco4cmb_amd import _lib
from pisco4cmb_amd._lib import check, lib, ptr_array


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


class Gpu:
    def __init__(self):
        self.ctx = ctypes.c_void_p()
        cfg = _lib.RunConfig(nStreams=1, maxMemUsage=1 << 40, deviceId=0, ptgPerConv=1 << 20,
                             pixelsPerDisc=0)
        check(lib.pisco_create(ctypes.byref(cfg), ctypes.byref(self.ctx)))
        self.stream = None   # the context's own stream; host() synchronises everything

    def close(self):
        if self.ctx:
            torch.cuda.synchronize()
            lib.pisco_destroy(self.ctx)
            self.ctx = None

    def sync(self):
        torch.cuda.synchronize()

    def libm(self, op, x, y=None):
        xd = dev(np.asarray(x, np.float64))
        yd = dev(np.asarray(y, np.float64)) if y is not None else None
        o = torch.empty_like(xd)
        check(lib.pisco_dev_libm(self.ctx, None, op, xd.numel(), xd.data_ptr(),
                                 yd.data_ptr() if yd is not None else None, o.data_ptr()))
        lib_sync(self)
        return host(o)

    def pix2ang(self, nside, pix):
        p = dev(np.asarray(pix, np.int32))
        th = torch.empty(p.numel(), dtype=torch.float64, device="cuda")
        ph = torch.empty_like(th)
        check(lib.pisco_dev_pix2ang(self.ctx, None, nside, p.numel(), p.data_ptr(),
                                    th.data_ptr(), ph.data_ptr()))
        lib_sync(self)
        return host(th), host(ph)

    def ang2pix(self, nside, theta, phi):
        th = dev(np.asarray(theta, np.float64))
        ph = dev(np.asarray(phi, np.float64))
        out = torch.empty(th.numel(), dtype=torch.int32, device="cuda")
        check(lib.pisco_dev_ang2pix(self.ctx, None, nside, th.numel(), th.data_ptr(),
                                    ph.data_ptr(), out.data_ptr()))
        lib_sync(self)
        return host(out)

    def get_interpol(self, nside, theta, phi):
        th = dev(np.asarray(theta, np.float64))
        ph = dev(np.asarray(phi, np.float64))
        n = th.numel()
        pix = torch.empty((n, 4), dtype=torch.int32, device="cuda")
        wgt = torch.empty((n, 4), dtype=torch.float64, device="cuda")
        check(lib.pisco_dev_get_interpol(self.ctx, None, nside, n, th.data_ptr(), ph.data_ptr(),
                                         pix.data_ptr(), wgt.data_ptr()))
        lib_sync(self)
        return host(pix), host(wgt)

    def query_disc(self, nside, ra, dec, radius):
        r = dev(np.asarray(ra, np.float32))
        d = dev(np.asarray(dec, np.float32))
        n = r.numel()
        counts = torch.empty(n, dtype=torch.int32, device="cuda")
        check(lib.pisco_dev_query_disc(self.ctx, None, nside, n, r.data_ptr(), d.data_ptr(),
                                       float(radius), None, counts.data_ptr(), None))
        lib_sync(self)
        c = host(counts).astype(np.int64)
        off = np.zeros(n + 1, np.int64)
        off[1:] = np.cumsum(c)
        offd = dev(off)
        pix = torch.empty(max(int(off[-1]), 1), dtype=torch.int32, device="cuda")
        check(lib.pisco_dev_query_disc(self.ctx, None, nside, n, r.data_ptr(), d.data_ptr(),
                                       float(radius), offd.data_ptr(), None, pix.data_ptr()))
        lib_sync(self)
        return off, host(pix)[:off[-1]]

    def match_fast(self, nside_sky, nside_beam, ra, dec, pa, off, pix, mode=1):
        """The kernels' mapping (mode 1: single precision, exact FP64 where it
        cannot certify; mode 2: single precision alone, bpix -2 where it
        declines): (bpix, cos 2chi, sin 2chi)."""
        r, d, p = (dev(np.asarray(x, np.float32)) for x in (ra, dec, pa))
        offd = dev(np.asarray(off, np.int64))
        pixd = dev(np.asarray(pix, np.int32))
        bpix = torch.empty(max(pix.size, 1), dtype=torch.int32, device="cuda")
        cs = torch.empty((max(pix.size, 1), 2), dtype=torch.float32, device="cuda")
        check(lib.pisco_dev_match_mode(self.ctx, None, mode, nside_sky, nside_beam, r.numel(),
                                       r.data_ptr(), d.data_ptr(), p.data_ptr(),
                                       offd.data_ptr(), pixd.data_ptr(), bpix.data_ptr(),
                                       cs.data_ptr()))
        lib_sync(self)
        cs = host(cs)[:pix.size]
        return host(bpix)[:pix.size], cs[:, 0], cs[:, 1]

    def match(self, nside_sky, nside_beam, ra, dec, pa, off, pix):
        r, d, p = (dev(np.asarray(x, np.float32)) for x in (ra, dec, pa))
        offd = dev(np.asarray(off, np.int64))
        pixd = dev(np.asarray(pix, np.int32))
        bpix = torch.empty(max(pix.size, 1), dtype=torch.int32, device="cuda")
        chi = torch.empty(max(pix.size, 1), dtype=torch.float64, device="cuda")
        check(lib.pisco_dev_match(self.ctx, None, nside_sky, nside_beam, r.numel(), r.data_ptr(),
                                  d.data_ptr(), p.data_ptr(), offd.data_ptr(), pixd.data_ptr(),
                                  bpix.data_ptr(), chi.data_ptr()))
        lib_sync(self)
        return host(bpix)[:pix.size], host(chi)[:pix.size]


def lib_sync(g):
    """The dev_* calls above run on the context's own (non-blocking) stream;
    torch.cuda.synchronize() waits for every stream of the device, which
    orders them before torch reads the results."""
    torch.cuda.synchronize()
