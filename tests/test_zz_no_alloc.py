"""pisco_dev_forward_noalloc (the single-pair forward with its workspace in
a fixed device array) against pisco_dev_forward on the same inputs: the same
kernels read the same packed sky and ring table, so the TOD must be equal
bit for bit; both are checked against the CPU oracle (relative RMS 1e-6) on
a prefix.  Three calls in a row, each on a smaller input than the one before
(sky NSIDE 512 with the 5-degree beam, then NSIDE 256, then NSIDE 64 with
fewer samples), so a call that read a previous call's leftovers would show.
A sky above the workspace's NSIDE 512 is refused with PISCO_ERR_INVALID.
Every call goes through the C-ABI (ctypes), with torch only supplying device
memory and the stream."""
import ctypes

import numpy as np
import pytest
import torch

from pisco4cmb_amd import inputs, torch_ops
from pisco4cmb_amd._lib import PISCO_ERR_INVALID, PISCO_OK, check, lib

pytestmark = pytest.mark.gpu

CASES = [  # sky NSIDE, beam NSIDE, cap deg, FWHM deg, samples, beam kind
    (512, 4096, 5.0, 0.33, 3000, "crosspolar"),
    (256, 2048, 1.5, 0.5, 2500, "copolar"),
    (64, 512, 4.0, 1.5, 700, "crosspolar"),
]


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def test_noalloc_forward_equals_the_allocating_forward(oracle):
    dev = torch.device("cuda:0")
    stream = torch_ops._stream(dev)
    for ns, nb, cap, fwhm, n, kind in CASES:
        print("no-alloc case: sky NSIDE %d, beam NSIDE %d, cap %.1f, %d samples, %s"
              % (ns, nb, cap, n, kind), flush=True)
        sky = inputs.iid_sky(ns, seed=ns + 1)
        beam = inputs.make_polbeam(kind, nb, cap, fwhm, 0.05, 0.01, nside_sky=ns)
        ra, dec, pa = inputs.ces_scan(n, t0=float(ns))
        I, Q, U = (t(x) for x in sky[:3])
        beam8 = torch_ops.pack_beam(t(np.stack(beam.aBeams)), t(np.stack(beam.bBeams)))
        rat, dect, pat = t(ra), t(dec), t(pa)
        args = (ns, I.data_ptr(), Q.data_ptr(), U.data_ptr(), nb, beam8.shape[0],
                beam8.data_ptr(), n, rat.data_ptr(), dect.data_ptr(), pat.data_ptr())
        a1 = torch.full((n,), float("nan"), device=dev)
        b1 = torch.full((n,), float("nan"), device=dev)
        check(lib.pisco_dev_forward_noalloc(stream, *args, a1.data_ptr(), b1.data_ptr()))
        a0 = torch.empty(n, device=dev)
        b0 = torch.empty(n, device=dev)
        check(lib.pisco_dev_forward(torch_ops._ctx(dev), stream, *args, a0.data_ptr(),
                                    b0.data_ptr()))
        torch.cuda.synchronize()
        for x, y in ((a1, a0), (b1, b0)):
            assert np.array_equal(x.cpu().numpy().view(np.int32), y.cpu().numpy().view(np.int32))
        m = min(n, 300)
        oa, ob = oracle.convolve(ns, *sky[:3], nb, list(beam.aBeams) + list(beam.bBeams),
                                 ra[:m], dec[:m], pa[:m])
        assert rel_rms(a1.cpu().numpy()[:m], oa) < 1e-6
        assert rel_rms(b1.cpu().numpy()[:m], ob) < 1e-6


def test_noalloc_forward_refuses_a_larger_sky():
    dev = torch.device("cuda:0")
    z = torch.zeros(12 * 1024 * 1024, device=dev)
    beam8 = torch.zeros((16, 8), device=dev)
    r = torch.zeros(4, device=dev)
    out = torch.zeros(4, device=dev)
    rc = lib.pisco_dev_forward_noalloc(torch_ops._stream(dev), 1024, z.data_ptr(), z.data_ptr(),
                                       z.data_ptr(), 64, 16, beam8.data_ptr(), 4, r.data_ptr(),
                                       r.data_ptr(), r.data_ptr(), out.data_ptr(),
                                       out.data_ptr())
    assert rc == PISCO_ERR_INVALID
    assert b"NSIDE 512" in lib.pisco_last_error()
    # a valid call still works afterwards
    beam8[:, 0] = 1.0
    rc = lib.pisco_dev_forward_noalloc(torch_ops._stream(dev), 8, z.data_ptr(), z.data_ptr(),
                                       z.data_ptr(), 64, 16, beam8.data_ptr(), 4, r.data_ptr(),
                                       r.data_ptr(), r.data_ptr(), out.data_ptr(),
                                       out.data_ptr())
    assert rc == PISCO_OK
    torch.cuda.synchronize()
    assert bool(torch.isfinite(out).all())
