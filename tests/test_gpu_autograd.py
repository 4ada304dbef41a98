"""Gradients of the torch.autograd binding (pisco4cmb_amd.torch_ops) on the
GPU: against the CPU oracle's FP64 adjoints (the CPU-double reference
gradient), by the dot-product identity, and by central finite differences
(eps = 1e-3) on 64 random sky pixels and 64 random beam entries of both
detectors; on a small shape and on the forward+adjoint configuration's
shape, each with co-polar and cross-polar beams.

The TOD is linear in the sky and in the beams, so a finite difference taken
around a map that is zero except at the perturbed entry is exact up to the
float rounding of that small TOD; the check is then sharp (relative 1e-4)
instead of being swamped by the rounding of a full-sky TOD.
"""
import math

import numpy as np
import pytest
import torch

from pisco4cmb_amd import inputs
from pisco4cmb_amd.torch_ops import beam_convolve

pytestmark = pytest.mark.gpu

def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def make_case(oracle, kind, ns, nb, cap, fwhm, ra, dec, pa, seed):
    """Cotangents g = TOD + noise of the same size (TOD from the oracle):
    <TOD, g> is then of the order of |TOD| |g|, so the dot-product checks
    measure the identity, not the cancellation of an uncorrelated sum."""
    sky = inputs.iid_sky(ns, seed=seed)
    beam = inputs.make_polbeam(kind, nb, cap, fwhm, 0.05, 0.01, nside_sky=ns)
    ta, tb = oracle.convolve(ns, *sky[:3], nb, list(beam.aBeams) + list(beam.bBeams), ra, dec,
                             pa, nthreads=8)
    rng = np.random.default_rng(seed + 1)
    ga = (ta + ta.std() * rng.normal(size=ra.size)).astype(np.float32)
    gb = (tb + tb.std() * rng.normal(size=ra.size)).astype(np.float32)
    return ns, nb, sky, beam, ra, dec, pa, ga, gb


def cuda(a, grad=False):
    t = torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda")
    return t.requires_grad_(grad)


def loss(ns, nb, I, Q, U, A, B, ra, dec, pa, ga, gb):
    ta, tb = beam_convolve(I, Q, U, A, B, ra, dec, pa, ns, nb)
    return (ta.double() * ga.double()).sum() + (tb.double() * gb.double()).sum()


def check_gradients(oracle, case, n_fd=64, fd_seed=9):
    """CPU-double gradients, both dot products, and central finite
    differences with eps = 1e-3 on n_fd sky pixels (over I, Q, U) and n_fd
    beam entries drawn from both detectors' maps 0-2."""
    ns, nb, sky, beam, ra, dec, pa, ga, gb = case
    I, Q, U = (cuda(x, True) for x in sky[:3])
    A = cuda(np.stack(beam.aBeams), True)
    B = cuda(np.stack(beam.bBeams), True)
    r, d, p, g1, g2 = (cuda(x) for x in (ra, dec, pa, ga, gb))
    L = loss(ns, nb, I, Q, U, A, B, r, d, p, g1, g2)
    L.backward()
    gsky = [t.grad.cpu().numpy() for t in (I, Q, U)]
    gA, gB = A.grad.cpu().numpy(), B.grad.cpu().numpy()

    # CPU-double reference gradients (the oracle's FP64 adjoints)
    beams = list(beam.aBeams) + list(beam.bBeams)
    ref = oracle.adjoint_map(ns, nb, beams, ra, dec, pa, ga, gb)
    for k in range(3):
        assert rel_rms(gsky[k], ref[k]) < 1e-6
    G = oracle.adjoint_beam(ns, *sky[:3], nb, beams, ra, dec, pa, ga, gb)
    for k, (got, col) in enumerate(((gA[0], 0), (gA[1], 1), (gA[2], 2),
                                     (gB[0], 3), (gB[1], 4), (gB[2], 5))):
        assert rel_rms(got, G[:, col]) < 1e-6
    assert np.all(gA[3] == 0) and np.all(gB[3] == 0)   # V is not read

    # dot-product identity: <tod, g> = <sky, grad_sky> = <beams, grad_beams>
    lhs = float(L.item())
    rhs_sky = sum(float(np.dot(gsky[k].astype(np.float64), sky[k])) for k in range(3))
    rhs_beam = float(np.sum(gA.astype(np.float64) * np.stack(beam.aBeams)) +
                     np.sum(gB.astype(np.float64) * np.stack(beam.bBeams)))
    assert abs(lhs - rhs_sky) <= 1e-6 * abs(lhs)
    assert abs(lhs - rhs_beam) <= 1e-6 * abs(lhs)

    # central finite differences, eps = 1e-3, around a single perturbed entry
    eps = 1e-3
    rng = np.random.default_rng(fd_seed)
    npix = 12 * ns * ns
    zero = cuda(np.zeros(npix, np.float32))
    with torch.no_grad():
        for comp in range(3):
            nz = np.flatnonzero(np.abs(gsky[comp]) > 1e-3 * np.abs(gsky[comp]).max())
            for pix in rng.choice(nz, n_fd // 3 + (comp == 0), replace=False):
                vals = []
                for s in (eps, -eps):
                    m = [zero.clone() for _ in range(3)]
                    m[comp][pix] = s
                    vals.append(float(loss(ns, nb, *m, A.detach(), B.detach(), r, d, p, g1,
                                           g2)))
                fd = (vals[0] - vals[1]) / (2 * eps)
                assert abs(fd - gsky[comp][pix]) <= 1e-4 * abs(gsky[comp]).max()
        za = torch.zeros_like(A)
        zb = torch.zeros_like(B)
        for j in range(n_fd):
            det = j % 2                       # detector a, then b
            grad = gA if det == 0 else gB
            k = int(rng.integers(0, 3))
            nzk = np.flatnonzero(np.abs(grad[k]) > 1e-3 * np.abs(grad[k]).max())
            bp = int(rng.choice(nzk))
            vals = []
            for s in (eps, -eps):
                a, b = za.clone(), zb.clone()
                (a if det == 0 else b)[k, bp] = s
                vals.append(float(loss(ns, nb, I.detach(), Q.detach(), U.detach(), a, b, r, d,
                                       p, g1, g2)))
            fd = (vals[0] - vals[1]) / (2 * eps)
            assert abs(fd - grad[k][bp]) <= 1e-4 * abs(grad[k]).max(), (det, k, bp)


@pytest.mark.parametrize("kind", ["copolar", "crosspolar"])
def test_gradients_vs_cpu_double_dot_product_and_fd(oracle, kind):
    """Sky NSIDE 64, beam NSIDE 512, 6 deg cap, 600 raster samples."""
    ra, dec, pa = inputs.raster_scan(600)
    check_gradients(oracle, make_case(oracle, kind, 64, 512, 6.0, 2.0, ra, dec, pa, 21))


@pytest.mark.parametrize("kind", ["copolar", "crosspolar"])
def test_gradients_at_the_headline_shape(oracle, kind):
    """The forward+adjoint configuration's shape: sky NSIDE 512, beam NSIDE
    4096, 5 deg cap, FWHM 0.33 deg, on 400 samples of the whole-sky CES scan
    (about 2.4e6 visits, so the CPU-double gradients take seconds); finite
    differences with eps = 1e-3 on 64 sky pixels and 64 beam entries of both
    detectors."""
    ra, dec, pa = inputs.ces_scan(400, fs=2.0, t0=120.0)
    check_gradients(oracle, make_case(oracle, kind, 512, 4096, 5.0, 0.33, ra, dec, pa, 23),
                    fd_seed=11)
