"""Gradients of the torch.autograd binding (pisco4cmb_amd.torch_ops) on the
GPU: against the CPU oracle's FP64 adjoints (the CPU-double reference
gradient), by the dot-product identity, and by central finite differences
(eps = 1e-3) on 64 random sky pixels and 64 random beam entries.

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

NS, NB = 64, 512


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


@pytest.fixture(scope="module")
def case(oracle):
    """Cotangents g = TOD + noise of the same size (TOD from the oracle):
    <TOD, g> is then of the order of |TOD| |g|, so the dot-product checks
    measure the identity, not the cancellation of an uncorrelated sum."""
    sky = inputs.iid_sky(NS, seed=21)
    beam = inputs.gaussian_polbeam(NB, 6.0, 2.0, 0.05, 0.01, nside_sky=NS)
    ra, dec, pa = inputs.raster_scan(600)
    ta, tb = oracle.convolve(NS, *sky[:3], NB, list(beam.aBeams) + list(beam.bBeams), ra, dec,
                             pa)
    rng = np.random.default_rng(5)
    ga = (ta + ta.std() * rng.normal(size=ra.size)).astype(np.float32)
    gb = (tb + tb.std() * rng.normal(size=ra.size)).astype(np.float32)
    return sky, beam, ra, dec, pa, ga, gb


def cuda(a, grad=False):
    t = torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda")
    return t.requires_grad_(grad)


def loss(I, Q, U, A, B, ra, dec, pa, ga, gb):
    ta, tb = beam_convolve(I, Q, U, A, B, ra, dec, pa, NS, NB)
    return (ta.double() * ga.double()).sum() + (tb.double() * gb.double()).sum()


def test_gradients_vs_cpu_double_dot_product_and_fd(case, oracle):
    sky, beam, ra, dec, pa, ga, gb = case
    I, Q, U = (cuda(x, True) for x in sky[:3])
    A = cuda(np.stack(beam.aBeams), True)
    B = cuda(np.stack(beam.bBeams), True)
    r, d, p, g1, g2 = (cuda(x) for x in (ra, dec, pa, ga, gb))
    L = loss(I, Q, U, A, B, r, d, p, g1, g2)
    L.backward()
    gsky = [t.grad.cpu().numpy() for t in (I, Q, U)]
    gA, gB = A.grad.cpu().numpy(), B.grad.cpu().numpy()

    # CPU-double reference gradients (the oracle's FP64 adjoints)
    beams = list(beam.aBeams) + list(beam.bBeams)
    ref = oracle.adjoint_map(NS, NB, beams, ra, dec, pa, ga, gb)
    for k in range(3):
        assert rel_rms(gsky[k], ref[k]) < 1e-6
    G = oracle.adjoint_beam(NS, *sky[:3], NB, beams, ra, dec, pa, ga, gb)
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
    rng = np.random.default_rng(9)
    npix = 12 * NS * NS
    zero = cuda(np.zeros(npix, np.float32))
    with torch.no_grad():
        for comp in range(3):
            nz = np.flatnonzero(np.abs(gsky[comp]) > 1e-3 * np.abs(gsky[comp]).max())
            for pix in rng.choice(nz, 64 // 3 + (comp == 0), replace=False):
                vals = []
                for s in (eps, -eps):
                    m = [zero.clone() for _ in range(3)]
                    m[comp][pix] = s
                    vals.append(float(loss(*m, A.detach(), B.detach(), r, d, p, g1, g2)))
                fd = (vals[0] - vals[1]) / (2 * eps)
                assert abs(fd - gsky[comp][pix]) <= 1e-4 * abs(gsky[comp]).max()
        za = torch.zeros_like(A)
        zb = torch.zeros_like(B)
        nzA = np.flatnonzero(np.abs(gA[0]) > 1e-3 * np.abs(gA[0]).max())
        for bp in rng.choice(nzA, 64, replace=False):
            k = int(rng.integers(0, 3))
            vals = []
            for s in (eps, -eps):
                a = za.clone()
                a[k, bp] = s
                vals.append(float(loss(I.detach(), Q.detach(), U.detach(), a, zb, r, d, p,
                                       g1, g2)))
            fd = (vals[0] - vals[1]) / (2 * eps)
            assert abs(fd - gA[k][bp]) <= 1e-4 * abs(gA[k]).max()
