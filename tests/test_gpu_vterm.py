"""The V-term option (the reference GPU kernel's V A3 and V B3,
cudaConv2.cu:192,204,214,217-218) against the oracle's V-term mode, on
cross-polar beams (A3, B3 non-zero) and a sky with non-zero V:

  - pisco_dev_forward_v with one sample per wave (5-degree discs) and with 8
    lanes per sample (small discs): relative RMS 1e-6 against the oracle and
    the FP64 oracle;
  - the host API with the v_term tuning key: exec_convolution for every
    polFlag, and exec_adjoint_beam's a[3]/b[3] gradients (g V at each beam
    pixel: the oracle's A0/B0 gradient columns for the sky I := V); with the
    key off, V is ignored bit for bit; refused with beam interpolation;
  - the torch binding with V: gradients of V, A3 and B3 against the oracle,
    the dot product over every input, and finite differences (eps = 1e-3).
"""
import math

import numpy as np
import pytest
import torch

from pisco4cmb_amd import inputs, pypisco, torch_ops
from pisco4cmb_amd._lib import check, lib

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda")


def sky_with_v(ns, seed):
    I, Q, U, _ = inputs.iid_sky(ns, seed)
    V = np.random.default_rng(seed + 100).normal(0.0, 3.0, I.size).astype(np.float32)
    return I, Q, U, V


def oracle_v(oracle, V, *args, **kw):
    oracle.set_v_term(V)
    try:
        return oracle.convolve(*args, **kw)
    finally:
        oracle.set_v_term(None)


@pytest.mark.parametrize("ns,nb,cap,fwhm,n", [(256, 2048, 5.0, 0.5, 600),
                                             (256, 2048, 1.5, 0.5, 3000)])
def test_dev_forward_v_against_oracle(oracle, ns, nb, cap, fwhm, n):
    I, Q, U, V = sky_with_v(ns, 61)
    beam = inputs.crosspolar_polbeam(nb, cap, fwhm, 0.05, 0.01, nside_sky=ns)
    ra, dec, pa = inputs.ces_scan(n, t0=33.0)
    beams = list(beam.aBeams) + list(beam.bBeams)
    dev = torch.device("cuda:0")
    beam8 = torch_ops.pack_beam(t(np.stack(beam.aBeams)), t(np.stack(beam.bBeams)))
    ta = torch.empty(n, device=dev)
    tb = torch.empty(n, device=dev)
    held = [t(x) for x in (I, Q, U, V, ra, dec, pa)]   # alive until the kernel has run
    It, Qt, Ut, Vt, rat, dect, pat = held
    check(lib.pisco_dev_forward_v(torch_ops._ctx(dev), torch_ops._stream(dev), ns,
                                  It.data_ptr(), Qt.data_ptr(), Ut.data_ptr(), Vt.data_ptr(),
                                  nb, beam8.shape[0], beam8.data_ptr(), n, rat.data_ptr(),
                                  dect.data_ptr(), pat.data_ptr(), ta.data_ptr(),
                                  tb.data_ptr()))
    a, b = ta.cpu().numpy(), tb.cpu().numpy()
    oa, ob = oracle_v(oracle, V, ns, I, Q, U, nb, beams, ra, dec, pa)
    assert rel_rms(a, oa) < 1e-6 and rel_rms(b, ob) < 1e-6
    da, db = oracle_v(oracle, V, ns, I, Q, U, nb, beams, ra, dec, pa, precision64=True)
    assert rel_rms(a, da) < 1e-6 and rel_rms(b, db) < 1e-6
    # the V terms are a visible part of the TOD
    na, nb_ = oracle.convolve(ns, I, Q, U, nb, beams, ra, dec, pa)
    assert rel_rms(na, oa) > 1e-3 and rel_rms(nb_, ob) > 1e-3


def conv_for(ns, sky, beam, ra, dec, pa, vterm):
    conv = pypisco.GPUConvolver(pypisco.default_run_config())
    conv.configure_execution()
    conv.set_tuning("v_term", vterm)
    conv.update_sky(pypisco.Sky(ns, *sky))
    conv.update_beam(beam)
    conv.update_scan(pypisco.Scan(ra.size, ra, dec, pa))
    return conv


def test_host_api_v_term(oracle):
    ns, nb, n = 128, 1024, 2000
    sky = sky_with_v(ns, 62)
    beam = inputs.crosspolar_polbeam(nb, 2.5, 1.0, 0.05, 0.01, nside_sky=ns)
    ra, dec, pa = inputs.raster_scan(n)
    beams = list(beam.aBeams) + list(beam.bBeams)
    oa, ob = oracle_v(oracle, sky[3], ns, *sky[:3], nb, beams, ra, dec, pa)
    conv = conv_for(ns, sky, beam, ra, dec, pa, 1)
    assert conv.get_tuning("v_term") == 1
    for flag in ("p", "a", "b"):
        a = np.zeros(n, np.float32)
        b = np.zeros(n, np.float32)
        conv.exec_convolution(a, b, flag)
        if flag != "b":
            assert rel_rms(a, oa) < 1e-6
        if flag != "a":
            assert rel_rms(b, ob) < 1e-6
    # beam gradients: a[3], b[3] = g V at each beam pixel
    rng = np.random.default_rng(63)
    ga, gb = (rng.normal(size=n).astype(np.float32) for _ in range(2))
    gA, gB = conv.exec_adjoint_beam(ga, gb)
    G = oracle.adjoint_beam(ns, *sky[:3], nb, beams, ra, dec, pa, ga, gb)
    GV = oracle.adjoint_beam(ns, sky[3], sky[3], sky[3], nb, beams, ra, dec, pa, ga, gb)
    for k in range(3):
        assert rel_rms(gA[k], G[:, k]) < 1e-6 and rel_rms(gB[k], G[:, 3 + k]) < 1e-6
    assert rel_rms(gA[3], GV[:, 0]) < 1e-6 and rel_rms(gB[3], GV[:, 3]) < 1e-6
    conv.close()
    # key off: V is not read
    off = conv_for(ns, sky, beam, ra, dec, pa, 0)
    nov = conv_for(ns, sky[:3] + (np.zeros_like(sky[3]),), beam, ra, dec, pa, 0)
    for c in (off, nov):
        a = np.zeros(n, np.float32)
        b = np.zeros(n, np.float32)
        c.exec_convolution(a, b)
        c.res = (a, b)
    assert all(np.array_equal(x, y) for x, y in zip(off.res, nov.res))
    gA, gB = off.exec_adjoint_beam(ga, gb)
    assert np.all(gA[3] == 0) and np.all(gB[3] == 0)
    off.close()
    nov.close()


def test_v_term_refused_with_interpolation():
    ns, nb = 32, 256
    sky = sky_with_v(ns, 64)
    beam = inputs.crosspolar_polbeam(nb, 4.0, 1.5, nside_sky=ns)
    ra, dec, pa = inputs.raster_scan(10)
    conv = conv_for(ns, sky, beam, ra, dec, pa, 1)
    conv.set_tuning("beam_interpolation", 1)
    with pytest.raises(ValueError):
        conv.exec_convolution(np.zeros(10, np.float32), np.zeros(10, np.float32))
    conv.close()


def test_autograd_with_v(oracle):
    """beam_convolve(..., V): gradients of I, Q, U, V and of all eight beam
    maps against the oracle (relative RMS 1e-6), the dot product over every
    input (1e-6), and central finite differences (eps = 1e-3) on 16 V pixels
    and 16 A3/B3 entries around single perturbed entries."""
    ns, nb, n = 64, 512, 600
    I, Q, U, V = sky_with_v(ns, 65)
    beam = inputs.crosspolar_polbeam(nb, 6.0, 2.0, 0.05, 0.01, nside_sky=ns)
    ra, dec, pa = inputs.raster_scan(n)
    beams = list(beam.aBeams) + list(beam.bBeams)
    oa, ob = oracle_v(oracle, V, ns, I, Q, U, nb, beams, ra, dec, pa)
    rng = np.random.default_rng(66)
    ga = (oa + oa.std() * rng.normal(size=n)).astype(np.float32)
    gb = (ob + ob.std() * rng.normal(size=n)).astype(np.float32)
    It, Qt, Ut, Vt = (t(x).requires_grad_(True) for x in (I, Q, U, V))
    A = t(np.stack(beam.aBeams)).requires_grad_(True)
    B = t(np.stack(beam.bBeams)).requires_grad_(True)
    r, d, p, g1, g2 = (t(x) for x in (ra, dec, pa, ga, gb))

    def loss(I_, Q_, U_, A_, B_, V_):
        x, y = torch_ops.beam_convolve(I_, Q_, U_, A_, B_, r, d, p, ns, nb, V=V_)
        return (x.double() * g1.double()).sum() + (y.double() * g2.double()).sum()
    L = loss(It, Qt, Ut, A, B, Vt)
    L.backward()
    gI, gQ, gU, gV = (x.grad.cpu().numpy() for x in (It, Qt, Ut, Vt))
    gA, gB = A.grad.cpu().numpy(), B.grad.cpu().numpy()
    ref = oracle.adjoint_map(ns, nb, beams, ra, dec, pa, ga, gb)
    for got, want in zip((gI, gQ, gU), ref):
        assert rel_rms(got, want) < 1e-6
    vbeam = [beams[3], beams[1], beams[2], beams[3], beams[7], beams[5], beams[6], beams[7]]
    assert rel_rms(gV, oracle.adjoint_map(ns, nb, vbeam, ra, dec, pa, ga, gb)[0]) < 1e-6
    G = oracle.adjoint_beam(ns, I, Q, U, nb, beams, ra, dec, pa, ga, gb)
    GV = oracle.adjoint_beam(ns, V, V, V, nb, beams, ra, dec, pa, ga, gb)
    for k in range(3):
        assert rel_rms(gA[k], G[:, k]) < 1e-6 and rel_rms(gB[k], G[:, 3 + k]) < 1e-6
    assert rel_rms(gA[3], GV[:, 0]) < 1e-6 and rel_rms(gB[3], GV[:, 3]) < 1e-6
    lhs = float(L.item())
    rhs_sky = sum(float(np.dot(g.astype(np.float64), x)) for g, x in
                  ((gI, I), (gQ, Q), (gU, U), (gV, V)))
    rhs_beam = float(np.sum(gA.astype(np.float64) * np.stack(beam.aBeams)) +
                     np.sum(gB.astype(np.float64) * np.stack(beam.bBeams)))
    assert abs(lhs - rhs_sky) <= 1e-6 * abs(lhs)
    assert abs(lhs - rhs_beam) <= 1e-6 * abs(lhs)
    eps = 1e-3
    zero = t(np.zeros(I.size, np.float32))
    with torch.no_grad():
        nz = np.flatnonzero(np.abs(gV) > 1e-3 * np.abs(gV).max())
        for pix in rng.choice(nz, 16, replace=False):
            vals = []
            for s in (eps, -eps):
                v = zero.clone()
                v[pix] = s
                vals.append(float(loss(zero, zero, zero, A.detach(), B.detach(), v)))
            assert abs((vals[0] - vals[1]) / (2 * eps) - gV[pix]) <= 1e-4 * np.abs(gV).max()
        za, zb = torch.zeros_like(A), torch.zeros_like(B)
        for j in range(16):
            grad = gA if j % 2 == 0 else gB
            bp = int(rng.choice(np.flatnonzero(np.abs(grad[3]) > 1e-3 * np.abs(grad[3]).max())))
            vals = []
            for s in (eps, -eps):
                a, b = za.clone(), zb.clone()
                (a if j % 2 == 0 else b)[3, bp] = s
                vals.append(float(loss(It.detach(), Qt.detach(), Ut.detach(), a, b,
                                       Vt.detach())))
            assert abs((vals[0] - vals[1]) / (2 * eps) - grad[3][bp]) <= \
                1e-4 * np.abs(grad[3]).max()
