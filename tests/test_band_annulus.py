"""The banded beam adjoint's annulus decomposition (pisco_band_annulus: the
sky pixels the band kernels walk for one band around one pointing), checked
on the CPU against the oracle's own query_disc and sky-to-beam matching
(oracle/pisco_oracle.c, which the GPU mapping equals bit for bit): every disc
visit whose beam pixel lies in a band's rows is in that band's annulus, for
any position angle; no annulus holds a pixel twice; the exact-span annulus
lies inside the disc; the float-span annulus (the form the kernels use)
holds the exact one and at most a pixel more per arc end.  Cases: beam
support of 10 deg at beam NSIDE 512 on a sky NSIDE 64 (24 bands of 1,000
rows), the same near the poles, 3 deg at beam NSIDE 1024 on sky NSIDE 128,
and 60 deg at beam NSIDE 32 on sky NSIDE 16, whose support reaches the beam
map's equatorial rings."""
import math

import numpy as np
import pytest

from pisco4cmb_amd import pypisco
from pisco4cmb_amd._lib import lib


@pytest.mark.parametrize("ns,nb,cap,rows,polar", [(64, 512, 10.0, 1000, False),
                                                  (64, 512, 10.0, 1000, True),
                                                  (128, 1024, 3.0, 2500, False),
                                                  (16, 32, 60.0, 300, False)])
def test_band_annuli_hold_every_visit(oracle, ns, nb, cap, rows, polar):
    npb = pypisco.beam_npix_for_radius(nb, math.radians(cap))
    rho = lib.pisco_beam_rho_max(nb, npb)
    rng = np.random.default_rng(ns + nb + int(polar))
    n = 30
    ra = rng.uniform(0, 2 * math.pi, n).astype(np.float32)
    if polar:
        dec = (np.where(np.arange(n) % 2, 1.0, -1.0) * (math.pi / 2 - rng.uniform(0, rho, n)))
    else:
        dec = np.arcsin(rng.uniform(-1, 1, n))
    dec = dec.astype(np.float32)
    pa = rng.uniform(-math.pi, math.pi, n).astype(np.float32)
    off, pix = oracle.query_disc_many(ns, ra, dec, rho)
    bp, _ = oracle.match(ns, nb, ra, dec, pa, off, pix)
    nband = (npb + rows - 1) // rows
    buf = np.empty(12 * ns * ns, np.int32)
    walked = [0, 0]
    visits = 0
    for i in range(n):
        dp, db = pix[off[i]:off[i + 1]], bp[off[i]:off[i + 1]]
        disc = set(dp.tolist())
        visits += int(((db >= 0) & (db < npb)).sum())
        for k in range(nband):
            want = set(dp[(db >= k * rows) & (db < min(npb, (k + 1) * rows))].tolist())
            got = []
            for cover in (0, 1):
                m = lib.pisco_band_annulus(ns, nb, npb, rows, k, float(ra[i]), float(dec[i]),
                                           cover, buf.size, buf.ctypes.data)
                assert m >= 0
                s = buf[:m].tolist()
                assert len(set(s)) == m, "a pixel twice in one annulus"
                assert want <= set(s), "a visit of the band outside its annulus"
                walked[cover] += m
                got.append(set(s))
            assert got[0] <= disc
            assert got[0] <= got[1]
            assert len(got[1]) - len(got[0]) <= 4 * (4 * ns - 1)
    print("in-beam visits %d, walked %d exact / %d float-span (x%.2f)"
          % (visits, walked[0], walked[1], walked[1] / max(visits, 1)))
    assert walked[1] <= 1.05 * walked[0] + 100


def test_band_annulus_bad_arguments():
    buf = np.empty(10, np.int32)
    assert lib.pisco_band_annulus(64, 512, 1000, 100, 10, 0.0, 0.0, 1, 10,
                                  buf.ctypes.data) == -1   # band past the beam
    assert lib.pisco_band_annulus(64, 512, 0, 100, 0, 0.0, 0.0, 1, 10, buf.ctypes.data) == -1
    assert lib.pisco_band_annulus(64, 512, 1000, 0, 0, 0.0, 0.0, 1, 10, buf.ctypes.data) == -1
    assert lib.pisco_band_annulus(64, 512, 20000, 1000, 3, 0.0, 0.0, 1, 10,
                                  buf.ctypes.data) == -2   # more pixels than cap
