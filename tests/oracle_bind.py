"""ctypes binding of the CPU oracle (oracle/liboracle*.so) for the tests.

The oracle is the checker: it restates the reference CPU path in C
(oracle/pisco_oracle.c).  Only tests, the smoke check and the CPU-baseline
leg of bench.py use it.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
c_long, c_int, c_double, c_char = ctypes.c_long, ctypes.c_int, ctypes.c_double, ctypes.c_char
P = ctypes.c_void_p


def _a(x):
    return None if x is None else x.ctypes.data


def _parr(arrs):
    out = (ctypes.c_void_p * len(arrs))()
    for i, a in enumerate(arrs):
        out[i] = _a(a)
    return out


class Oracle:
    def __init__(self, glibc=False):
        name = "liboracle_glibc.so" if glibc else "liboracle.so"
        path = os.path.join(ORACLE_DIR, name)
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle (%s)" % path)
        L = ctypes.CDLL(path)
        self.L = L
        sig = {
            "or_pix2ang": (None, [c_int, c_long, P, P, P]),
            "or_ang2pix": (None, [c_int, c_long, P, P, P]),
            "or_ring_above": (None, [c_int, c_long, P, P]),
            "or_get_interpol": (None, [c_int, c_long, P, P, P, P]),
            "or_query_disc_ranges": (c_long, [c_int, c_double, c_double, c_double, P, c_long]),
            "or_query_disc_many": (c_long, [c_int, c_long, P, P, c_double, P, P]),
            "or_rho_sigma_chi": (None, [c_long, P, P, P, P, P, P, P, P]),
            "or_match": (None, [c_int, c_int, c_long, P, P, P, P, P, P, P]),
            "or_beam_rho_max": (c_double, [c_int, c_long]),
            "or_beam_from_fields": (None, [c_long, P, P, P, P, P, P, P, P]),
            "or_gaussian_fields": (None, [c_int, c_long, c_double, c_double, c_double, P, P, P, P]),
            "or_polbeam_build": (None, [c_long, c_int, c_int, c_double, c_char, P, P]),
            "or_convolve": (c_int, [c_int, P, P, P, c_int, c_long, P, c_long, P, P, P, c_char,
                                    c_int, c_int, P, P, P, P]),
            "or_adjoint_map": (None, [c_int, c_int, c_long, P, c_long, P, P, P, P, P, P]),
            "or_adjoint_beam": (None, [c_int, P, P, P, c_int, c_long, P, c_long, P, P, P, P, P, P]),
            "or_disc_sizes": (c_long, [c_int, c_double, c_long, P, P, P]),
            "or_check_cap_division": (c_long, [c_int]),
            "or_vec_sin": (None, [c_long, P, P]),
            "or_vec_cos": (None, [c_long, P, P]),
            "or_vec_acos": (None, [c_long, P, P]),
            "or_vec_asin": (None, [c_long, P, P]),
            "or_vec_atan2": (None, [c_long, P, P, P]),
            "or_skymap_accumulate": (c_long, [c_int, c_long, P, P, P, c_long, c_int, P, P, P, P]),
            "or_skymap_solve": (c_long, [c_int, P, P, P, P, P, P]),
            "or_dgesv3_many": (None, [c_long, P, P, P]),
            "or_set_beam_interpolation": (None, [c_int]),
            "or_set_v_term": (None, [P]),
            "or_focal_plane_to_sky": (None, [c_int, P, P, P, c_long, P, P, P, c_int, P, P, P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a

    # --- HEALPix
    def pix2ang(self, nside, pix):
        pix = np.ascontiguousarray(pix, np.int32)
        th = np.empty(pix.size)
        ph = np.empty(pix.size)
        self.L.or_pix2ang(nside, pix.size, _a(pix), _a(th), _a(ph))
        return th, ph

    def ang2pix(self, nside, theta, phi):
        th = np.ascontiguousarray(theta, np.float64)
        ph = np.ascontiguousarray(phi, np.float64)
        out = np.empty(th.size, np.int32)
        self.L.or_ang2pix(nside, th.size, _a(th), _a(ph), _a(out))
        return out

    def ring_above(self, nside, z):
        z = np.ascontiguousarray(z, np.float64)
        out = np.empty(z.size, np.int32)
        self.L.or_ring_above(nside, z.size, _a(z), _a(out))
        return out

    def get_interpol(self, nside, theta, phi):
        th = np.ascontiguousarray(theta, np.float64)
        ph = np.ascontiguousarray(phi, np.float64)
        pix = np.empty((th.size, 4), np.int32)
        wgt = np.empty((th.size, 4))
        self.L.or_get_interpol(nside, th.size, _a(th), _a(ph), _a(pix), _a(wgt))
        return pix, wgt

    def query_disc(self, nside, theta, phi, radius):
        cap = 4096
        while True:
            rr = np.empty(2 * cap, np.int32)
            n = self.L.or_query_disc_ranges(nside, theta, phi, radius, _a(rr), cap)
            if n <= cap:
                return rr[:2 * n].reshape(-1, 2)
            cap = n

    def query_disc_many(self, nside, ra, dec, radius):
        ra = np.ascontiguousarray(ra, np.float32)
        dec = np.ascontiguousarray(dec, np.float32)
        n = ra.size
        off = np.empty(n + 1, np.int64)
        total = self.L.or_query_disc_many(nside, n, _a(ra), _a(dec), radius, _a(off), None)
        pix = np.empty(max(total, 1), np.int32)
        self.L.or_query_disc_many(nside, n, _a(ra), _a(dec), radius, _a(off), _a(pix))
        return off, pix[:total]

    def rho_sigma_chi(self, ra_bc, dec_bc, psi_bc, ra_pix, dec_pix):
        arrs = [np.ascontiguousarray(x, np.float64) for x in (ra_bc, dec_bc, psi_bc, ra_pix, dec_pix)]
        n = arrs[0].size
        out = [np.empty(n) for _ in range(3)]
        self.L.or_rho_sigma_chi(n, *[_a(x) for x in arrs], *[_a(o) for o in out])
        return out

    def match(self, nside_sky, nside_beam, ra, dec, pa, off, pix):
        ra, dec, pa = (np.ascontiguousarray(x, np.float32) for x in (ra, dec, pa))
        bpix = np.empty(pix.size, np.int32)
        chi = np.empty(pix.size)
        self.L.or_match(nside_sky, nside_beam, ra.size, _a(ra), _a(dec), _a(pa), _a(off),
                        _a(pix), _a(bpix), _a(chi))
        return bpix, chi

    # --- PolBeam
    def beam_rho_max(self, nside_beam, npix):
        return self.L.or_beam_rho_max(nside_beam, npix)

    def gaussian_polbeam(self, nside_beam, npix, fwhmx, fwhmy, phi0, epsilon, dets, nside_sky):
        f = [np.empty(npix) for _ in range(4)]
        self.L.or_gaussian_fields(nside_beam, npix, fwhmx, fwhmy, phi0, *[_a(x) for x in f])
        tilde = [np.zeros(npix, np.float32) for _ in range(8)]
        if dets in "ap":
            self.L.or_beam_from_fields(npix, *[_a(x) for x in f], *[_a(t) for t in tilde[:4]])
        if dets in "bp":
            self.L.or_beam_from_fields(npix, *[_a(x) for x in f], *[_a(t) for t in tilde[4:]])
        out = [np.empty(npix, np.float32) for _ in range(8)]
        self.L.or_polbeam_build(npix, nside_beam, nside_sky, epsilon, dets.encode(),
                                _parr(tilde), _parr(out))
        return out, tilde, f

    def beam_from_fields(self, npix, magEco, phsEco, magEcx, phsEcx):
        """tilde I, Q, U, V (float32) of one detector from its E fields."""
        f = [np.ascontiguousarray(x, np.float64) for x in (magEco, phsEco, magEcx, phsEcx)]
        tilde = [np.zeros(npix, np.float32) for _ in range(4)]
        self.L.or_beam_from_fields(npix, *[_a(x) for x in f], *[_a(t) for t in tilde])
        return tilde

    def polbeam_build(self, npix, nside_beam, nside_sky, epsilon, dets, tilde):
        """aBeams[0..3] + bBeams[0..3] from tilde Ia..Va, Ib..Vb."""
        tilde = [np.ascontiguousarray(t, np.float32) for t in tilde]
        out = [np.empty(npix, np.float32) for _ in range(8)]
        self.L.or_polbeam_build(npix, nside_beam, nside_sky, epsilon, dets.encode(),
                                _parr(tilde), _parr(out))
        return out

    def polbeam_from_fields(self, npix, nside_beam, nside_sky, epsilon, fields_a, fields_b):
        """A detector pair built from E fields: beam_from_fields for a and b,
        then build (polbeam.cpp:117-288)."""
        tilde = (self.beam_from_fields(npix, *fields_a) +
                 self.beam_from_fields(npix, *fields_b))
        return self.polbeam_build(npix, nside_beam, nside_sky, epsilon, "p", tilde)

    # --- Convolver
    def convolve(self, nside_sky, I, Q, U, nside_beam, beams, ra, dec, pa, polflag="p",
                 precision64=False, nthreads=0):
        beams = [np.ascontiguousarray(b, np.float32) for b in beams]
        npix = beams[0].size
        I, Q, U = (np.ascontiguousarray(x, np.float32) for x in (I, Q, U))
        ra, dec, pa = (np.ascontiguousarray(x, np.float32) for x in (ra, dec, pa))
        n = ra.size
        ta = np.zeros(n, np.float32)
        tb = np.zeros(n, np.float32)
        da = np.zeros(n)
        db = np.zeros(n)
        rc = self.L.or_convolve(nside_sky, _a(I), _a(Q), _a(U), nside_beam, npix, _parr(beams),
                                n, _a(ra), _a(dec), _a(pa), polflag.encode(), int(precision64),
                                nthreads, _a(ta), _a(tb), _a(da), _a(db))
        if rc != 0:
            raise ValueError("valid polFlags are 'a', 'b' and 'p'")
        return (da, db) if precision64 else (ta, tb)

    def adjoint_map(self, nside_sky, nside_beam, beams, ra, dec, pa, g_a, g_b):
        beams = [np.ascontiguousarray(b, np.float32) for b in beams]
        ra, dec, pa = (np.ascontiguousarray(x, np.float32) for x in (ra, dec, pa))
        g_a = None if g_a is None else np.ascontiguousarray(g_a, np.float32)
        g_b = None if g_b is None else np.ascontiguousarray(g_b, np.float32)
        gmap = np.empty((12 * nside_sky ** 2, 3))
        self.L.or_adjoint_map(nside_sky, nside_beam, beams[0].size, _parr(beams), ra.size,
                              _a(ra), _a(dec), _a(pa), _a(g_a), _a(g_b), _a(gmap))
        return gmap[:, 0].copy(), gmap[:, 1].copy(), gmap[:, 2].copy()

    def adjoint_beam(self, nside_sky, I, Q, U, nside_beam, beams, ra, dec, pa, g_a, g_b):
        beams = [np.ascontiguousarray(b, np.float32) for b in beams]
        I, Q, U = (np.ascontiguousarray(x, np.float32) for x in (I, Q, U))
        ra, dec, pa = (np.ascontiguousarray(x, np.float32) for x in (ra, dec, pa))
        g_a = None if g_a is None else np.ascontiguousarray(g_a, np.float32)
        g_b = None if g_b is None else np.ascontiguousarray(g_b, np.float32)
        npix = beams[0].size
        g = np.empty((npix, 6))
        self.L.or_adjoint_beam(nside_sky, _a(I), _a(Q), _a(U), nside_beam, npix, _parr(beams),
                               ra.size, _a(ra), _a(dec), _a(pa), _a(g_a), _a(g_b), _a(g))
        return g

    def disc_sizes(self, nside_sky, radius, ra, dec):
        ra, dec = (np.ascontiguousarray(x, np.float32) for x in (ra, dec))
        sizes = np.empty(ra.size, np.int32)
        total = self.L.or_disc_sizes(nside_sky, radius, ra.size, _a(ra), _a(dec), _a(sizes))
        return total, sizes

    # --- SkyMap
    def skymap_accumulate(self, nside, ra, dec, pa, pol_angles, data, AtA=None, AtD=None,
                          shared_pointing=False):
        """ra, dec, pa: [ndet, n] (or [n] with shared_pointing); data [ndet, n].
        Returns (AtA [npix, 9], AtD [npix, 3], skipped)."""
        data = np.ascontiguousarray(data, np.float32)
        ndet, n = data.shape
        ra, dec, pa = (np.ascontiguousarray(x, np.float32) for x in (ra, dec, pa))
        pol = np.ascontiguousarray(pol_angles, np.float32)
        npix = 12 * nside * nside
        AtA = np.zeros((npix, 9)) if AtA is None else AtA
        AtD = np.zeros((npix, 3)) if AtD is None else AtD
        skipped = self.L.or_skymap_accumulate(nside, n, _a(ra), _a(dec), _a(pa),
                                              0 if shared_pointing else n, ndet, _a(pol),
                                              _a(data), _a(AtA), _a(AtD))
        return AtA, AtD, skipped

    def skymap_solve(self, nside, AtA, AtD):
        npix = 12 * nside * nside
        AtA = np.ascontiguousarray(AtA, np.float64)
        AtD = np.ascontiguousarray(AtD, np.float64)
        I, Q, U = (np.empty(npix, np.float32) for _ in range(3))
        hits = np.empty(npix, np.int32)
        singular = self.L.or_skymap_solve(nside, _a(AtA), _a(AtD), _a(I), _a(Q), _a(U),
                                          _a(hits))
        return I, Q, U, hits, singular

    def dgesv3(self, a, b):
        """a: [n, 3, 3] (row i, column j), b: [n, 3] -> (x, info)."""
        a = np.ascontiguousarray(np.swapaxes(np.asarray(a, np.float64), 1, 2))  # column-major
        b = np.array(b, np.float64, order="C", copy=True)
        info = np.empty(a.shape[0], np.int32)
        self.L.or_dgesv3_many(a.shape[0], _a(a), _a(b), _a(info))
        return b, info

    def set_beam_interpolation(self, on):
        """Selects the oracle's beam interpolation mode for convolve and the
        adjoints (process-wide; tests restore it)."""
        self.L.or_set_beam_interpolation(int(bool(on)))

    def set_v_term(self, V):
        """V-term mode of the forward (process-wide; tests restore it with
        None): the sky V map, or None for the CPU Convolver."""
        self._v = None if V is None else np.ascontiguousarray(V, np.float32)
        self.L.or_set_v_term(_a(self._v))

    # --- BPoint
    def focal_plane_to_sky(self, dx, dy, pol, x_ac, y_ac, rot_ac, flag):
        dx, dy, pol = (np.ascontiguousarray(v, np.float64) for v in (dx, dy, pol))
        x_ac, y_ac, rot_ac = (np.ascontiguousarray(v, np.float64) for v in (x_ac, y_ac, rot_ac))
        nd, n = dx.size, x_ac.size
        out = [np.empty((nd, n)) for _ in range(3)]
        self.L.or_focal_plane_to_sky(nd, _a(dx), _a(dy), _a(pol), n, _a(x_ac), _a(y_ac),
                                     _a(rot_ac), int(flag), *[_a(o) for o in out])
        return out

    def vec(self, name, x, y=None):
        x = np.ascontiguousarray(x, np.float64)
        o = np.empty_like(x)
        if y is None:
            getattr(self.L, "or_vec_" + name)(x.size, _a(x), _a(o))
        else:
            y = np.ascontiguousarray(y, np.float64)
            getattr(self.L, "or_vec_" + name)(x.size, _a(x), _a(y), _a(o))
        return o


def reference_sphtrigo():
    """The reference's own rho_sigma_chi_pix (src/Sphtrigo/sphtrigo.cpp),
    built by oracle/Makefile into oracle/_ref/ when the reference tree was
    present at build time; None otherwise."""
    path = os.path.join(ORACLE_DIR, "_ref", "libsphtrigo_ref.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    f = L.ref_rho_sigma_chi
    f.restype = None
    f.argtypes = [c_long] + [P] * 8

    def rsc(ra_bc, dec_bc, psi_bc, ra_pix, dec_pix):
        arrs = [np.ascontiguousarray(x, np.float64) for x in (ra_bc, dec_bc, psi_bc, ra_pix, dec_pix)]
        n = arrs[0].size
        out = [np.empty(n) for _ in range(3)]
        f(n, *[_a(x) for x in arrs], *[_a(o) for o in out])
        return out
    return rsc
