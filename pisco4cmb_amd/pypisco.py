"""pypisco: Python mirror of PISCO's convolution interface over libpisco.

The classes keep the reference's names, argument meaning and call order:

    sky  = Sky(nside, I, Q, U, V)                      # src/Sky/sky.hpp:15-20
    scan = Scan(nsamples, ra, dec, pa)                 # src/Scan/scan.hpp:14-18
    beam = PolBeam(nside_beam, nPixels, epsilon, 'p')  # src/Polbeam/polbeam.hpp:26-29
    beam.make_unpol_gaussian_elliptical_beams(fwhmx, fwhmy, phi0)
    beam.build(nside_sky)
    conv = GPUConvolver(RunConfig(...))                # gpuconvolver.hpp:21
    conv.configure_execution()
    conv.update_sky(sky); conv.update_beam(beam); conv.update_scan(scan)
    conv.exec_convolution(data_a, data_b)              # fills float32 arrays

Convolver(scan, sky, beam).exec_convolution(polFlag, data_a, data_b) keeps the
CPU class's signature (convolver.hpp:18-20) and runs on the GPU through the
same library.  Errors: invalid configurations or flags raise ValueError (the
reference throws std::invalid_argument), call-order errors raise
RuntimeError; nothing exits the process.

Arrays are numpy float32 (pointing in radians, pa counter-clockwise).
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import RunConfig, check, lib, ptr_array

__all__ = ["RunConfig", "Sky", "Scan", "PolBeam", "GPUConvolver", "Convolver", "SkyMap", "BPoint",
           "read_txt_columns", "write_txt_columns", "read_sky_txt", "write_sky_txt",
           "default_run_config", "beam_npix_for_radius", "ang2pix", "pix2ang"]


def _f32(a, n=None, name="array"):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None and a.size < n:
        raise ValueError("%s has %d elements, %d needed" % (name, a.size, n))
    return a


def _addr(a):
    return a.ctypes.data if a is not None else None


def default_run_config(ptgPerConv=100000, pixelsPerDisc=0, deviceId=0,
                       maxMemUsage=1 << 40):
    """A RunConfig whose memory estimate is valid (fields as cudaConv2.h:13-26)."""
    return RunConfig(nStreams=1, maxMemUsage=maxMemUsage, deviceId=deviceId,
                     gridSizeX=0, gridSizeY=0, blockSizeX=0, blockSizeY=0,
                     ptgPerConv=ptgPerConv, pixelsPerDisc=pixelsPerDisc)


def ang2pix(nside, theta, phi):
    """Healpix_Base::ang2pix(pointing(theta, phi)) on the host (RING)."""
    r = lib.pisco_ang2pix(int(nside), float(theta), float(phi))
    if r < 0:
        raise ValueError("invalid nside or angle")
    return int(r)


def pix2ang(nside, pix):
    """Healpix_Base::pix2ang on the host (RING) -> (theta, phi)."""
    th, ph = ctypes.c_double(), ctypes.c_double()
    check(lib.pisco_pix2ang(int(nside), int(pix), ctypes.addressof(th),
                            ctypes.addressof(ph)))
    return th.value, ph.value


def beam_npix_for_radius(nside_beam, radius):
    """Number of beam pixels of a cap of the given radius (radians):
    ang2pix(nside, radius, 0), as scripts/fields2beams.py:102 sizes beams."""
    return ang2pix(nside_beam, radius, 0.0)


class Sky:
    """Non-owning view of a RING HEALPix I,Q,U,V sky (src/Sky/sky.cpp:4-17)."""

    def __init__(self, nside, I, Q, U, V=None):
        self.nside = int(nside)
        self.nPixels = 12 * self.nside * self.nside
        self.sI = _f32(I, self.nPixels, "I")
        self.sQ = _f32(Q, self.nPixels, "Q")
        self.sU = _f32(U, self.nPixels, "U")
        self.sV = _f32(V, self.nPixels, "V") if V is not None else np.zeros(
            self.nPixels, np.float32)

    def get_nside(self):
        return self.nside

    def size(self):
        return self.nPixels


class Scan:
    """Non-owning view of ra, dec, pa (src/Scan/scan.cpp:4-29)."""

    def __init__(self, nsamples, ra, dec, pa):
        self.nsamp = int(nsamples)
        self.ra = _f32(ra, self.nsamp, "ra")
        self.dec = _f32(dec, self.nsamp, "dec")
        self.pa = _f32(pa, self.nsamp, "pa")

    def get_nsamples(self):
        return self.nsamp

    def get_ra_ptr(self):
        return self.ra

    def get_dec_ptr(self):
        return self.dec

    def get_pa_ptr(self):
        return self.pa


class PolBeam:
    """Polarised beams of a detector pair on the first nPixels RING pixels of
    a map of NSIDE nside (src/Polbeam/polbeam.cpp:13-372)."""

    def __init__(self, nside, nPixels, epsilon, enabledDets):
        if enabledDets not in ("a", "b", "p"):
            raise ValueError("enabledDets must be 'a', 'b' or 'p'")
        self.nside = int(nside)
        self.nPixels = int(nPixels)
        self.epsilon = float(epsilon)
        self.enabledDets = enabledDets
        rho = lib.pisco_beam_rho_max(self.nside, self.nPixels)
        if rho < 0:
            raise ValueError("invalid beam geometry (nside=%d, nPixels=%d)"
                             % (self.nside, self.nPixels))
        self.rhoMax = rho
        n = self.nPixels
        self.Ia, self.Qa, self.Ua, self.Va = (np.zeros(n, np.float32) for _ in range(4))
        self.Ib, self.Qb, self.Ub, self.Vb = (np.zeros(n, np.float32) for _ in range(4))
        self.aBeams = [np.zeros(n, np.float32) for _ in range(4)]
        self.bBeams = [np.zeros(n, np.float32) for _ in range(4)]

    def get_nside(self):
        return self.nside

    def size(self):
        return self.nPixels

    def get_rho_max(self):
        return self.rhoMax

    def beam_from_fields(self, polFlag, magEco, phaseEco, magEcx, phaseEcx):
        """tilde I,Q,U,V of detector polFlag from its E fields (polbeam.cpp:117-184)."""
        if polFlag == "a":
            out = (self.Ia, self.Qa, self.Ua, self.Va)
        elif polFlag == "b":
            out = (self.Ib, self.Qb, self.Ub, self.Vb)
        else:
            raise ValueError("polFlag must be 'a' or 'b'")
        f = [np.ascontiguousarray(x, dtype=np.float64) for x in
             (magEco, phaseEco, magEcx, phaseEcx)]
        for x in f:
            if x.size < self.nPixels:
                raise ValueError("field arrays need nPixels elements")
        check(lib.pisco_beam_from_fields(self.nPixels, *[_addr(x) for x in f],
                                         *[_addr(o) for o in out]))

    def build(self, nsideSky):
        """Normalise and mix with leakage epsilon (polbeam.cpp:186-288)."""
        tilde = ptr_array([_addr(x) for x in (self.Ia, self.Qa, self.Ua, self.Va,
                                              self.Ib, self.Qb, self.Ub, self.Vb)])
        out = ptr_array([_addr(x) for x in self.aBeams + self.bBeams])
        check(lib.pisco_polbeam_build(self.nPixels, self.nside, int(nsideSky),
                                      self.epsilon, self.enabledDets.encode(),
                                      tilde, out))

    def _tilde(self, polFlag):
        if polFlag == "a":
            return (self.Ia, self.Qa, self.Ua, self.Va)
        if polFlag == "b":
            return (self.Ib, self.Qb, self.Ub, self.Vb)
        raise ValueError("polFlag must be 'a' or 'b'")

    def read_tilde_beams(self, path, polFlag):
        """tilde I, Q, U, V of detector polFlag from a text file of nPixels
        rows, the file scripts/fields2beams.py writes (hpx_tilde_beams_*);
        build() then normalises them."""
        cols = read_txt_columns(path, 4)
        if cols[0].size != self.nPixels:
            raise ValueError("%s has %d rows, the beam has %d pixels"
                             % (path, cols[0].size, self.nPixels))
        for dst, src in zip(self._tilde(polFlag), cols):
            dst[:] = src

    def write_tilde_beams(self, path, polFlag):
        write_txt_columns(path, self._tilde(polFlag))

    def make_unpol_gaussian_elliptical_beams(self, fwhmx, fwhmy, phi0):
        """Co-polarised elliptical Gaussian fields (polbeam.cpp:290-372); FWHM
        and phi0 in degrees."""
        n = self.nPixels
        f = [np.empty(n, np.float64) for _ in range(4)]
        check(lib.pisco_beam_gaussian_fields(self.nside, n, float(fwhmx), float(fwhmy),
                                             float(phi0), *[_addr(x) for x in f]))
        if self.enabledDets in ("a", "p"):
            self.beam_from_fields("a", *f)
        if self.enabledDets in ("b", "p"):
            self.beam_from_fields("b", *f)


class GPUConvolver:
    """Drop-in for GPUConvolver (src/GpuConvolver/gpuconvolver.hpp:17-82)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.ctx = None
        self.nside_sky = None
        self.npix_beam = None
        self.nsamples = None

    def configure_execution(self):
        ctx = ctypes.c_void_p()
        rc = lib.pisco_create(ctypes.byref(self.cfg), ctypes.byref(ctx))
        if rc in (_lib.PISCO_ERR_INVALID, _lib.PISCO_ERR_MEMORY):
            raise ValueError("FATAL: GPUConvolver has an invalid configuration: "
                             + lib.pisco_last_error().decode())
        check(rc)
        self.ctx = ctx

    def _need_ctx(self):
        if self.ctx is None:
            raise RuntimeError("configure_execution() must be called first")

    def set_tuning(self, key, value):
        """Launch-shape tuning of this context (see pisco_set_tuning)."""
        self._need_ctx()
        rc = lib.pisco_set_tuning(self.ctx, key.encode(), int(value))
        if rc == _lib.PISCO_ERR_INVALID:
            raise ValueError(lib.pisco_last_error().decode())
        check(rc)

    def get_tuning(self, key):
        self._need_ctx()
        v = ctypes.c_long()
        check(lib.pisco_get_tuning(self.ctx, key.encode(), ctypes.byref(v)))
        return v.value

    def update_sky(self, sky):
        self._need_ctx()
        check(lib.pisco_update_sky(self.ctx, sky.nside, _addr(sky.sI), _addr(sky.sQ),
                                   _addr(sky.sU), _addr(sky.sV)))
        self.nside_sky = sky.nside

    def update_beam(self, beam):
        self._need_ctx()
        a = ptr_array([_addr(x) for x in beam.aBeams])
        b = ptr_array([_addr(x) for x in beam.bBeams])
        check(lib.pisco_update_beam(self.ctx, beam.nside, beam.nPixels, a, b))
        self.npix_beam = beam.nPixels

    def update_scan(self, scan):
        self._need_ctx()
        check(lib.pisco_update_scan(self.ctx, scan.nsamp, _addr(scan.ra), _addr(scan.dec),
                                    _addr(scan.pa)))
        self.nsamples = scan.nsamp

    def _outputs(self, *arrs):
        for a in arrs:
            if a is not None and (a.dtype != np.float32 or not a.flags.c_contiguous
                                  or a.size < (self.nsamples or 0)):
                raise ValueError("outputs must be contiguous float32 arrays of "
                                 "nsamples elements")

    def exec_convolution(self, data_a, data_b, polFlag="p"):
        self._need_ctx()
        self._outputs(data_a, data_b)
        rc = lib.pisco_exec_convolution(self.ctx, polFlag.encode(), _addr(data_a),
                                        _addr(data_b))
        if rc == _lib.PISCO_ERR_INVALID:
            raise ValueError(lib.pisco_last_error().decode())
        check(rc)

    def exec_adjoint_map(self, g_a, g_b):
        """A^T applied to TOD cotangents -> (gI, gQ, gU) float32 maps."""
        self._need_ctx()
        n = 12 * self.nside_sky ** 2
        ga = _f32(g_a, self.nsamples) if g_a is not None else None
        gb = _f32(g_b, self.nsamples) if g_b is not None else None
        out = [np.empty(n, np.float32) for _ in range(3)]
        check(lib.pisco_exec_adjoint_map(self.ctx, _addr(ga), _addr(gb),
                                         *[_addr(o) for o in out]))
        return tuple(out)

    def exec_adjoint_beam(self, g_a, g_b):
        """Gradient of <g, TOD> w.r.t. aBeams[0..3], bBeams[0..3]."""
        self._need_ctx()
        n = self.npix_beam
        ga = _f32(g_a, self.nsamples) if g_a is not None else None
        gb = _f32(g_b, self.nsamples) if g_b is not None else None
        oa = [np.empty(n, np.float32) for _ in range(4)]
        ob = [np.empty(n, np.float32) for _ in range(4)]
        check(lib.pisco_exec_adjoint_beam(self.ctx, _addr(ga), _addr(gb),
                                          ptr_array([_addr(o) for o in oa]),
                                          ptr_array([_addr(o) for o in ob])))
        return oa, ob

    def close(self):
        if self.ctx is not None:
            lib.pisco_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Convolver:
    """Convolver(Scan*, Sky*, PolBeam*) (src/Convolver/convolver.hpp:14-36),
    computed on the GPU."""

    def __init__(self, scan, sky, beam, cfg=None):
        self.scan, self.sky, self.beam = scan, sky, beam
        self._gpu = GPUConvolver(cfg if cfg is not None else default_run_config())
        self._gpu.configure_execution()
        self._gpu.update_sky(sky)
        self._gpu.update_beam(beam)
        self._gpu.update_scan(scan)

    def exec_convolution(self, polFlag, data_a, data_b):
        if polFlag not in ("a", "b", "p"):
            raise ValueError("valid polFlags are 'a', 'b' and 'p'")
        self._gpu.exec_convolution(data_a if polFlag != "b" else None,
                                   data_b if polFlag != "a" else None, polFlag)


class SkyMap:
    """SkyMap(nside) (src/SkyMap/skymap.hpp:13-55): binned I, Q, U map-maker,
    its normal equations held on a GPU.

        m = SkyMap(nside)
        m.accumulate_data(nsamples, ra, dec, pa, ndetectors, detPolAngles, data)
        m.solve_map()
        I, Q, U, hits = m.get_stokes_I(), m.get_stokes_Q(), m.get_stokes_U(), m.get_hitmap()

    ra, dec, pa and data hold ndetectors * nsamples float32 values,
    detector-major (skymap.cpp:91-97).  The accumulators start at zero.  A
    sample with theta = pi/2 - dec outside [0, pi] (where the reference's
    Healpix_Base::ang2pix throws) raises ValueError after the call has
    accumulated every other sample."""

    INSUFFCOV = 0

    def __init__(self, nside, device=0):
        self.nside = int(nside)
        self.nPixels = 12 * self.nside * self.nside
        h = ctypes.c_void_p()
        check(lib.pisco_skymap_create(int(device), self.nside, ctypes.byref(h)))
        self.handle = h
        self.stokesI = np.zeros(self.nPixels, np.float32)
        self.stokesQ = np.zeros(self.nPixels, np.float32)
        self.stokesU = np.zeros(self.nPixels, np.float32)
        self.hitsMap = np.zeros(self.nPixels, np.int32)
        self.singular = 0

    def get_nside(self):
        return self.nside

    def size(self):
        return self.nPixels

    def accumulate_data(self, nsamples, ra, dec, pa, ndetectors, detPolAngles, data):
        n = int(nsamples) * int(ndetectors)
        ra, dec, pa, data = (_f32(a, n, nm) for a, nm in ((ra, "ra"), (dec, "dec"), (pa, "pa"),
                                                          (data, "data")))
        pol = _f32(detPolAngles, int(ndetectors), "detPolAngles")
        rc = lib.pisco_skymap_accumulate(self.handle, int(nsamples), _addr(ra), _addr(dec),
                                         _addr(pa), int(ndetectors), _addr(pol), _addr(data))
        if rc == _lib.PISCO_ERR_INVALID:
            raise ValueError(lib.pisco_last_error().decode())
        check(rc)

    def solve_map(self):
        sing = ctypes.c_long()
        check(lib.pisco_skymap_solve(self.handle, _addr(self.stokesI), _addr(self.stokesQ),
                                     _addr(self.stokesU), _addr(self.hitsMap),
                                     ctypes.byref(sing)))
        self.singular = sing.value

    def get_stokes_I(self):
        return self.stokesI

    def get_stokes_Q(self):
        return self.stokesQ

    def get_stokes_U(self):
        return self.stokesU

    def get_hitmap(self):
        return self.hitsMap

    def get_normal(self):
        """(AtA [npix, 9], AtD [npix, 3]) as the reference stores them."""
        AtA = np.empty((self.nPixels, 9))
        AtD = np.empty((self.nPixels, 3))
        check(lib.pisco_skymap_get_normal(self.handle, _addr(AtA), _addr(AtD)))
        return AtA, AtD

    def reset(self):
        check(lib.pisco_skymap_reset(self.handle))

    def set_chunk(self, samples):
        check(lib.pisco_skymap_set_chunk(self.handle, int(samples)))

    def close(self):
        if getattr(self, "handle", None) is not None:
            lib.pisco_skymap_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BPoint:
    """BPoint::focal_plane_to_sky (src/Bpoint/bpoint.cpp:553-674) computed on
    the GPU in FP64 (the reference's SOFA-based astrometry is not part of
    this library).  Outputs are filled in place, detector-major."""

    def __init__(self, deviceId=0):
        self._gpu = GPUConvolver(default_run_config(deviceId=deviceId))
        self._gpu.configure_execution()

    def focal_plane_to_sky(self, ndets, det_dx, det_dy, det_pol_angle, nsamples, x_ac, y_ac,
                           rot_ac, flag, x_det, y_det, rot_det):
        nd, n = int(ndets), int(nsamples)
        ins = [np.ascontiguousarray(a, np.float64) for a in (det_dx, det_dy, det_pol_angle)]
        ins += [np.ascontiguousarray(a, np.float64) for a in (x_ac, y_ac, rot_ac)]
        for a, want in zip(ins, (nd, nd, nd, n, n, n)):
            if a.size < want:
                raise ValueError("input array too short")
        outs = (x_det, y_det, rot_det)
        for o in outs:
            if (not isinstance(o, np.ndarray) or o.dtype != np.float64 or o.size < nd * n
                    or not o.flags.c_contiguous):
                raise ValueError("outputs must be C-contiguous float64 arrays of ndets*nsamples")
        self._gpu._need_ctx()
        rc = lib.pisco_focal_plane_to_sky(self._gpu.ctx, nd, *[_addr(a) for a in ins[:3]],
                                          n, *[_addr(a) for a in ins[3:]], int(flag),
                                          *[_addr(o) for o in outs])
        if rc == _lib.PISCO_ERR_INVALID:
            raise ValueError(lib.pisco_last_error().decode())
        check(rc)

    def close(self):
        self._gpu.close()


# ---- text files of the reference's formats -------------------------------

def read_txt_columns(path, ncols):
    """Columns of a whitespace-separated text file (numpy.savetxt format) as
    float32 arrays, parsed natively (pisco_read_txt_columns)."""
    n = ctypes.c_long()
    p = str(path).encode()
    rc = lib.pisco_read_txt_columns(p, int(ncols), 0, None, ctypes.byref(n))
    if rc == _lib.PISCO_ERR_INVALID:
        raise ValueError(lib.pisco_last_error().decode())
    check(rc)
    cols = [np.empty(n.value, np.float32) for _ in range(int(ncols))]
    rc = lib.pisco_read_txt_columns(p, int(ncols), n.value, ptr_array([_addr(c) for c in cols]),
                                    ctypes.byref(n))
    if rc == _lib.PISCO_ERR_INVALID:
        raise ValueError(lib.pisco_last_error().decode())
    check(rc)
    return cols


def write_txt_columns(path, cols):
    """Writes equal-length float32 columns as numpy.savetxt does."""
    cols = [np.ascontiguousarray(c, np.float32) for c in cols]
    if len({c.size for c in cols}) != 1:
        raise ValueError("columns must have one length")
    rc = lib.pisco_write_txt_columns(str(path).encode(), len(cols), cols[0].size,
                                     ptr_array([_addr(c) for c in cols]))
    if rc == _lib.PISCO_ERR_INVALID:
        raise ValueError(lib.pisco_last_error().decode())
    check(rc)


def read_sky_txt(path):
    """A Sky from a text map of I Q U V rows per RING pixel, the file
    scripts/create_cmb.py writes (numpy.savetxt of [I, Q, U, V].T); NSIDE
    from the row count."""
    I, Q, U, V = read_txt_columns(path, 4)
    nside = int(round(math.sqrt(I.size / 12.0)))
    if 12 * nside * nside != I.size or nside < 1:
        raise ValueError("%s has %d rows, not 12 nside^2" % (path, I.size))
    return Sky(nside, I, Q, U, V)


def write_sky_txt(path, sky):
    write_txt_columns(path, [sky.sI, sky.sQ, sky.sU,
                             sky.sV if sky.sV is not None else np.zeros_like(sky.sI)])
