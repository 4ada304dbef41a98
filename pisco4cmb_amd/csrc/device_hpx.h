/*
 * device_hpx.h -- HEALPix RING geometry and the per-visit beam-frame mapping
 * for gfx950 kernels.
 *
 * Semantics follow the CPU path of PISCO (Convolver, convolver.cpp:74-183),
 * which takes pix2ang / ang2pix / query_disc from healpix_cxx's Healpix_Base
 * (RING scheme, non-inclusive disc), and the rotation of
 * SphericalTransformations::rho_sigma_chi_pix (sphtrigo.cpp:4-55).  The device
 * port of those routines in the reference (cudaHealpix.cu:30-259) differs from
 * the CPU library in three places (its fmod, its polar "sth" rule and the
 * missing imodulo); the CPU behaviour is the one reproduced here.
 *
 * All FP64 arithmetic is written so that, compiled with -ffp-contract=off and
 * evaluated with pisco_libm.h, it yields the same bits as the CPU oracle:
 * pixel indices are therefore bit-exact, not merely close.
 */
#ifndef PISCO_DEVICE_HPX_H
#define PISCO_DEVICE_HPX_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pisco_libm.h"

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

namespace pisco {

constexpr double kPi = 3.141592653589793238462643383279502884197;
constexpr double kTwoPi = 6.283185307179586476925286766559005768394;
constexpr double kHalfPi = 1.570796326794896619231321691639751442099;
constexpr double kInvHalfPi = 0.6366197723675813430755350534900574;
constexpr double kInvTwoPi = 1.0 / 6.283185307179586476925286766559005768394;
constexpr double kTwoThird = 2.0 / 3.0;

/* Geometry of one RING-ordered HEALPix map (NSIDE <= 8192, int32 pixels,
 * as Healpix_Base<int>). */
struct Hpx {
    int nside;
    int ncap;   /* pixels in the north polar cap */
    int npix;
    double fact1, fact2;
};

__host__ __device__ inline Hpx hpx_make(int nside)
{
    Hpx h;
    h.nside = nside;
    h.ncap = 2 * (nside * nside - nside);
    h.npix = 12 * nside * nside;
    h.fact2 = 4. / h.npix;
    h.fact1 = (nside << 1) * h.fact2;
    return h;
}

/* fmod(x, y) for finite x and y > 0, exact (IEEE fmod is exact, and so is
 * this: the candidate quotient k is off by at most one and x - k*y is
 * representable for the right k, so one fma returns it unrounded). */
__host__ __device__ inline double exact_fmod(double x, double y)
{
    double ax = fabs(x);
    if (ax < y)
        return x;
    double k = trunc(ax / y);
    double r = fma(-k, y, ax);
    if (r < 0.0) {
        k -= 1.0;
        r = fma(-k, y, ax);
    } else if (r >= y) {
        k += 1.0;
        r = fma(-k, y, ax);
    }
    return copysign(r, x);
}

/* healpix_cxx fmodulo: result in [0, v2). */
__host__ __device__ inline double fmodulo(double v1, double v2)
{
    if (v1 >= 0)
        return (v1 < v2) ? v1 : exact_fmod(v1, v2);
    double tmp = exact_fmod(v1, v2) + v2;
    return (tmp == v2) ? 0. : tmp;
}

__host__ __device__ inline int isqrt_i(int v)
{
    return (int)sqrt((double)v + 0.5);
}

/* Ring index (1..4N-1) of a RING pixel. */
__host__ __device__ inline int pix2ring(const Hpx &h, int pix)
{
    if (pix < h.ncap)
        return (1 + isqrt_i(1 + 2 * pix)) >> 1;
    if (pix < h.npix - h.ncap)
        return (pix - h.ncap) / (4 * h.nside) + h.nside;
    int ip = h.npix - pix;
    return 4 * h.nside - ((1 + isqrt_i(2 * ip - 1)) >> 1);
}

/* First pixel, number of pixels and shift flag of ring r
 * (Healpix_Base::get_ring_info_small). */
__host__ __device__ inline void ring_info(const Hpx &h, int r, int &startpix, int &ringpix,
                                 bool &shifted)
{
    if (r < h.nside) {
        shifted = true;
        ringpix = 4 * r;
        startpix = 2 * r * (r - 1);
    } else if (r < 3 * h.nside) {
        shifted = ((r - h.nside) & 1) == 0;
        ringpix = 4 * h.nside;
        startpix = h.ncap + (r - h.nside) * ringpix;
    } else {
        shifted = true;
        int nr = 4 * h.nside - r;
        ringpix = 4 * nr;
        startpix = h.npix - 2 * nr * (nr + 1);
    }
}

/* z and (if |z| > 0.99) sin(theta) of the pixels of ring r, as pix2loc
 * computes them. */
__host__ __device__ inline void ring_z(const Hpx &h, int r, double &z, double &sth,
                              bool &have_sth)
{
    have_sth = false;
    sth = 0.0;
    int n = h.nside;
    if (r < n) {
        double tmp = (r * r) * h.fact2;
        z = 1.0 - tmp;
        if (z > 0.99) {
            sth = sqrt(tmp * (2.0 - tmp));
            have_sth = true;
        }
    } else if (r <= 3 * n) {
        z = (2 * n - r) * h.fact1;
    } else {
        int ir = 4 * n - r;
        double tmp = (ir * ir) * h.fact2;
        z = tmp - 1.0;
        if (z < -0.99) {
            sth = sqrt(tmp * (2.0 - tmp));
            have_sth = true;
        }
    }
}

/* theta of ring r (pix2ang). */
__host__ __device__ inline double ring_theta(const Hpx &h, int r)
{
    double z, sth;
    bool hs;
    ring_z(h, r, z, sth, hs);
    return hs ? pl_atan2(sth, z) : pl_acos(z);
}

/* phi of pixel pix lying in ring r with first pixel startpix (pix2loc). */
__host__ __device__ inline double pixel_phi(const Hpx &h, int r, int pix, int startpix)
{
    int iphi = pix - startpix + 1;
    int n = h.nside;
    if (r < n)
        return (iphi - 0.5) * kHalfPi / r;
    if (r <= 3 * n) {
        double fodd = ((r + n) & 1) ? 1 : 0.5;
        return (iphi - fodd) * kPi * 0.75 * h.fact1;
    }
    return (iphi - 0.5) * kHalfPi / (4 * n - r);
}

__host__ __device__ inline void pix2ang(const Hpx &h, int pix, double &theta, double &phi)
{
    int r = pix2ring(h, pix);
    int sp, nr;
    bool sh;
    ring_info(h, r, sp, nr, sh);
    theta = ring_theta(h, r);
    phi = pixel_phi(h, r, pix, sp);
}

/* Per-ring constants the kernels read once per visit: cos and sin of the
 * ring's declination (dec = pi/2 - theta, exactly as pix2ang + the rotation
 * compute them), the first pixel, and for polar-cap rings the divisor of
 * the phi formula (ring index from the nearest pole) with its reciprocal. */
struct RingRec {
    double cdp, sdp;
    double inv;       /* 1.0 / rdiv (cap rings), 0 in the equatorial belt */
    int startpix;
    int rdiv;         /* r (north cap), 4N - r (south cap), 0 otherwise */
};

__host__ __device__ inline RingRec ring_record(const Hpx &h, int r)
{
    RingRec rr;
    double th = ring_theta(h, r);
    double dec = kPi / 2.0 - th;
    pl_sincos(dec, &rr.sdp, &rr.cdp);
    int nr;
    bool sh;
    ring_info(h, r, rr.startpix, nr, sh);
    rr.rdiv = (r < h.nside) ? r : ((r > 3 * h.nside) ? 4 * h.nside - r : 0);
    rr.inv = rr.rdiv ? 1.0 / rr.rdiv : 0.0;
    return rr;
}

/* pixel_phi from a ring record.  In the caps ((iphi - 0.5) * pi/2) / rdiv is
 * formed as q = x * inv, then q + (x - q*rdiv) * inv with fma: this equals
 * the IEEE quotient for every cap pixel of every NSIDE up to 8192 (checked
 * exhaustively, 134,201,344 cases), and costs no division. */
__host__ __device__ inline double ring_phi(const Hpx &h, const RingRec &rr, int r, int pix)
{
    int iphi = pix - rr.startpix + 1;
    if (rr.rdiv > 0) {
        double x = (iphi - 0.5) * kHalfPi;
        double q = x * rr.inv;
        double e = fma(-q, (double)rr.rdiv, x);
        return fma(e, rr.inv, q);
    }
    double fodd = ((r + h.nside) & 1) ? 1 : 0.5;
    return (iphi - fodd) * kPi * 0.75 * h.fact1;
}

/* Healpix_Base::loc2pix, RING. */
__host__ __device__ inline int loc2pix(const Hpx &h, double z, double phi, double sth,
                              bool have_sth)
{
    double za = fabs(z);
    double tt = fmodulo(phi * kInvHalfPi, 4.0);
    int nside = h.nside;
    if (za <= kTwoThird) {
        int nl4 = 4 * nside;
        double temp1 = nside * (0.5 + tt);
        double temp2 = nside * z * 0.75;
        int jp = (int)(temp1 - temp2);
        int jm = (int)(temp1 + temp2);
        int ir = nside + 1 + jp - jm;
        int kshift = 1 - (ir & 1);
        int t1 = jp + jm - nside + kshift + 1 + nl4 + nl4;
        int ip = (t1 >> 1) % nl4;
        return h.ncap + (ir - 1) * nl4 + ip;
    }
    double tp = tt - (int)tt;
    double tmp = ((za < 0.99) || (!have_sth)) ? nside * sqrt(3 * (1 - za))
                                              : nside * sth / sqrt((1. + za) / 3.);
    int jp = (int)(tp * tmp);
    int jm = (int)((1.0 - tp) * tmp);
    int ir = jp + jm + 1;
    int ip = (int)(tt * ir);
    return (z > 0) ? 2 * ir * (ir - 1) + ip : h.npix - 2 * ir * (ir + 1) + ip;
}

/* Healpix_Base::ang2pix(pointing): sin(theta) is used only within 0.01 rad
 * of a pole. */
__host__ __device__ inline int ang2pix(const Hpx &h, double theta, double phi)
{
    if ((theta < 0.01) || (theta > 3.14159 - 0.01)) {
        double s, c;
        pl_sincos(theta, &s, &c);
        return loc2pix(h, c, phi, s, true);
    }
    return loc2pix(h, pl_cos(theta), phi, 0., false);
}

__host__ __device__ inline int ring_above(const Hpx &h, double z)
{
    double az = fabs(z);
    if (az <= kTwoThird)
        return (int)(h.nside * (2 - 1.5 * z));
    int iring = (int)(h.nside * sqrt(3 * (1 - az)));
    return (z > 0) ? iring : 4 * h.nside - iring - 1;
}

__host__ __device__ inline double ring2z(const Hpx &h, int ring)
{
    if (ring < h.nside)
        return 1 - ring * ring * h.fact2;
    if (ring <= 3 * h.nside)
        return (2 * h.nside - ring) * h.fact1;
    ring = 4 * h.nside - ring;
    return ring * ring * h.fact2 - 1;
}

/* ---- query_disc, decomposed by ring ------------------------------------ */

/* Per-pointing part of the non-inclusive RING query_disc
 * (Healpix_Base::query_disc with fact = 0). Rings are r_first..r_last;
 * rings outside [irmin, irmax] are whole (the polar blocks). */
struct Disc {
    double phi;       /* normalised phi of the centre */
    double z0, xa;
    double cosr;
    int irmin, irmax;
    int r_first, r_last;
};

__host__ __device__ inline void disc_setup(const Hpx &h, double theta, double phi,
                                  double radius, double cosr, Disc &d)
{
    /* pointing::normalize */
    theta = fmodulo(theta, kTwoPi);
    if (theta > kPi) {
        phi += kPi;
        theta = kTwoPi - theta;
    }
    phi = fmodulo(phi, kTwoPi);
    d.phi = phi;
    d.cosr = cosr;
    if (radius >= kPi) {            /* whole sky */
        d.irmin = 4 * h.nside;
        d.irmax = 0;
        d.r_first = 1;
        d.r_last = 4 * h.nside - 1;
        d.z0 = 0;
        d.xa = 0;
        return;
    }
    d.z0 = pl_cos(theta);
    d.xa = 1. / sqrt((1 - d.z0) * (1 + d.z0));
    double rlat1 = theta - radius;
    d.irmin = ring_above(h, pl_cos(rlat1)) + 1;
    bool north = (rlat1 <= 0) && (d.irmin > 1);
    double rlat2 = theta + radius;
    d.irmax = ring_above(h, pl_cos(rlat2));
    bool south = (rlat2 >= kPi) && (d.irmax + 1 < 4 * h.nside);
    d.r_first = north ? 1 : d.irmin;
    d.r_last = south ? 4 * h.nside - 1 : d.irmax;
}

/* Pixel ranges of ring r inside the disc, in ascending order: returns the
 * number of ranges (0..2); range k is [pix0[k], pix0[k] + len[k]). */
__host__ __device__ inline int disc_ring(const Hpx &h, const Disc &d, int r, int pix0[2],
                                int len[2])
{
    int ipix1, nr;
    bool shifted;
    ring_info(h, r, ipix1, nr, shifted);
    if (r < d.irmin || r > d.irmax) {
        pix0[0] = ipix1;
        len[0] = nr;
        return 1;
    }
    double z = ring2z(h, r);
    double x = (d.cosr - z * d.z0) * d.xa;
    double ysq = 1 - z * z - x * x;
    double dphi = (ysq <= 0) ? 0.0 : pl_atan2(sqrt(ysq), x);
    if (!(dphi > 0))
        return 0;
    double shift = shifted ? 0.5 : 0.;
    double a = nr * kInvTwoPi * (d.phi - dphi) - shift;
    double b = nr * kInvTwoPi * (d.phi + dphi) - shift;
    int ip_lo = ((a >= 0) ? (int)a : (int)a - 1) + 1;
    int ip_hi = (b >= 0) ? (int)b : (int)b - 1;
    if (ip_lo > ip_hi)
        return 0;
    if (ip_hi >= nr) {
        ip_lo -= nr;
        ip_hi -= nr;
    }
    if (ip_lo < 0) {   /* (constant indices: a runtime index would put the arrays in scratch) */
        if (ip_hi + 1 > 0) {
            pix0[0] = ipix1;
            len[0] = ip_hi + 1;
            pix0[1] = ipix1 + ip_lo + nr;
            len[1] = -ip_lo;
            return 2;
        }
        pix0[0] = ipix1 + ip_lo + nr;
        len[0] = -ip_lo;
        return 1;
    }
    pix0[0] = ipix1 + ip_lo;
    len[0] = ip_hi - ip_lo + 1;
    return 1;
}

/* The same ring of the disc as one unwrapped index range: returns 0 (empty),
 * 1 (ring positions ip_lo..ip_hi, which may run below 0 or past nr-1 and
 * are taken modulo nr) or 2 (whole ring). */
__host__ __device__ inline int disc_ring_span(const Hpx &h, const Disc &d, int r, int &ip_lo,
                                              int &ip_hi, int &ipix1, int &nr)
{
    bool shifted;
    ring_info(h, r, ipix1, nr, shifted);
    if (r < d.irmin || r > d.irmax)
        return 2;
    double z = ring2z(h, r);
    double x = (d.cosr - z * d.z0) * d.xa;
    double ysq = 1 - z * z - x * x;
    double dphi = (ysq <= 0) ? 0.0 : pl_atan2(sqrt(ysq), x);
    if (!(dphi > 0))
        return 0;
    double shift = shifted ? 0.5 : 0.;
    double a = nr * kInvTwoPi * (d.phi - dphi) - shift;
    double b = nr * kInvTwoPi * (d.phi + dphi) - shift;
    ip_lo = ((a >= 0) ? (int)a : (int)a - 1) + 1;
    ip_hi = (b >= 0) ? (int)b : (int)b - 1;
    if (ip_lo > ip_hi)
        return 0;
    if (ip_hi - ip_lo + 1 >= nr)
        return 2;
    return 1;
}

/* disc_ring_span with the arc's half-width dphi from a float atan2, for
 * callers that need a superset (grow) or a subset (!grow) of the exact span
 * rather than the span itself.  The float dphi is within 4e-6 rad of the
 * FP64 one (float rounding of x and sqrt(ysq), <= 1.2e-7 rad; plf_atan2_core
 * with the hardware reciprocal, ~1e-6), which moves the floor arguments a, b
 * by at most m = nr/(2 pi) 4e-6 pixels.  A floor whose argument lies farther
 * than m from an integer equals the exact one; otherwise it is moved one
 * pixel outward (grow) or inward.  Everything else (the ring range, ysq and
 * the empty-arc test) is the exact path's.  Rings of more than 2^17 pixels
 * (m > 0.08) take the exact path. */
__host__ __device__ inline int disc_ring_span_cover(const Hpx &h, const Disc &d, int r, bool grow,
                                                    int &ip_lo, int &ip_hi, int &ipix1, int &nr)
{
    bool shifted;
    ring_info(h, r, ipix1, nr, shifted);
    if (nr > (1 << 17))
        return disc_ring_span(h, d, r, ip_lo, ip_hi, ipix1, nr);
    if (r < d.irmin || r > d.irmax)
        return 2;
    double z = ring2z(h, r);
    double x = (d.cosr - z * d.z0) * d.xa;
    double ysq = 1 - z * z - x * x;
    if (!(ysq > 0))
        return 0;   /* dphi = 0 on the exact path: empty */
    const float yf = (float)sqrt(ysq), xf = (float)x;
    /* The device's reciprocal estimate is within 1 ulp of the IEEE quotient
     * the host builds (the checks of the cover) use; through plf_atan2_core
     * that moves dphi by under 1e-7 rad, and the 4e-6 rad margin below
     * covers both forms. */
#if defined(__HIP_DEVICE_COMPILE__)
    const float rm = __builtin_amdgcn_rcpf(fmaxf(fabsf(xf), yf));
#else
    const float rm = 1.0f / fmaxf(fabsf(xf), yf);
#endif
    const float dphi = plf_atan2_core(yf, xf, rm);
    const double k = nr * kInvTwoPi, shift = shifted ? 0.5 : 0.;
    const double a = k * (d.phi - (double)dphi) - shift;
    const double b = k * (d.phi + (double)dphi) - shift;
    const double fa = floor(a), fb = floor(b), m = k * 4e-6;
    const double da = a - fa, db = b - fb;
    ip_lo = (int)fa + 1;
    ip_hi = (int)fb;
    if (da < m || da > 1.0 - m)
        ip_lo += grow ? -1 : 1;
    if (db < m || db > 1.0 - m)
        ip_hi += grow ? 1 : -1;
    if (ip_lo > ip_hi)
        return 0;
    if (ip_hi - ip_lo + 1 >= nr)
        return 2;
    return 1;
}

/* Pixel runs of ring r whose centres lie in the annulus dout minus din (two
 * discs about the same centre, din the smaller; inner == false: no inner
 * disc): up to two arcs, each split at the ring's first pixel, so up to 4
 * runs [p0[k], p0[k] + len[k]).  Returns the number of runs.  Both spans come
 * from disc_ring_span about the same centre, so the inner span lies inside
 * the outer one in the same unwrapped positions.  COVER: the spans come from
 * disc_ring_span_cover, the outer one grown and the inner one shrunk, so the
 * runs hold every pixel of the exact annulus (and, near a tie of a floor,
 * a pixel more); the nesting of the spans is kept. */
template <bool COVER = false>
__host__ __device__ inline int annulus_ring(const Hpx &h, const Disc &dout, const Disc &din,
                                            bool inner, int r, int p0[4], int len[4])
{
    int lo_o = 0, hi_o = -1, ip1 = 0, nr = 0;
    int kind_o;
    if constexpr (COVER)
        kind_o = disc_ring_span_cover(h, dout, r, true, lo_o, hi_o, ip1, nr);
    else
        kind_o = disc_ring_span(h, dout, r, lo_o, hi_o, ip1, nr);
    if (kind_o == 0)
        return 0;
    int kind_i = 0, lo_i = 0, hi_i = -1;
    if (inner && r >= din.r_first && r <= din.r_last) {
        int ip1b, nrb;
        if constexpr (COVER)
            kind_i = disc_ring_span_cover(h, din, r, false, lo_i, hi_i, ip1b, nrb);
        else
            kind_i = disc_ring_span(h, din, r, lo_i, hi_i, ip1b, nrb);
    }
    if (kind_i == 2)
        return 0;   /* the whole ring lies inside the inner disc */
    int ia[2], ib[2], ni = 0;
    if (kind_i == 0) {
        ia[0] = kind_o == 2 ? 0 : lo_o;
        ib[0] = kind_o == 2 ? nr - 1 : hi_o;
        ni = 1;
    } else if (kind_o == 2) {   /* the ring minus the inner span */
        ia[0] = hi_i + 1;
        ib[0] = lo_i - 1 + nr;
        ni = 1;
    } else {
        if (lo_i > lo_o) {
            ia[ni] = lo_o;
            ib[ni] = lo_i - 1;
            ni++;
        }
        if (hi_o > hi_i) {
            ia[ni] = hi_i + 1;
            ib[ni] = hi_o;
            ni++;
        }
    }
    int n = 0;
    for (int k = 0; k < ni; k++) {
        const int l = ib[k] - ia[k] + 1;   /* 1 .. nr */
        if (l <= 0)
            continue;
        int st = ia[k] % nr;
        if (st < 0)
            st += nr;
        if (st + l <= nr) {
            p0[n] = ip1 + st;
            len[n] = l;
            n++;
        } else {
            p0[n] = ip1 + st;
            len[n] = nr - st;
            n++;
            p0[n] = ip1;
            len[n] = l - (nr - st);
            n++;
        }
    }
    return n;
}

/* The annulus of sky radii that holds every visit whose beam pixel lies in
 * rows [lo, lo + nrow) of a beam map hb of support radius rho_max (cos_rmax
 * its cosine): a point in a pixel of beam ring R lies between the centres of
 * rings R-1 and R+1, so rows of rings Rf..Rl need rho in
 * [theta(Rf-1), theta(Rl+1)], widened by 1e-9 and clipped to rho_max.
 * inner == false: no inner disc (Rf is ring 1). */
__host__ __device__ inline void band_annulus_radii(const Hpx &hb, double rho_max, double cos_rmax,
                                                   int lo, int nrow, bool &inner, double &r_in,
                                                   double &c_in, double &r_out, double &c_out)
{
    const int rf = pix2ring(hb, lo), rl = pix2ring(hb, lo + nrow - 1);
    inner = rf > 1;
    r_in = inner ? ring_theta(hb, rf - 1) - 1e-9 : 0.0;
    r_out = rl + 1 <= 4 * hb.nside - 1 ? ring_theta(hb, rl + 1) + 1e-9 : rho_max;
    if (r_out >= rho_max) {
        r_out = rho_max;
        c_out = cos_rmax;
    } else {
        c_out = pl_cos(r_out);
    }
    c_in = pl_cos(r_in);
}

/* ---- the beam-frame mapping of one visit -------------------------------- */

/* Per-sample constants of the rotation (sphtrigo.cpp:29-30 evaluated once
 * per sample: they depend only on the beam centre). */
struct Centre {
    double ra, dec, psi;   /* ra, dec of the beam centre; psi = -pa */
    double cdc, sdc;
};

__host__ __device__ inline Centre centre_make(float ra, float dec, float pa)
{
    Centre c;
    c.ra = (double)ra;
    c.dec = (double)dec;
    c.psi = -(double)pa;
    pl_sincos(c.dec, &c.sdc, &c.cdc);
    return c;
}

/* rho, sigma and chi of the sky pixel at (ra_pix, cos/sin dec_pix) seen
 * from the beam centre c: rho_sigma_chi_pix (sphtrigo.cpp:27-54). */
__host__ __device__ inline void visit_rsc(const Centre &c, double ra_pix, double cdp, double sdp,
                                          double &rho, double &sigma, double &chi)
{
    double sd, cd;
    pl_sincos(ra_pix - c.ra, &sd, &cd);
    double crho = c.sdc * sdp + c.cdc * cdp * cd;
    rho = (crho >= 1.0) ? 0.0 : pl_acos(crho);
    double alpha, beta;
    if (rho > 1.0e-6) {
        beta = pl_atan2(sd * c.cdc, c.sdc * cdp - c.cdc * sdp * cd);
        alpha = pl_atan2(sd * cdp, sdp * c.cdc - cdp * c.sdc * cd);
    } else {
        beta = kHalfPi;
        alpha = kHalfPi;
    }
    sigma = kPi - alpha - c.psi;
    if (sigma < 0.0)
        sigma += 2.0 * kPi;
    else if (sigma > 2.0 * kPi)
        sigma -= 2.0 * kPi;
    chi = beta - sigma;
    if (chi < -kPi)
        chi += 2.0 * kPi;
}

/* Beam pixel (ang2pix of (rho, sigma) on the beam map) and chi for the sky
 * pixel at (ra_pix, cos/sin dec_pix): rho_sigma_chi_pix followed by
 * Healpix_Base::ang2pix (convolver.cpp:148-149). */
__host__ __device__ inline int beam_pixel(const Hpx &hb, const Centre &c, double ra_pix,
                                 double cdp, double sdp, double &chi)
{
    double rho, sigma;
    visit_rsc(c, ra_pix, cdp, sdp, rho, sigma, chi);
    return ang2pix(hb, rho, sigma);
}

/* ---- single-precision form of the same mapping, certified ---------------
 *
 * The kernels evaluate the visit geometry in float32 and keep the beam pixel
 * it gives only where no floor of ang2pix can flip; any other visit is
 * re-evaluated by the exact FP64 path (beam_pixel).  Either way the pixel is
 * the one the CPU path computes.
 *
 * Geometry in haversine form (no cancellation).  With dD = dec_p - dec_c and
 * h2 = 1 - cos(dra) = 2 sin^2(dra/2):
 *   1 - cos rho = (1 - cos dD) + cos dec_c cos dec_p h2,
 *   1 - cos dD  = sin^2 dD / (1 + cos dD),
 *   ya = sin dra cos dec_p,  xa =  sin dD + cos dec_p sin dec_c h2,
 *   yb = sin dra cos dec_c,  xb = -sin dD + cos dec_c sin dec_p h2,
 * alpha = atan2(ya, xa), beta = atan2(yb, xb) (sphtrigo.cpp:41-42).  sin dD
 * and cos dD come from FP64 products of the ring's and the sample's sin/cos
 * dec; everything after that is float.  In the polar branch of ang2pix
 * (cos rho > 2/3) the floors are jp = (int)(tp*tmp), jm = (int)((1-tp)*tmp)
 * and ip = (int)(tt*(jp+jm+1)), with tt = (2 - 2 alpha/pi - 2 psi/pi) mod 4,
 * tp = frac(tt), tmp = nside*sqrt(3 (1 - cos rho)).
 *
 * Error of the float values against the CPU path's (float32 emulation over
 * 4e6 random visits, discs of 5 degrees, a quarter of them near the poles):
 *   tt   <= 6.0e-7 absolute              bound used: 4e-6
 *   tmp  <= 2.6e-7 relative              bound used: 2e-6
 * plus the CPU path's own rounding of acos(crho) (|crho error| <= 7e-16), which
 * moves its tmp by <= 1.05e-15 nside / sqrt(3 (1 - cos rho)); bound used: 100
 * times that (marg_k).  A floor is kept only if its argument is farther from
 * an integer than these bounds.  rho > 1e-6 (sphtrigo.cpp:40) is decided in
 * float outside 0.5e-6 < rho < 1.5e-6 and by the exact path inside.
 *
 * chi is needed as float cos(2 chi), sin(2 chi): with chi = beta + alpha + psi
 * - pi (mod 2 pi), e^{2 i chi} = w^2/|w|^2 e^{2 i psi}, w = (xb + i yb)(xa +
 * i ya), evaluated in float (error <= 9.2e-7 measured). */
struct FastConsts {
    float omc_lo;   /* 1 - cos(0.5e-6): below, rho <= 1e-6 for sure */
    float omc_hi;   /* 1 - cos(1.5e-6): above, rho >  1e-6 for sure */
    float marg_k;   /* 1.2e-13 * nside_beam */
    float nside;    /* nside_beam */
};

/* Wave-uniform float constants of one sample. */
struct SampleF {
    float cdc, sdc;   /* cos, sin dec of the beam centre */
    float t0;         /* (pi - psi) * 2/pi mod 4 */
    float c2p, s2p;   /* cos, sin 2 psi */
};

__host__ __device__ inline SampleF sample_f32(const Centre &c, double c2psi, double s2psi)
{
    SampleF f;
    f.cdc = (float)c.cdc;
    f.sdc = (float)c.sdc;
    f.t0 = (float)fmodulo((kPi - c.psi) * kInvHalfPi, 4.0);
    f.c2p = (float)c2psi;
    f.s2p = (float)s2psi;
    return f;
}

constexpr int kNeedExact = -2;   /* beam_pixel_f32: the floors are not certain */

struct ExactVisit {
    int bp;
    float c2, s2;
};

/* The exact FP64 mapping (beam_pixel) of one visit, with (float)cos/sin 2chi
 * as the CPU path forms them.  Out of line, scalar arguments only (a
 * reference argument would force the caller's parameters into scratch). */
__device__ __noinline__ ExactVisit beam_visit_exact(int nside_beam, double ra_c, double psi,
                                                    double cdc, double sdc, double ra_pix,
                                                    double cdp, double sdp)
{
    Centre c;
    c.ra = ra_c;
    c.dec = 0.0;   /* not read by beam_pixel */
    c.psi = psi;
    c.cdc = cdc;
    c.sdc = sdc;
    double chi, s, co;
    ExactVisit e;
    e.bp = beam_pixel(hpx_make(nside_beam), c, ra_pix, cdp, sdp, chi);
    pl_sincos(2.0 * chi, &s, &co);
    e.c2 = (float)co;
    e.s2 = (float)s;
    return e;
}

/* The part of a visit's float geometry that reads the sky pixel's ring
 * (FP64 in, float out); the rest (beam_pixel_f32_tail) is float only. */
struct VisitRing {
    float xh;         /* dra / 2, dra = ra_pix - ra_centre in [-pi, pi] */
    float sD, opcD;   /* sin(dec_p - dec_c), 1 + cos(dec_p - dec_c) */
    float cdp, sdp;   /* cos, sin dec_p */
};

__device__ inline VisitRing visit_ring_f32(const Centre &c, double ra_pix, double cdp, double sdp)
{
    double dra = ra_pix - c.ra;
    dra = fma(-kTwoPi, rint(dra * kInvTwoPi), dra);
    VisitRing g;
    g.xh = (float)(0.5 * dra);
    g.sD = (float)fma(sdp, c.cdc, -(cdp * c.sdc));
    g.opcD = 1.0f + (float)fma(cdp, c.cdc, sdp * c.sdc);
    g.cdp = (float)cdp;
    g.sdp = (float)sdp;
    return g;
}

/* The part of the float mapping that does not depend on the detector's
 * position angle: 1 - cos rho, alpha, and e^{2i(alpha+beta)} (er, ei). */
struct VisitGeo {
    float omc;        /* 1 - cos rho */
    float x3y;        /* sqrt(3 (1 - cos rho)) */
    float y;          /* 1/sqrt(3 (1 - cos rho)) */
    float alpha;
    float er, ei;     /* e^{2i(alpha+beta)} */
    bool exact;       /* rho within 0.5e-6 of 1e-6, or degenerate: exact path */
};

__device__ inline VisitGeo visit_geo_f32(const FastConsts &fc, const SampleF &sf,
                                         const VisitRing &g)
{
    VisitGeo v;
    float sh, ch;
    plf_sincos(g.xh, &sh, &ch);
    const float sdra = 2.0f * (sh * ch);
    const float h2 = 2.0f * (sh * sh);
    const float omD = (g.sD * g.sD) * __builtin_amdgcn_rcpf(g.opcD);
    v.omc = fmaf(h2, sf.cdc * g.cdp, omD);
    const float ya = sdra * g.cdp, xa = fmaf(h2, g.cdp * sf.sdc, g.sD);
    const float yb = sdra * sf.cdc, xb = fmaf(h2, sf.cdc * g.sdp, -g.sD);
    const bool big = v.omc > fc.omc_hi;
    v.exact = !big && !(v.omc < fc.omc_lo);   /* rho within 0.5e-6 of 1e-6 */
    v.alpha = 0x1.921fb6p+0f;                 /* pi/2 when rho <= 1e-6 */
    v.er = 1.0f;
    v.ei = 0.0f;
    if (big) {
        v.alpha = plf_atan2_core(ya, xa, __builtin_amdgcn_rcpf(fmaxf(fabsf(xa), fabsf(ya))));
        const float wr = xb * xa - yb * ya, wi = xb * ya + yb * xa;
        const float n = wr * wr + wi * wi;
        if (!(n > 1e-36f))
            v.exact = true;   /* |w|^2 >= 5e-24 for rho > 1.5e-6 */
        const float inv = __builtin_amdgcn_rcpf(n);
        v.er = (wr * wr - wi * wi) * inv;
        v.ei = 2.0f * (wr * wi) * inv;
    }
    const float x3 = 3.0f * v.omc;
    v.y = __builtin_amdgcn_rsqf(x3);
    v.x3y = x3 * v.y;
    return v;
}

/* The detector's part: its beam pixel (beam map of NSIDE nside with margin
 * factor marg_k) and float cos/sin 2chi, given t0 = (pi - psi) 2/pi mod 4 and
 * cos/sin 2psi of its position angle; kNeedExact where a floor is not
 * certain. */
__device__ inline int beam_pixel_f32_det(const VisitGeo &v, float t0, float c2p, float s2p,
                                         float nside, float marg_k, float &c2, float &s2)
{
    c2 = v.er * c2p - v.ei * s2p;
    s2 = v.er * s2p + v.ei * c2p;
    float tt = fmaf(-v.alpha, 0x1.45f306p-1f /* 2/pi */, t0);
    tt = (tt < 0.0f) ? tt + 4.0f : tt;
    tt = (tt >= 4.0f) ? tt - 4.0f : tt;
    const float tmp = nside * v.x3y;
    const float tp = tt - (float)(int)tt;
    const float u1 = tp * tmp, u2 = (1.0f - tp) * tmp;
    const int jp = (int)u1, jm = (int)u2;
    const float d1 = u1 - (float)jp, d2 = u2 - (float)jm;
    const float md = fmaf(tmp, 8e-6f, fmaf(marg_k, v.y, 1e-6f));
    const int ir = jp + jm + 1;
    const float vv = tt * (float)ir;
    const int ip = (int)vv;
    const float dv = vv - (float)ip;
    const float mv = fmaf((float)ir, 4.5e-6f, 1e-6f);
    const bool ok = !v.exact && (v.omc > 0.0f) && (v.omc < 0.33332f) && (d1 > md) &&
                    (d1 < 1.0f - md) && (d2 > md) && (d2 < 1.0f - md) && (dv > mv) &&
                    (dv < 1.0f - mv);
    return ok ? 2 * ir * (ir - 1) + ip : kNeedExact;
}

/* Beam pixel and float cos/sin 2chi of a visit, or kNeedExact (c2, s2 then
 * unset). */
__device__ inline int beam_pixel_f32_tail(const FastConsts &fc, const SampleF &sf,
                                          const VisitRing &g, float &c2, float &s2)
{
    const VisitGeo v = visit_geo_f32(fc, sf, g);
    if (v.exact)
        return kNeedExact;
    return beam_pixel_f32_det(v, sf.t0, sf.c2p, sf.s2p, fc.nside, fc.marg_k, c2, s2);
}

/* Beam pixel and float cos/sin 2chi of the visit of the sky pixel at
 * (ra_pix, cos/sin dec_pix), or kNeedExact (c2, s2 then unset). */
__device__ inline int beam_pixel_f32(const FastConsts &fc, const Centre &c, const SampleF &sf,
                                     double ra_pix, double cdp, double sdp, float &c2,
                                     float &s2)
{
    return beam_pixel_f32_tail(fc, sf, visit_ring_f32(c, ra_pix, cdp, sdp), c2, s2);
}

/* ---- certified double-precision form of the same mapping -----------------
 *
 * A second certified form, used by the forward kernel: on the 5-degree,
 * NSIDE 512 workload the forward measured faster with it than with the
 * single-precision form (its beam gathers saw a 77% instead of 54% L2 hit
 * rate).  Its margins are ~1e-12 relative, so it practically never needs the
 * exact path, and it evaluates that path in place instead of deferring. */

/* The CPU path obtains the beam pixel from rho = acos(crho) through
 * z = cos(rho) (and sin(rho) within 0.01 rad of the pole); in the polar branch
 * of ang2pix the two floors that depend on rho are jp = (int)(tp*tmp) and
 * jm = (int)((1-tp)*tmp) with tmp = nside*sqrt(3*(1-z)) (or, within 0.01 rad,
 * nside*sin(rho)/sqrt((1+z)/3)).  Both forms equal nside*sqrt(3*(1-crho))
 * up to the rounding of acos/cos/sin, which is bounded by
 *     |tmp_exact - tmp_fast| <= nside*4.4e-16/sqrt(1-crho) + 2e-15*tmp
 * (pisco_libm's acos, cos and sin are within 2 ulp).  beam_pixel_f64
 * evaluates the cheap form and accepts its floors only if both fractional
 * parts are farther than 100 times that bound (plus 1e-10) from an integer;
 * otherwise it evaluates the exact path.  Either way the returned pixel is
 * the one beam_pixel returns.  sigma (hence tt, tp and ip) is computed with
 * the same operations as the exact path.
 *
 * chi is needed only as float cos(2 chi), sin(2 chi).  With
 * beta = atan2(yb, xb), alpha = atan2(ya, xa) and chi = beta + alpha + psi - pi
 * (mod 2 pi; sphtrigo.cpp:41-53), e^{2 i chi} = w^2/|w|^2 * e^{2 i psi} with
 * w = (xb + i yb)(xa + i ya): no transcendental is needed (FP64 algebra,
 * within ~1e-15 of cos/sin(2 chi) before rounding to float). */
struct FastConsts64 {
    double crho_big;    /* crho <  crho_big  => rho > 1e-6 for sure */
    double crho_small;  /* crho >= crho_small => rho <= 1e-6 for sure */
    double marg_k;      /* 100 * 4.4e-16 * nside * sqrt(3) */
};

/* Rare paths of beam_pixel_f64, kept out of line so that they do not
 * occupy registers in the hot loop. */
/* (NSIDE by value: a reference argument of an out-of-line call would force
 * the caller's parameter block into addressable scratch memory.) */
__device__ __noinline__ int beam_pixel_exact_rho(int nside_beam, double crho, double sigma)
{
    double rho = (crho >= 1.0) ? 0.0 : pl_acos(crho);
    return ang2pix(hpx_make(nside_beam), rho, sigma);
}

__device__ __noinline__ bool rho_above_1e6(double crho)
{
    return (crho >= 1.0) ? false : (pl_acos(crho) > 1.0e-6);
}

__device__ __noinline__ double2 e2i_angles(double yb, double xb, double alpha)
{
    double beta = pl_atan2(yb, xb);
    double s2c, c2c;
    pl_sincos(2.0 * (beta + alpha), &s2c, &c2c);
    return make_double2(c2c, s2c);
}

/* pl_sincos for a wavefront: when no lane needs argument reduction the
 * reduction step is skipped for the whole wave.  For |x| < 0.78 the
 * reduction computes n = rint(2x/pi) = 0 and returns y = x, ylo = +0 exactly,
 * so the results are the same bits either way. */
__device__ inline void wave_sincos(double x, double *s, double *c)
{
    if (__any(!(fabs(x) < 0.78))) {
        pl_sincos(x, s, c);
        return;
    }
    *s = pl_sin_kernel(x, 0.0);
    *c = pl_cos_kernel(x, 0.0);
}

/* 1/n to ~1e-15 relative: hardware reciprocal estimate + one Newton step
 * (used only where the value feeds float outputs, never an index). */
__device__ inline double recip_fast(double n)
{
    double r = __builtin_amdgcn_rcp(n);
    return fma(r, fma(-n, r, 1.0), r);
}

/* Bound on |alpha_f - alpha| for alpha_f = plf_atan2_core of the FP64
 * arguments rounded to float (float32 emulation over 4e6 arguments of every
 * direction and magnitude, the reciprocal estimate within +-1 ulp: at most
 * 3.6e-7 rad), with a factor of four of margin. */
constexpr double kAlphaF32Err = 1.5e-6;

/* sigma -> the polar-branch floors of loc2pix (beam_pixel_f64's tail).
 * et bounds the error of tt = sigma 2/pi (mod 4) coming from an approximate
 * alpha (0 for the exact alpha): every floor that depends on tt must then be
 * farther than its error from an integer.  An uncertain floor gives
 * kNeedExact when DEFER, else the exact path's pixel. */
template <bool DEFER>
__device__ inline int beam_floors(const Hpx &hb, double crho, double sigma, double tmp, double y,
                                  double marg_k, double et)
{
    double tt = fmodulo(sigma * kInvHalfPi, 4.0);
    double tp = tt - (int)tt;
    double u1 = tp * tmp;
    double u2 = (1.0 - tp) * tmp;
    int jp = (int)u1, jm = (int)u2;
    double d1 = u1 - jp, d2 = u2 - jm;
    /* margin: nside * 4.4e-14 / sqrt(omc) + 3e-13 tmp + 1e-10, with
     * 1/sqrt(omc) = sqrt(3) * y; plus et * tmp for the floors of tp * tmp */
    double md = fma(marg_k, y, fma(fma(et, 1.0, 3e-13), tmp, 1e-10));
    bool ok = (crho > kTwoThird + 1e-12) && (d1 > md) && (d1 < 1.0 - md) && (d2 > md) &&
              (d2 < 1.0 - md);
    int ir = jp + jm + 1;
    double vv = tt * ir;
    int ip = (int)vv;
    if (et > 0.0) {
        /* the quadrant (int)tt and ip = (int)(tt ir) */
        double mv = fma(et, (double)ir, 1e-12);
        double dv = vv - ip;
        ok = ok && (tp > et) && (tp < 1.0 - et) && (dv > mv) && (dv < 1.0 - mv);
    }
    if (ok)
        return 2 * ir * (ir - 1) + ip;
    if constexpr (DEFER)
        return kNeedExact;
    else
        return beam_pixel_exact_rho(hb.nside, crho, sigma);
}

__device__ inline double sigma_of(double alpha, double psi)
{
    double sigma = kPi - alpha - psi;
    if (sigma < 0.0)
        sigma += 2.0 * kPi;
    else if (sigma > 2.0 * kPi)
        sigma -= 2.0 * kPi;
    return sigma;
}

/* FLOAT_ALPHA: alpha from the float atan2 of the FP64 arguments, its floors
 * certified against kAlphaF32Err, and kNeedExact for a visit whose floors
 * cannot be certified (the caller defers it to this function's exact form);
 * otherwise alpha from the FP64 atan2, the CPU path's own (sphtrigo.cpp:42),
 * and the exact path in place for an uncertain floor. */
template <bool FLOAT_ALPHA = false>
__device__ inline int beam_pixel_f64(const Hpx &hb, const FastConsts64 &fc, const Centre &c,
                                      double c2psi, double s2psi, double ra_pix,
                                      double cdp, double sdp, float &c2, float &s2)
{
    double sd, cd;
    wave_sincos(ra_pix - c.ra, &sd, &cd);
    double crho = c.sdc * sdp + c.cdc * cdp * cd;
    bool big;   /* rho > 1e-6 (sphtrigo.cpp:40) */
    if (crho < fc.crho_big)
        big = true;
    else if (crho >= fc.crho_small)
        big = false;
    else
        big = rho_above_1e6(crho);
    double yb = sd * c.cdc, xb = c.sdc * cdp - c.cdc * sdp * cd;
    double ya = sd * cdp, xa = sdp * c.cdc - cdp * c.sdc * cd;

    /* cos/sin 2chi */
    double er = 1.0, ei = 0.0;   /* e^{2i(alpha+beta)}; alpha = beta = pi/2 -> 1 */
    if (big) {
        double wr = xb * xa - yb * ya, wi = xb * ya + yb * xa;
        double n = wr * wr + wi * wi;
        if (n > 1e-280) {
            double inv = recip_fast(n);
            er = (wr * wr - wi * wi) * inv;
            ei = 2.0 * wr * wi * inv;
        } else {   /* degenerate geometry: take the angles themselves */
            double2 e = e2i_angles(yb, xb, pl_atan2(ya, xa));
            er = e.x;
            ei = e.y;
        }
    }
    c2 = (float)(er * c2psi - ei * s2psi);
    s2 = (float)(er * s2psi + ei * c2psi);

    /* beam pixel: polar branch of loc2pix with the certified floors.
     * sqrt(3 omc) from the hardware rsqrt estimate and two Newton steps:
     * within 1e-15 relative, an error the margin includes. */
    double omc = 1.0 - crho;
    omc = omc > 0.0 ? omc : 0.0;
    double x3 = 3.0 * omc;
    double y = __builtin_amdgcn_rsq(x3);
    y = y * fma(-0.5 * x3, y * y, 1.5);
    y = y * fma(-0.5 * x3, y * y, 1.5);
    double tmp = (x3 > 0.0) ? hb.nside * (x3 * y) : 0.0;
    if constexpr (FLOAT_ALPHA) {
        double alpha = kHalfPi;
        if (big) {
            const float yf = (float)ya, xf = (float)xa;
            alpha = (double)plf_atan2_core(
                yf, xf, __builtin_amdgcn_rcpf(fmaxf(fabsf(xf), fabsf(yf))));
        }
        return beam_floors<true>(hb, crho, sigma_of(alpha, c.psi), tmp, y, fc.marg_k,
                                 big ? kAlphaF32Err * kInvHalfPi : 0.0);
    } else {
        double alpha = big ? pl_atan2(ya, xa) : kHalfPi;
        return beam_floors<false>(hb, crho, sigma_of(alpha, c.psi), tmp, y, fc.marg_k, 0.0);
    }
}

/* get_interpol (Healpix_Base::get_interpol; the device port in the reference
 * is cudaHealpix.cu:198-259). */
__host__ __device__ inline void ring_info2(const Hpx &h, int ring, int &startpix,
                                  int &ringpix, double &theta, bool &shifted)
{
    int northring = (ring > 2 * h.nside) ? 4 * h.nside - ring : ring;
    if (northring < h.nside) {
        double tmp = northring * northring * h.fact2;
        double costheta = 1 - tmp;
        double sintheta = sqrt(tmp * (2 - tmp));
        theta = pl_atan2(sintheta, costheta);
        ringpix = 4 * northring;
        shifted = true;
        startpix = 2 * northring * (northring - 1);
    } else {
        theta = pl_acos((2 * h.nside - northring) * h.fact1);
        ringpix = 4 * h.nside;
        shifted = ((northring - h.nside) & 1) == 0;
        startpix = h.ncap + (northring - h.nside) * ringpix;
    }
    if (northring != ring) {
        theta = kPi - theta;
        startpix = h.npix - startpix - ringpix;
    }
}

__host__ __device__ inline void get_interpol(const Hpx &h, double theta, double phi,
                                    int pix[4], double wgt[4])
{
    double z = pl_cos(theta);
    int ir1 = ring_above(h, z);
    int ir2 = ir1 + 1;
    double theta1 = 0, theta2 = 0, w1, tmp, dphi;
    int sp, nr, i1, i2;
    bool shift;
    if (ir1 > 0) {
        ring_info2(h, ir1, sp, nr, theta1, shift);
        dphi = kTwoPi / nr;
        tmp = (phi / dphi - .5 * shift);
        i1 = (tmp < 0) ? (int)tmp - 1 : (int)tmp;
        w1 = (phi - (i1 + .5 * shift) * dphi) / dphi;
        i2 = i1 + 1;
        if (i1 < 0) i1 += nr;
        if (i2 >= nr) i2 -= nr;
        pix[0] = sp + i1;
        pix[1] = sp + i2;
        wgt[0] = 1 - w1;
        wgt[1] = w1;
    }
    if (ir2 < (4 * h.nside)) {
        ring_info2(h, ir2, sp, nr, theta2, shift);
        dphi = kTwoPi / nr;
        tmp = (phi / dphi - .5 * shift);
        i1 = (tmp < 0) ? (int)tmp - 1 : (int)tmp;
        w1 = (phi - (i1 + .5 * shift) * dphi) / dphi;
        i2 = i1 + 1;
        if (i1 < 0) i1 += nr;
        if (i2 >= nr) i2 -= nr;
        pix[2] = sp + i1;
        pix[3] = sp + i2;
        wgt[2] = 1 - w1;
        wgt[3] = w1;
    }
    if (ir1 == 0) {
        double wtheta = theta / theta2;
        wgt[2] *= wtheta;
        wgt[3] *= wtheta;
        double fac = (1 - wtheta) * 0.25;
        wgt[0] = fac;
        wgt[1] = fac;
        wgt[2] += fac;
        wgt[3] += fac;
        pix[0] = (pix[2] + 2) & 3;
        pix[1] = (pix[3] + 2) & 3;
    } else if (ir2 == 4 * h.nside) {
        double wtheta = (theta - theta1) / (kPi - theta1);
        wgt[0] *= (1 - wtheta);
        wgt[1] *= (1 - wtheta);
        double fac = wtheta * 0.25;
        wgt[0] += fac;
        wgt[1] += fac;
        wgt[2] = fac;
        wgt[3] = fac;
        pix[2] = ((pix[0] + 2) & 3) + h.npix - 4;
        pix[3] = ((pix[1] + 2) & 3) + h.npix - 4;
    } else {
        double wtheta = (theta - theta1) / (theta2 - theta1);
        wgt[0] *= (1 - wtheta);
        wgt[1] *= (1 - wtheta);
        wgt[2] *= wtheta;
        wgt[3] *= wtheta;
    }
}

} /* namespace pisco */

#endif /* PISCO_DEVICE_HPX_H */
