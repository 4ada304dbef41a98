/*
 * pisco_oracle.c -- CPU oracle for the pixel-space beam convolution.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the test suite, the
 * smoke check and the CPU-baseline leg of bench.py compare the HIP kernels
 * against.  Nothing in the shipped library links or calls it.
 *
 * It restates, in plain C, the reference algorithm of PISCO's CPU path:
 *   - Convolver::exec_convolution / beam_times_sky
 *       (src/Convolver/convolver.cpp:22-72, :74-183)
 *   - SphericalTransformations::rho_sigma_chi_pix (src/Sphtrigo/sphtrigo.cpp:4-55)
 *   - PolBeam ctor / beam_from_fields / build / make_unpol_gaussian_elliptical_beams
 *       (src/Polbeam/polbeam.cpp:13-38, :117-184, :186-288, :290-372)
 *   - the HEALPix RING routines the reference takes from healpix_cxx
 *     (Healpix_Base::pix2ang, ang2pix, query_disc, get_interpol; third party,
 *     not vendored, version unpinned in cmake_modules/FindHealpix.cmake:19-31).
 *     They are restated from the published healpix_cxx 3.x algorithm
 *     (Gorski et al. 2005, ApJ 622, 759; healpix_base.cc): pix2loc/loc2pix with
 *     the theta<0.01 "sth" rule and a non-negative fmodulo, ring_above,
 *     ring2z, get_ring_info_small/get_ring_info2, the non-inclusive RING
 *     query_disc and the bilinear get_interpol
 *     (the device port src/GpuConvolver/cuda/cudaHealpix.cu:30-259 follows the
 *     same healpix_base.cc routines).
 * plus the adjoints the build adds (TOD -> I,Q,U map; TOD -> beam maps), in
 * double precision, a full-FP64 forward used to state the error of each
 * result against an FP64 oracle, and the V-term mode of the reference's GPU
 * kernel (V A3, V B3; cudaConv2.cu:204,214,217-218), off by default.
 *
 * Transcendentals: by default sin/cos/acos/atan2 come from pisco_libm.h, the
 * FP64 library the GPU kernels use, so that pixel indices are comparable bit
 * for bit.  Built with -DORACLE_GLIBC they come from the system libm instead;
 * the test suite compares the two builds (index mismatch rate) as a
 * diagnostic of how far the shared library moves results from glibc.
 *
 * How each part is pinned (tests/ under the CPU marker):
 *   pix2ang / ang2pix  round trip of every pixel (NSIDE <= 256, sampled 4096
 *                      and 8192), closed-form ring z and phi, published NSIDE 1
 *                      values, equal-area sampling by brute force
 *   query_disc         brute force over all pixels (NSIDE <= 64), area law
 *   get_interpol       weights sum to one, non-negative, exact at pixel centres
 *   rho_sigma_chi      bit-exact against golden vectors written by the
 *                      reference's own sphtrigo.cpp (tests/golden), an
 *                      independent 3-vector construction, on-axis closed forms
 *   Convolver forward  point source traces the beam, uniform sky sums the
 *                      beam, linearity, a/b antisymmetry, 90-degree rotation
 *   PolBeam            closed form of the elliptical Gaussian
 *   adjoints           dot-product identity against the forward
 * The TOD sums and query_disc have no golden values in the reference; beyond
 * these properties their bits are not pinned against the original build.
 * Compile with -ffp-contract=off (see Makefile).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_GLIBC
#define O_SIN sin
#define O_COS cos
#define O_ACOS acos
#define O_ASIN asin
#define O_ATAN2 atan2
#else
#include "../pisco4cmb_amd/csrc/pisco_libm.h"
#define O_SIN pl_sin
#define O_COS pl_cos
#define O_ACOS pl_acos
#define O_ASIN pl_asin
#define O_ATAN2 pl_atan2
#endif

static const double O_PI = 3.141592653589793238462643383279502884197;
static const double O_TWOPI = 6.283185307179586476925286766559005768394;
static const double O_HALFPI = 1.570796326794896619231321691639751442099;
static const double O_INV_HALFPI = 0.6366197723675813430755350534900574;
static const double O_INV_TWOPI = 1.0 / 6.283185307179586476925286766559005768394;
static const double O_TWOTHIRD = 2.0 / 3.0;

/* ------------------------------------------------------------------ */
/* HEALPix RING, healpix_cxx semantics                                 */
/* ------------------------------------------------------------------ */

typedef struct {
    int nside;
    int npface;
    int ncap;
    int npix;
    double fact1, fact2;
} hpx_t;

static hpx_t hpx_make(int nside)
{
    hpx_t h;
    h.nside = nside;
    h.npface = nside * nside;
    h.ncap = (h.npface - nside) << 1;
    h.npix = 12 * h.npface;
    h.fact2 = 4. / h.npix;
    h.fact1 = (nside << 1) * h.fact2;
    return h;
}

/* healpix_cxx fmodulo: result in [0, v2) */
static double o_fmodulo(double v1, double v2)
{
    if (v1 >= 0)
        return (v1 < v2) ? v1 : fmod(v1, v2);
    double tmp = fmod(v1, v2) + v2;
    return (tmp == v2) ? 0. : tmp;
}

static int o_isqrt(int v)
{
    /* integer square root, exact for v < 2^31 */
    int r = (int)sqrt((double)v + 0.5);
    return r;
}

static void hpx_pix2loc(const hpx_t *h, int pix, double *z, double *phi,
                        double *sth, int *have_sth)
{
    *have_sth = 0;
    *sth = 0.0;
    if (pix < h->ncap) {                    /* north polar cap */
        int iring = (1 + o_isqrt(1 + 2 * pix)) >> 1;
        int iphi = (pix + 1) - 2 * iring * (iring - 1);
        double tmp = (iring * iring) * h->fact2;
        *z = 1.0 - tmp;
        if (*z > 0.99) {
            *sth = sqrt(tmp * (2.0 - tmp));
            *have_sth = 1;
        }
        *phi = (iphi - 0.5) * O_HALFPI / iring;
    } else if (pix < (h->npix - h->ncap)) { /* equatorial region */
        int nl4 = 4 * h->nside;
        int ip = pix - h->ncap;
        int tmp = ip / nl4;
        int iring = tmp + h->nside;
        int iphi = ip - nl4 * tmp + 1;
        double fodd = ((iring + h->nside) & 1) ? 1 : 0.5;
        *z = (2 * h->nside - iring) * h->fact1;
        *phi = (iphi - fodd) * O_PI * 0.75 * h->fact1;
    } else {                                /* south polar cap */
        int ip = h->npix - pix;
        int iring = (1 + o_isqrt(2 * ip - 1)) >> 1;
        int iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
        double tmp = (iring * iring) * h->fact2;
        *z = tmp - 1.0;
        if (*z < -0.99) {
            *sth = sqrt(tmp * (2.0 - tmp));
            *have_sth = 1;
        }
        *phi = (iphi - 0.5) * O_HALFPI / iring;
    }
}

static void hpx_pix2ang(const hpx_t *h, int pix, double *theta, double *phi)
{
    double z, sth;
    int have_sth;
    hpx_pix2loc(h, pix, &z, phi, &sth, &have_sth);
    *theta = have_sth ? O_ATAN2(sth, z) : O_ACOS(z);
}

static int hpx_loc2pix(const hpx_t *h, double z, double phi, double sth,
                       int have_sth)
{
    double za = fabs(z);
    double tt = o_fmodulo(phi * O_INV_HALFPI, 4.0);
    int nside = h->nside;
    if (za <= O_TWOTHIRD) {
        int nl4 = 4 * nside;
        double temp1 = nside * (0.5 + tt);
        double temp2 = nside * z * 0.75;
        int jp = (int)(temp1 - temp2);
        int jm = (int)(temp1 + temp2);
        int ir = nside + 1 + jp - jm;
        int kshift = 1 - (ir & 1);
        int t1 = jp + jm - nside + kshift + 1 + nl4 + nl4;
        int ip = (t1 >> 1) % nl4;
        return h->ncap + (ir - 1) * nl4 + ip;
    } else {
        double tp = tt - (int)tt;
        double tmp = ((za < 0.99) || (!have_sth))
                         ? nside * sqrt(3 * (1 - za))
                         : nside * sth / sqrt((1. + za) / 3.);
        int jp = (int)(tp * tmp);
        int jm = (int)((1.0 - tp) * tmp);
        int ir = jp + jm + 1;
        int ip = (int)(tt * ir);
        return (z > 0) ? 2 * ir * (ir - 1) + ip
                       : h->npix - 2 * ir * (ir + 1) + ip;
    }
}

static int hpx_ang2pix(const hpx_t *h, double theta, double phi)
{
    if ((theta < 0.01) || (theta > 3.14159 - 0.01))
        return hpx_loc2pix(h, O_COS(theta), phi, O_SIN(theta), 1);
    return hpx_loc2pix(h, O_COS(theta), phi, 0., 0);
}

static int hpx_ring_above(const hpx_t *h, double z)
{
    double az = fabs(z);
    if (az <= O_TWOTHIRD)
        return (int)(h->nside * (2 - 1.5 * z));
    int iring = (int)(h->nside * sqrt(3 * (1 - az)));
    return (z > 0) ? iring : 4 * h->nside - iring - 1;
}

static double hpx_ring2z(const hpx_t *h, int ring)
{
    if (ring < h->nside)
        return 1 - ring * ring * h->fact2;
    if (ring <= 3 * h->nside)
        return (2 * h->nside - ring) * h->fact1;
    ring = 4 * h->nside - ring;
    return ring * ring * h->fact2 - 1;
}

static void hpx_ring_info_small(const hpx_t *h, int ring, int *startpix,
                                int *ringpix, int *shifted)
{
    if (ring < h->nside) {
        *shifted = 1;
        *ringpix = 4 * ring;
        *startpix = 2 * ring * (ring - 1);
    } else if (ring < 3 * h->nside) {
        *shifted = ((ring - h->nside) & 1) == 0;
        *ringpix = 4 * h->nside;
        *startpix = h->ncap + (ring - h->nside) * (*ringpix);
    } else {
        *shifted = 1;
        int nr = 4 * h->nside - ring;
        *ringpix = 4 * nr;
        *startpix = h->npix - 2 * nr * (nr + 1);
    }
}

static void hpx_ring_info2(const hpx_t *h, int ring, int *startpix,
                           int *ringpix, double *theta, int *shifted)
{
    int northring = (ring > 2 * h->nside) ? 4 * h->nside - ring : ring;
    if (northring < h->nside) {
        double tmp = northring * northring * h->fact2;
        double costheta = 1 - tmp;
        double sintheta = sqrt(tmp * (2 - tmp));
        *theta = O_ATAN2(sintheta, costheta);
        *ringpix = 4 * northring;
        *shifted = 1;
        *startpix = 2 * northring * (northring - 1);
    } else {
        *theta = O_ACOS((2 * h->nside - northring) * h->fact1);
        *ringpix = 4 * h->nside;
        *shifted = ((northring - h->nside) & 1) == 0;
        *startpix = h->ncap + (northring - h->nside) * (*ringpix);
    }
    if (northring != ring) {
        *theta = O_PI - *theta;
        *startpix = h->npix - *startpix - *ringpix;
    }
}

/* Append [a, b) to a sorted range list, merging adjacent ranges. */
typedef struct {
    int *r;       /* 2*n entries: begin,end */
    long n;
    long cap;
} rset_t;

static void rset_append(rset_t *s, int a, int b)
{
    if (a >= b)
        return;
    if (s->n > 0 && s->r[2 * s->n - 1] == a) {
        s->r[2 * s->n - 1] = b;
        return;
    }
    if (s->n == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 64;
        s->r = (int *)realloc(s->r, sizeof(int) * 2 * s->cap);
    }
    s->r[2 * s->n] = a;
    s->r[2 * s->n + 1] = b;
    s->n++;
}

/* pointing::normalize (theta into [0,pi], phi into [0,2pi)) */
static void o_normalize(double *theta, double *phi)
{
    *theta = o_fmodulo(*theta, O_TWOPI);
    if (*theta > O_PI) {
        *phi += O_PI;
        *theta = O_TWOPI - *theta;
    }
    *phi = o_fmodulo(*phi, O_TWOPI);
}

static int o_ifloor(double x)
{
    return (x >= 0) ? (int)x : (int)x - 1;
}

/* Non-inclusive RING query_disc (Healpix_Base::query_disc, fact = 0). */
static void hpx_query_disc(const hpx_t *h, double theta, double phi,
                           double radius, rset_t *out)
{
    out->n = 0;
    o_normalize(&theta, &phi);
    double rsmall = radius, rbig = radius;
    if (rsmall >= O_PI) {
        rset_append(out, 0, h->npix);
        return;
    }
    rbig = (rbig < O_PI) ? rbig : O_PI;
    double cosrbig = O_COS(rbig);
    double z0 = O_COS(theta);
    double xa = 1. / sqrt((1 - z0) * (1 + z0));
    double rlat1 = theta - rsmall;
    double zmax = O_COS(rlat1);
    int irmin = hpx_ring_above(h, zmax) + 1;
    if ((rlat1 <= 0) && (irmin > 1)) {          /* north pole in the disc */
        int sp, rp, sh;
        hpx_ring_info_small(h, irmin - 1, &sp, &rp, &sh);
        rset_append(out, 0, sp + rp);
    }
    double rlat2 = theta + rsmall;
    double zmin = O_COS(rlat2);
    int irmax = hpx_ring_above(h, zmin);
    for (int iz = irmin; iz <= irmax; ++iz) {
        double z = hpx_ring2z(h, iz);
        double x = (cosrbig - z * z0) * xa;
        double ysq = 1 - z * z - x * x;
        double dphi = -1;
        if (ysq <= 0)
            dphi = 0;                           /* fct == 1 */
        else
            dphi = O_ATAN2(sqrt(ysq), x);
        if (dphi > 0) {
            int nr, ipix1, shifted;
            hpx_ring_info_small(h, iz, &ipix1, &nr, &shifted);
            double shift = shifted ? 0.5 : 0.;
            int ipix2 = ipix1 + nr - 1;
            int ip_lo = o_ifloor(nr * O_INV_TWOPI * (phi - dphi) - shift) + 1;
            int ip_hi = o_ifloor(nr * O_INV_TWOPI * (phi + dphi) - shift);
            if (ip_lo <= ip_hi) {
                if (ip_hi >= nr) {
                    ip_lo -= nr;
                    ip_hi -= nr;
                }
                if (ip_lo < 0) {
                    rset_append(out, ipix1, ipix1 + ip_hi + 1);
                    rset_append(out, ipix1 + ip_lo + nr, ipix2 + 1);
                } else {
                    rset_append(out, ipix1 + ip_lo, ipix1 + ip_hi + 1);
                }
            }
        }
    }
    if ((rlat2 >= O_PI) && (irmax + 1 < 4 * h->nside)) { /* south pole */
        int sp, rp, sh;
        hpx_ring_info_small(h, irmax + 1, &sp, &rp, &sh);
        rset_append(out, sp, h->npix);
    }
}

static void hpx_get_interpol(const hpx_t *h, double theta, double phi,
                             int pix[4], double wgt[4])
{
    double z = O_COS(theta);
    int ir1 = hpx_ring_above(h, z);
    int ir2 = ir1 + 1;
    double theta1 = 0, theta2 = 0, w1, tmp, dphi;
    int sp, nr, shift, i1, i2;
    if (ir1 > 0) {
        hpx_ring_info2(h, ir1, &sp, &nr, &theta1, &shift);
        dphi = O_TWOPI / nr;
        tmp = (phi / dphi - .5 * shift);
        i1 = (tmp < 0) ? (int)tmp - 1 : (int)tmp;
        w1 = (phi - (i1 + .5 * shift) * dphi) / dphi;
        i2 = i1 + 1;
        if (i1 < 0) i1 += nr;
        if (i2 >= nr) i2 -= nr;
        pix[0] = sp + i1; pix[1] = sp + i2;
        wgt[0] = 1 - w1; wgt[1] = w1;
    }
    if (ir2 < (4 * h->nside)) {
        hpx_ring_info2(h, ir2, &sp, &nr, &theta2, &shift);
        dphi = O_TWOPI / nr;
        tmp = (phi / dphi - .5 * shift);
        i1 = (tmp < 0) ? (int)tmp - 1 : (int)tmp;
        w1 = (phi - (i1 + .5 * shift) * dphi) / dphi;
        i2 = i1 + 1;
        if (i1 < 0) i1 += nr;
        if (i2 >= nr) i2 -= nr;
        pix[2] = sp + i1; pix[3] = sp + i2;
        wgt[2] = 1 - w1; wgt[3] = w1;
    }
    if (ir1 == 0) {
        double wtheta = theta / theta2;
        wgt[2] *= wtheta; wgt[3] *= wtheta;
        double fac = (1 - wtheta) * 0.25;
        wgt[0] = fac; wgt[1] = fac; wgt[2] += fac; wgt[3] += fac;
        pix[0] = (pix[2] + 2) & 3;
        pix[1] = (pix[3] + 2) & 3;
    } else if (ir2 == 4 * h->nside) {
        double wtheta = (theta - theta1) / (O_PI - theta1);
        wgt[0] *= (1 - wtheta); wgt[1] *= (1 - wtheta);
        double fac = wtheta * 0.25;
        wgt[0] += fac; wgt[1] += fac; wgt[2] = fac; wgt[3] = fac;
        pix[2] = ((pix[0] + 2) & 3) + h->npix - 4;
        pix[3] = ((pix[1] + 2) & 3) + h->npix - 4;
    } else {
        double wtheta = (theta - theta1) / (theta2 - theta1);
        wgt[0] *= (1 - wtheta); wgt[1] *= (1 - wtheta);
        wgt[2] *= wtheta; wgt[3] *= wtheta;
    }
}

/* ------------------------------------------------------------------ */
/* Sphtrigo (src/Sphtrigo/sphtrigo.cpp:4-55)                           */
/* ------------------------------------------------------------------ */

static void o_rho_sigma_chi(double *rho, double *sigma, double *chi,
                            double ra_bc, double dec_bc, double psi_bc,
                            double ra_pix, double dec_pix)
{
    double alpha, beta;
    double cdc = O_COS(dec_bc);
    double sdc = O_SIN(dec_bc);
    double cdp = O_COS(dec_pix);
    double sdp = O_SIN(dec_pix);
    double dra = ra_pix - ra_bc;
    double sd = O_SIN(dra);
    double cd = O_COS(dra);
    double crho = sdc * sdp + cdc * cdp * cd;
    if (crho >= 1.0)
        *rho = 0.0;
    else
        *rho = O_ACOS(crho);
    if (*rho > 1.0e-6) {
        beta = O_ATAN2(sd * cdc, sdc * cdp - cdc * sdp * cd);
        alpha = O_ATAN2(sd * cdp, sdp * cdc - cdp * sdc * cd);
    } else {
        beta = O_HALFPI;
        alpha = O_HALFPI;
    }
    *sigma = O_PI - alpha - psi_bc;
    if (*sigma < 0.0)
        *sigma += 2.0 * O_PI;
    else if (*sigma > 2.0 * O_PI)
        *sigma -= 2.0 * O_PI;
    *chi = beta - *sigma;
    if (*chi < -O_PI)
        *chi += 2.0 * O_PI;
}

/* ------------------------------------------------------------------ */
/* Exported: HEALPix                                                   */
/* ------------------------------------------------------------------ */

void or_pix2ang(int nside, long n, const int *pix, double *theta, double *phi)
{
    hpx_t h = hpx_make(nside);
    for (long i = 0; i < n; i++)
        hpx_pix2ang(&h, pix[i], &theta[i], &phi[i]);
}

void or_ang2pix(int nside, long n, const double *theta, const double *phi,
                int *pix)
{
    hpx_t h = hpx_make(nside);
    for (long i = 0; i < n; i++)
        pix[i] = hpx_ang2pix(&h, theta[i], phi[i]);
}

void or_ring_above(int nside, long n, const double *z, int *ring)
{
    hpx_t h = hpx_make(nside);
    for (long i = 0; i < n; i++)
        ring[i] = hpx_ring_above(&h, z[i]);
}

void or_get_interpol(int nside, long n, const double *theta, const double *phi,
                     int *pix, double *wgt)
{
    hpx_t h = hpx_make(nside);
    for (long i = 0; i < n; i++)
        hpx_get_interpol(&h, theta[i], phi[i], pix + 4 * i, wgt + 4 * i);
}

/* Ranges of the non-inclusive disc.  Writes at most max_ranges [begin,end)
 * pairs into ranges and returns the number of ranges (which may exceed
 * max_ranges; then only the first max_ranges were written). */
long or_query_disc_ranges(int nside, double theta, double phi, double radius,
                          int *ranges, long max_ranges)
{
    hpx_t h = hpx_make(nside);
    rset_t s = {0, 0, 0};
    hpx_query_disc(&h, theta, phi, radius, &s);
    long n = s.n;
    for (long i = 0; i < n && i < max_ranges; i++) {
        ranges[2 * i] = s.r[2 * i];
        ranges[2 * i + 1] = s.r[2 * i + 1];
    }
    free(s.r);
    return n;
}

/* Pixel lists of the discs of many pointings, flattened: offsets[n+1],
 * pixels[...].  Pointing given as the reference passes it
 * (convolver.cpp:99: theta = pi/2 - dec, phi = ra, with float ra/dec).
 * Returns the total count; writes pixels only if pixels != NULL. */
long or_query_disc_many(int nside, long n, const float *ra, const float *dec,
                        double radius, long *offsets, int *pixels)
{
    hpx_t h = hpx_make(nside);
    rset_t s = {0, 0, 0};
    long total = 0;
    for (long i = 0; i < n; i++) {
        double dec_bc = dec[i];
        double ra_bc = (float)ra[i];
        hpx_query_disc(&h, O_HALFPI - dec_bc, ra_bc, radius, &s);
        if (offsets)
            offsets[i] = total;
        for (long r = 0; r < s.n; r++)
            for (int p = s.r[2 * r]; p < s.r[2 * r + 1]; p++) {
                if (pixels)
                    pixels[total] = p;
                total++;
            }
    }
    if (offsets)
        offsets[n] = total;
    free(s.r);
    return total;
}

/* ------------------------------------------------------------------ */
/* Exported: Sphtrigo and the sky->beam matching                       */
/* ------------------------------------------------------------------ */

void or_rho_sigma_chi(long n, const double *ra_bc, const double *dec_bc,
                      const double *psi_bc, const double *ra_pix,
                      const double *dec_pix, double *rho, double *sigma,
                      double *chi)
{
    for (long i = 0; i < n; i++)
        o_rho_sigma_chi(&rho[i], &sigma[i], &chi[i], ra_bc[i], dec_bc[i],
                        psi_bc[i], ra_pix[i], dec_pix[i]);
}

/* For each pointing i and each sky pixel of its disc (or_query_disc_many
 * layout), the beam pixel ang2pix(rho, sigma) and chi, as the CPU path
 * computes them (convolver.cpp:112-149; sphtrigo.cpp:27-54). */
void or_match(int nside_sky, int nside_beam, long n, const float *ra,
              const float *dec, const float *pa, const long *offsets,
              const int *pixels, int *bpix, double *chi_out)
{
    hpx_t hs = hpx_make(nside_sky);
    hpx_t hb = hpx_make(nside_beam);
    for (long i = 0; i < n; i++) {
        double ra_bc = (float)ra[i];
        double dec_bc = dec[i];
        double pa_bc = -pa[i];
        for (long k = offsets[i]; k < offsets[i + 1]; k++) {
            double th, ph, rho, sigma, chi;
            hpx_pix2ang(&hs, pixels[k], &th, &ph);
            o_rho_sigma_chi(&rho, &sigma, &chi, ra_bc, dec_bc, pa_bc, ph,
                            O_PI / 2.0 - th);
            bpix[k] = hpx_ang2pix(&hb, rho, sigma);
            chi_out[k] = chi;
        }
    }
}

/* ------------------------------------------------------------------ */
/* PolBeam (src/Polbeam/polbeam.cpp)                                   */
/* ------------------------------------------------------------------ */

/* rhoMax = pix2ang(nPixels).theta at the beam NSIDE (polbeam.cpp:35-36) */
double or_beam_rho_max(int nside_beam, long npix_beam)
{
    hpx_t h = hpx_make(nside_beam);
    double th, ph;
    hpx_pix2ang(&h, (int)npix_beam, &th, &ph);
    return th;
}

/* beam_from_fields (polbeam.cpp:117-184): tilde I,Q,U,V from E fields. */
void or_beam_from_fields(long npix, const double *magEco, const double *phsEco,
                         const double *magEcx, const double *phsEcx, float *I,
                         float *Q, float *U, float *V)
{
    for (long i = 0; i < npix; i++) {
        double cr = magEco[i] * cos(phsEco[i]), ci = magEco[i] * sin(phsEco[i]);
        double xr = magEcx[i] * cos(phsEcx[i]), xi = magEcx[i] * sin(phsEcx[i]);
        double nco = cr * cr + ci * ci;
        double ncx = xr * xr + xi * xi;
        /* Eco * conj(Ecx) */
        double pr = cr * xr - ci * (-xi);
        double pi_ = cr * (-xi) + ci * xr;
        I[i] = (float)(nco + ncx);
        Q[i] = (float)(nco - ncx);
        U[i] = (float)(2 * pr);
        V[i] = (float)(-2 * pi_);
    }
}

/* make_unpol_gaussian_elliptical_beams (polbeam.cpp:290-372): the E-field
 * magnitude of a perfectly co-polarised elliptical Gaussian. */
void or_gaussian_fields(int nside_beam, long npix, double fwhmx, double fwhmy,
                        double phi0, double *magEco, double *phsEco,
                        double *magEcx, double *phsEcx)
{
    hpx_t h = hpx_make(nside_beam);
    double deg2rad = M_PI / 180.0;
    double sigma_x = (deg2rad * fwhmx) / 2.35482;
    double sigma_y = (deg2rad * fwhmy) / 2.35482;
    double phi_0 = deg2rad * phi0;
    double a = (cos(phi_0) * cos(phi_0)) / (2 * sigma_x * sigma_x) +
               (sin(phi_0) * sin(phi_0)) / (2 * sigma_y * sigma_y);
    double b = -(sin(2 * phi_0)) / (4 * sigma_x * sigma_x) +
               (sin(2 * phi_0)) / (4 * sigma_y * sigma_y);
    double c = (sin(phi_0) * sin(phi_0)) / (2 * sigma_x * sigma_x) +
               (cos(phi_0) * cos(phi_0)) / (2 * sigma_y * sigma_y);
    for (long bpix = 0; bpix < npix; bpix++) {
        double rho, sig;
        hpx_pix2ang(&h, (int)bpix, &rho, &sig);
        double val = exp(-(a * cos(sig) * cos(sig) + 2.0 * b * cos(sig) * sin(sig) +
                           c * sin(sig) * sin(sig)) * rho * rho);
        magEco[bpix] = sqrt(val);
        magEcx[bpix] = 0.0;
        phsEco[bpix] = 0.0;
        phsEcx[bpix] = 0.0;
    }
}

/* PolBeam::build (polbeam.cpp:186-288).  tilde[8] = Ia,Qa,Ua,Va,Ib,Qb,Ub,Vb;
 * out[8] = aBeams[0..3], bBeams[0..3]. */
void or_polbeam_build(long npix, int nside_beam, int nside_sky, double epsilon,
                      char dets, const float *const *tilde, float *const *out)
{
    const float *Ia = tilde[0], *Qa = tilde[1], *Ua = tilde[2], *Va = tilde[3];
    const float *Ib = tilde[4], *Qb = tilde[5], *Ub = tilde[6], *Vb = tilde[7];
    double sumIa = 0, sumIb = 0;
    for (long i = 0; i < npix; i++) {
        if (dets == 'a' || dets == 'p')
            sumIa += Ia[i];
        if (dets == 'b' || dets == 'p')
            sumIb += Ib[i];
    }
    double ocmp = nside_sky / (double)nside_beam;
    ocmp = ocmp * ocmp;
    double aNorm = sumIa * ocmp;
    double bNorm = sumIb * ocmp;
    if (dets == 'a')
        bNorm += 1e-8;
    if (dets == 'b')
        aNorm += 1e-8;
    for (long i = 0; i < npix; i++) {
        out[0][i] = (float)((Ia[i] + epsilon * Ib[i]) / aNorm);
        out[1][i] = (float)((Qa[i] - epsilon * Qb[i]) / aNorm);
        out[2][i] = (float)((Ua[i] - epsilon * Ub[i]) / aNorm);
        out[3][i] = (float)((Va[i] + epsilon * Vb[i]) / aNorm);
        out[4][i] = (float)((Ib[i] + epsilon * Ia[i]) / bNorm);
        out[5][i] = (float)((Qb[i] - epsilon * Qa[i]) / bNorm);
        out[6][i] = (float)((Ub[i] - epsilon * Ua[i]) / bNorm);
        out[7][i] = (float)((Vb[i] + epsilon * Va[i]) / bNorm);
    }
}

/* ------------------------------------------------------------------ */
/* Convolver (src/Convolver/convolver.cpp)                             */
/* ------------------------------------------------------------------ */

typedef struct {
    hpx_t hs, hb;
    const float *I, *Q, *U;
    const float *V;   /* V-term mode: the sky V map (NULL: off) */
    long npix_beam;
    const float *a[3], *b[3];
    const float *a3, *b3;   /* aBeams[3], bBeams[3] (V-term mode) */
    double rho_max;
    int interp;   /* beam interpolated from the get_interpol neighbours */
} conv_t;

/* Beam interpolation mode of the forward and adjoints (or_set_beam_interpolation). */
static int g_interp = 0;

void or_set_beam_interpolation(int on) { g_interp = on != 0; }

/* V-term mode of the forward (or_set_v_term): the reference's GPU kernel
 * adds V A3 to d_a and V B3 to d_b (cudaConv2.cu:204,214,217-218), which the
 * CPU Convolver does not (convolver.cpp:157-164).  With a non-NULL sky V map
 * the forwards append sV * A3 (sV * B3) as the last term of the per-visit
 * float expression (FP64 in the FP64 forward); NULL restores the CPU
 * Convolver.  Process-wide; the nearest-pixel mode only. */
static const float *g_skyV = 0;

void or_set_v_term(const float *V) { g_skyV = V; }

/* The beam at (rho, sigma) interpolated from its get_interpol neighbours
 * inside the beam map, normalised by their weight sum: the interpolating
 * form of convolver.cpp:124-147 (there commented out), in its precision:
 * float values accumulated from double products, divided by the double
 * weight sum; the values start from zero at every visit.  Returns 0 when no
 * neighbour lies in the beam; nb, w, *ws describe the neighbours used. */
static int o_interp_beam(const conv_t *cv, double rho, double sigma, float va[3], float vb[3],
                         int nb[4], double w[4], double *ws_out)
{
    hpx_get_interpol(&cv->hb, rho, sigma, nb, w);
    double ws = 0.0;
    for (int b = 0; b < 3; b++) {
        va[b] = 0.0f;
        vb[b] = 0.0f;
    }
    for (int k = 0; k < 4; k++) {
        if (nb[k] < cv->npix_beam) {
            ws += w[k];
            for (int b = 0; b < 3; b++) {
                va[b] += (double)cv->a[b][nb[k]] * w[k];
                vb[b] += (double)cv->b[b][nb[k]] * w[k];
            }
        }
    }
    *ws_out = ws;
    if (!(ws > 0.0))
        return 0;
    for (int b = 0; b < 3; b++) {
        va[b] /= ws;
        vb[b] /= ws;
    }
    return 1;
}

/* Visit callback kinds */
enum { V_FWD = 0, V_FWD64 = 1, V_ADJ_MAP = 2, V_ADJ_BEAM = 3 };

/* Walks the disc of one sample in ascending pixel order and applies the
 * requested operation to every visit whose beam pixel is inside the beam
 * (convolver.cpp:98-165). */
static void o_sample(const conv_t *cv, double ra_bc, double dec_bc,
                     double pa_bc, int kind, double *da, double *db,
                     double ga, double gb, double *gmap, double *gbeam,
                     rset_t *rs)
{
    double data_a = 0.0, data_b = 0.0;
    hpx_query_disc(&cv->hs, O_HALFPI - dec_bc, ra_bc, cv->rho_max, rs);
    for (long rn = 0; rn < rs->n; rn++) {
        for (int skyPix = rs->r[2 * rn]; skyPix < rs->r[2 * rn + 1]; skyPix++) {
            double th, ph, rho = 0, sigma = 0, chi = 0;
            hpx_pix2ang(&cv->hs, skyPix, &th, &ph);
            double ra_pix = ph;
            double dec_pix = O_PI / 2.0 - th;
            o_rho_sigma_chi(&rho, &sigma, &chi, ra_bc, dec_bc, pa_bc, ra_pix,
                            dec_pix);
            double c2d = O_COS(2.0 * chi), s2d = O_SIN(2.0 * chi);
            float c2chi = (float)c2d, s2chi = (float)s2d;
            int bpix = 0, inside;
            float va[3], vb[3];
            int nb[4];
            double w[4], ws = 0.0;
            if (cv->interp) {
                inside = o_interp_beam(cv, rho, sigma, va, vb, nb, w, &ws);
            } else {
                bpix = hpx_ang2pix(&cv->hb, rho, sigma);
                inside = bpix < cv->npix_beam;
                for (int b = 0; inside && b < 3; b++) {
                    va[b] = cv->a[b][bpix];
                    vb[b] = cv->b[b][bpix];
                }
            }
            if (inside) {
                float ba0 = va[0], ba1 = va[1], ba2 = va[2];
                float bb0 = vb[0], bb1 = vb[1], bb2 = vb[2];
                float sI = cv->I[skyPix], sQ = cv->Q[skyPix], sU = cv->U[skyPix];
                if (kind == V_FWD && cv->V) {
                    /* the same with the V terms appended (cudaConv2.cu:217-218) */
                    float sV = cv->V[skyPix], ba3 = cv->a3[bpix], bb3 = cv->b3[bpix];
                    float ta = sI * (ba0) + sQ * (ba1 * c2chi - ba2 * s2chi) +
                               sU * (ba2 * c2chi + ba1 * s2chi) + sV * ba3;
                    float tb = sI * (bb0) + sQ * (-bb1 * c2chi + bb2 * s2chi) +
                               sU * (-bb2 * c2chi - bb1 * s2chi) + sV * bb3;
                    data_a += ta;
                    data_b += tb;
                } else if (kind == V_FWD) {
                    /* float expression, double accumulation (convolver.cpp:157-164) */
                    float ta = sI * (ba0) + sQ * (ba1 * c2chi - ba2 * s2chi) +
                               sU * (ba2 * c2chi + ba1 * s2chi);
                    float tb = sI * (bb0) + sQ * (-bb1 * c2chi + bb2 * s2chi) +
                               sU * (-bb2 * c2chi - bb1 * s2chi);
                    data_a += ta;
                    data_b += tb;
                } else if (kind == V_FWD64) {
                    data_a += (double)sI * ba0 + (double)sQ * (ba1 * c2d - ba2 * s2d) +
                              (double)sU * (ba2 * c2d + ba1 * s2d);
                    data_b += (double)sI * bb0 + (double)sQ * (-bb1 * c2d + bb2 * s2d) +
                              (double)sU * (-bb2 * c2d - bb1 * s2d);
                    if (cv->V) {
                        data_a += (double)cv->V[skyPix] * cv->a3[bpix];
                        data_b += (double)cv->V[skyPix] * cv->b3[bpix];
                    }
                } else if (kind == V_ADJ_MAP) {
                    double c = c2chi, s = s2chi;
                    double vI = ga * ba0 + gb * bb0;
                    double vQ = ga * (ba1 * c - ba2 * s) + gb * (-bb1 * c + bb2 * s);
                    double vU = ga * (ba2 * c + ba1 * s) + gb * (-bb2 * c - bb1 * s);
                    double *g = gmap + 3 * (long)skyPix;
#ifdef _OPENMP
#pragma omp atomic
#endif
                    g[0] += vI;
#ifdef _OPENMP
#pragma omp atomic
#endif
                    g[1] += vQ;
#ifdef _OPENMP
#pragma omp atomic
#endif
                    g[2] += vU;
                } else {
                    double c = c2chi, s = s2chi;
                    double q = sQ, u = sU;
                    double v[6] = {ga * sI, ga * (q * c + u * s), ga * (-q * s + u * c),
                                   gb * sI, gb * (-q * c - u * s), gb * (q * s - u * c)};
                    /* nearest: one row; interpolated: each neighbour in the
                     * beam with weight w / ws */
                    for (int n = 0; n < (cv->interp ? 4 : 1); n++) {
                        long row = cv->interp ? nb[n] : bpix;
                        if (row >= cv->npix_beam)
                            continue;
                        double f = cv->interp ? w[n] / ws : 1.0;
                        double *g = gbeam + 6 * row;
                        for (int k = 0; k < 6; k++) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                            g[k] += v[k] * f;
                        }
                    }
                }
            }
        }
    }
    if (da) *da = data_a;
    if (db) *db = data_b;
}

static conv_t o_conv_make(int nside_sky, const float *I, const float *Q,
                          const float *U, int nside_beam, long npix_beam,
                          const float *const *beams)
{
    conv_t cv;
    cv.hs = hpx_make(nside_sky);
    cv.hb = hpx_make(nside_beam);
    cv.I = I; cv.Q = Q; cv.U = U;
    cv.npix_beam = npix_beam;
    for (int k = 0; k < 3; k++) {
        cv.a[k] = beams[k];
        cv.b[k] = beams[4 + k];
    }
    cv.a3 = beams[3];
    cv.b3 = beams[7];
    cv.V = (g_skyV && beams[3] && beams[7] && !g_interp) ? g_skyV : 0;
    cv.rho_max = or_beam_rho_max(nside_beam, npix_beam);
    cv.interp = g_interp;
    return cv;
}

/* Convolver::exec_convolution (convolver.cpp:22-72).  beams[8] =
 * aBeams[0..3], bBeams[0..3].  precision64 != 0 computes the sum entirely
 * in FP64 (the FP64 oracle used to state errors).  Returns -1 on an
 * invalid polFlag (the reference throws std::invalid_argument). */
int or_convolve(int nside_sky, const float *I, const float *Q, const float *U,
                int nside_beam, long npix_beam, const float *const *beams,
                long nsamples, const float *ra, const float *dec,
                const float *pa, char polflag, int precision64, int nthreads,
                float *tod_a, float *tod_b, double *tod64_a, double *tod64_b)
{
    if (polflag != 'a' && polflag != 'b' && polflag != 'p')
        return -1;
    conv_t cv = o_conv_make(nside_sky, I, Q, U, nside_beam, npix_beam, beams);
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        rset_t rs = {0, 0, 0};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (long i = 0; i < nsamples; i++) {
            double da = 0.0, db = 0.0;
            float ra_bc = ra[i];
            double dec_bc = dec[i];
            double pa_bc = -pa[i];
            o_sample(&cv, ra_bc, dec_bc, pa_bc, precision64 ? V_FWD64 : V_FWD,
                     &da, &db, 0, 0, 0, 0, &rs);
            if (polflag == 'a' || polflag == 'p') {
                if (tod_a) tod_a[i] = (float)da;
                if (tod64_a) tod64_a[i] = da;
            }
            if (polflag == 'b' || polflag == 'p') {
                if (tod_b) tod_b[i] = (float)db;
                if (tod64_b) tod64_b[i] = db;
            }
        }
        free(rs.r);
    }
    return 0;
}

/* Adjoint of the forward map with respect to the sky, in FP64:
 * gmap[3*p + {0,1,2}] += sum over visits of the TOD cotangents times the
 * beam response (the transpose of convolver.cpp:157-164, with the beam
 * products and c2chi/s2chi as float values exactly as the forward uses). */
void or_adjoint_map(int nside_sky, int nside_beam, long npix_beam,
                    const float *const *beams, long nsamples, const float *ra,
                    const float *dec, const float *pa, const float *g_a,
                    const float *g_b, double *gmap)
{
    conv_t cv = o_conv_make(nside_sky, NULL, NULL, NULL, nside_beam, npix_beam,
                            beams);
    memset(gmap, 0, sizeof(double) * 3 * (size_t)cv.hs.npix);
    /* o_sample loads sky values for every visit; the map adjoint does not
     * use them, so it is given an all-zero sky. */
    float *zs = (float *)calloc((size_t)cv.hs.npix, sizeof(float));
    cv.I = zs; cv.Q = zs; cv.U = zs;
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        rset_t rt = {0, 0, 0};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (long i = 0; i < nsamples; i++) {
            float ra_bc = ra[i];
            double dec_bc = dec[i];
            double pa_bc = -pa[i];
            o_sample(&cv, ra_bc, dec_bc, pa_bc, V_ADJ_MAP, 0, 0,
                     g_a ? g_a[i] : 0.0, g_b ? g_b[i] : 0.0, gmap, 0, &rt);
        }
        free(rt.r);
    }
    free(zs);
}

/* Adjoint with respect to the beam maps, in FP64: gbeam[6*b + k] for
 * k = A0,A1,A2,B0,B1,B2 (the V maps get no gradient: the CPU forward does
 * not read them, convolver.cpp:157-164). */
void or_adjoint_beam(int nside_sky, const float *I, const float *Q,
                     const float *U, int nside_beam, long npix_beam,
                     const float *const *beams, long nsamples, const float *ra,
                     const float *dec, const float *pa, const float *g_a,
                     const float *g_b, double *gbeam)
{
    conv_t cv = o_conv_make(nside_sky, I, Q, U, nside_beam, npix_beam, beams);
    memset(gbeam, 0, sizeof(double) * 6 * (size_t)npix_beam);
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        rset_t rt = {0, 0, 0};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (long i = 0; i < nsamples; i++) {
            float ra_bc = ra[i];
            double dec_bc = dec[i];
            double pa_bc = -pa[i];
            o_sample(&cv, ra_bc, dec_bc, pa_bc, V_ADJ_BEAM, 0, 0,
                     g_a ? g_a[i] : 0.0, g_b ? g_b[i] : 0.0, 0, gbeam, &rt);
        }
        free(rt.r);
    }
}

/* Number of disc pixels per sample (exact |D_i|), for the byte model. */
long or_disc_sizes(int nside_sky, double radius, long nsamples,
                   const float *ra, const float *dec, int *sizes)
{
    hpx_t h = hpx_make(nside_sky);
    rset_t rs = {0, 0, 0};
    long total = 0;
    for (long i = 0; i < nsamples; i++) {
        double dec_bc = dec[i];
        double ra_bc = (float)ra[i];
        hpx_query_disc(&h, O_HALFPI - dec_bc, ra_bc, radius, &rs);
        long c = 0;
        for (long r = 0; r < rs.n; r++)
            c += rs.r[2 * r + 1] - rs.r[2 * r];
        if (sizes)
            sizes[i] = (int)c;
        total += c;
    }
    free(rs.r);
    return total;
}

/* Elementary functions as the oracle evaluates them (for the libm tests). */
void or_vec_sin(long n, const double *x, double *o) { for (long i = 0; i < n; i++) o[i] = O_SIN(x[i]); }
void or_vec_cos(long n, const double *x, double *o) { for (long i = 0; i < n; i++) o[i] = O_COS(x[i]); }
void or_vec_acos(long n, const double *x, double *o) { for (long i = 0; i < n; i++) o[i] = O_ACOS(x[i]); }
void or_vec_asin(long n, const double *x, double *o) { for (long i = 0; i < n; i++) o[i] = O_ASIN(x[i]); }
void or_vec_atan2(long n, const double *y, const double *x, double *o) { for (long i = 0; i < n; i++) o[i] = O_ATAN2(y[i], x[i]); }

/* The kernels form phi = ((iphi - 0.5) * pi/2) / rdiv in the polar caps as
 * q = x * inv, q + (x - q * rdiv) * inv with inv = 1.0 / rdiv.  Returns the
 * number of cap pixels of NSIDE <= nmax for which that differs from the
 * IEEE quotient (every rdiv < nmax and every iphi in [1, 4 rdiv]). */
long or_check_cap_division(int nmax)
{
    long bad = 0;
    for (int r = 1; r < nmax; r++) {
        double inv = 1.0 / r;
        double rd = r;
        for (int iphi = 1; iphi <= 4 * r; iphi++) {
            double x = (iphi - 0.5) * O_HALFPI;
            double q = x * inv;
            double e = fma(-q, rd, x);
            if (fma(e, inv, q) != x / r)
                bad++;
        }
    }
    return bad;
}

/* ------------------------------------------------------------------ */
/* Binned map-maker: SkyMap (src/SkyMap/skymap.cpp)                    */
/* ------------------------------------------------------------------ */

/* SkyMap::accumulate_data (skymap.cpp:68-116), with the reference's storage:
 * AtA[9 * pix + k] row-major 3x3 and AtD[3 * pix + k].  ra, dec, pa and data
 * are detector-major [ndet][nsamples]; ptg_stride is nsamples there, or 0 when
 * every detector shares one pointing (the same sums as passing the pointing
 * replicated per detector).  Accumulates onto what AtA and AtD hold.  Returns
 * the number of samples skipped: a theta outside [0, pi], where
 * Healpix_Base::ang2pix fails its planck_assert, or a non-finite ra, for which
 * it would return an arbitrary index. */
long or_skymap_accumulate(int nside, long nsamples, const float *ra, const float *dec,
                          const float *pa, long ptg_stride, int ndet,
                          const float *det_pol_angles, const float *data, double *AtA,
                          double *AtD)
{
    hpx_t h = hpx_make(nside);
    long skipped = 0;
    for (int det = 0; det < ndet; det++) {
        double pol_angle = det_pol_angles[det];
        for (long s = 0; s < nsamples; s++) {
            long ip = det * ptg_stride + s;
            double phi = ra[ip];
            double theta = O_HALFPI - (double)dec[ip];
            double psi = -(pol_angle + (double)pa[ip]);
            if (!(theta >= 0.0 && theta <= O_PI && fabs(phi) < INFINITY)) {
                skipped++;
                continue;
            }
            long p = hpx_ang2pix(&h, theta, phi);
            double d = data[det * nsamples + s];
            double twopsi = 2.0 * psi;
            double c = O_COS(twopsi);
            double sn = O_SIN(twopsi);
            AtD[p * 3 + 0] += 1.0 * d;
            AtD[p * 3 + 1] += c * d;
            AtD[p * 3 + 2] += sn * d;
            AtA[p * 9 + 0] += 1.0;
            AtA[p * 9 + 1] += c;
            AtA[p * 9 + 2] += sn;
            AtA[p * 9 + 3] += c;
            AtA[p * 9 + 4] += c * c;
            AtA[p * 9 + 5] += c * sn;
            AtA[p * 9 + 6] += sn;
            AtA[p * 9 + 7] += c * sn;
            AtA[p * 9 + 8] += sn * sn;
        }
    }
    return skipped;
}

/* LAPACK DGESV for n = 3, nrhs = 1, as the Netlib reference implementation
 * computes it: DGETF2 (right-looking LU, partial pivoting with IDAMAX's first
 * maximum, the column below the pivot scaled by the reciprocal of the pivot
 * when |pivot| >= DLAMCH('S'), rank-1 update by DGER, which skips a column
 * whose multiplier row entry is zero), then DGETRS (DLASWP, then DTRSM lower
 * unit and upper non-unit, each skipping a zero right-hand-side entry).
 * a is column-major a[i + 3 j]; on return b holds the solution unless the
 * result is non-zero (an exactly zero pivot: b is then left as it was, as
 * DGESV leaves it). */
static int o_dgesv3(double a[9], double b[3])
{
    const int n = 3;
    const double sfmin = 2.2250738585072014e-308;   /* DLAMCH('S') */
    int ipiv[3];
    int info = 0;
    for (int j = 0; j < n; j++) {
        int jp = j;
        double amax = fabs(a[j + 3 * j]);
        for (int i = j + 1; i < n; i++)
            if (fabs(a[i + 3 * j]) > amax) {
                amax = fabs(a[i + 3 * j]);
                jp = i;
            }
        ipiv[j] = jp;
        if (a[jp + 3 * j] != 0.0) {
            if (jp != j)
                for (int k = 0; k < n; k++) {
                    double t = a[j + 3 * k];
                    a[j + 3 * k] = a[jp + 3 * k];
                    a[jp + 3 * k] = t;
                }
            if (j < n - 1) {
                if (fabs(a[j + 3 * j]) >= sfmin) {
                    double r = 1.0 / a[j + 3 * j];
                    for (int i = j + 1; i < n; i++)
                        a[i + 3 * j] *= r;
                } else {
                    for (int i = j + 1; i < n; i++)
                        a[i + 3 * j] = a[i + 3 * j] / a[j + 3 * j];
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
        if (j < n - 1)
            for (int k = j + 1; k < n; k++) {
                if (a[j + 3 * k] != 0.0) {
                    double temp = -a[j + 3 * k];
                    for (int i = j + 1; i < n; i++)
                        a[i + 3 * k] = a[i + 3 * k] + a[i + 3 * j] * temp;
                }
            }
    }
    if (info != 0)
        return info;
    for (int i = 0; i < n; i++)
        if (ipiv[i] != i) {
            double t = b[i];
            b[i] = b[ipiv[i]];
            b[ipiv[i]] = t;
        }
    for (int k = 0; k < n; k++)
        if (b[k] != 0.0)
            for (int i = k + 1; i < n; i++)
                b[i] = b[i] - b[k] * a[i + 3 * k];
    for (int k = n - 1; k >= 0; k--)
        if (b[k] != 0.0) {
            b[k] = b[k] / a[k + 3 * k];
            for (int i = 0; i < k; i++)
                b[i] = b[i] - b[k] * a[i + 3 * k];
        }
    return 0;
}

/* SkyMap::solve_map (skymap.cpp:134-224): per pixel, hits = int(AtA[0]);
 * with at least 3 hits the 3x3 system AtA x = AtD is solved as DGESV does
 * (the matrix assembled column-major as at skymap.cpp:159-172); a singular
 * system leaves I, Q, U = AtD and sets hits to 0; fewer than 3 hits give
 * I = Q = U = 0 (INSUFFCOV) and keep the hits.  Returns the number of
 * singular pixels (the reference prints each one). */
long or_skymap_solve(int nside, const double *AtA, const double *AtD, float *I, float *Q,
                     float *U, int *hits)
{
    long npix = 12L * nside * nside;
    long singular = 0;
    for (long p = 0; p < npix; p++) {
        const double *m = AtA + 9 * p;
        double a[9] = {m[0], m[3], m[6], m[1], m[4], m[7], m[2], m[5], m[8]};
        double b[3] = {AtD[3 * p + 0], AtD[3 * p + 1], AtD[3 * p + 2]};
        double ii = 0.0, qq = 0.0, uu = 0.0;
        int h = (int)m[0];
        if (h >= 3) {
            int info = o_dgesv3(a, b);
            ii = b[0];
            qq = b[1];
            uu = b[2];
            if (info != 0) {
                h = 0;
                singular++;
            }
        }
        I[p] = (float)ii;
        Q[p] = (float)qq;
        U[p] = (float)uu;
        hits[p] = h;
    }
    return singular;
}

/* The 3x3 solve alone, for pinning against LAPACK: n systems, a column-major
 * [n][9], b [n][3] overwritten; info [n]. */
void or_dgesv3_many(long n, double *a, double *b, int *info)
{
    for (long i = 0; i < n; i++)
        info[i] = o_dgesv3(a + 9 * i, b + 3 * i);
}

/* ------------------------------------------------------------------ */
/* Focal plane to sky: BPoint::focal_plane_to_sky                      */
/* ------------------------------------------------------------------ */

/* src/Bpoint/bpoint.cpp:553-674: detector offsets (dx right, dy up, in the
 * receiver frame looking out, radians) and polarisation angles, array-centre
 * coordinates per sample -> detector coordinates per (detector, sample),
 * outputs detector-major [ndets][nsamples].  flag 0: az/el (x to the West),
 * 1: ra/dec (x to the East), 2: source-relative (x = -position angle).
 * Every expression keeps the reference's operation order. */
void or_focal_plane_to_sky(int ndets, const double *det_dx, const double *det_dy,
                           const double *det_pol_angle, long nsamples, const double *x_ac,
                           const double *y_ac, const double *rot_ac, int flag, double *x_det,
                           double *y_det, double *rot_det)
{
    for (int det = 0; det < ndets; det++) {
        double dx = det_dx[det];
        double cdx = O_COS(dx), sdx = O_SIN(dx);
        double cdy = O_COS(det_dy[det]), sdy = O_SIN(det_dy[det]);
        double polang = det_pol_angle[det];
        double cr = cdx * cdy;                         /* cos of the offset radius */
        double sr = sqrt((1.0 - cr) * (1.0 + cr));
        double alpha = O_ATAN2(sdy, sdx * cdy);        /* offset angle from +x */
        for (long s = 0; s < nsamples; s++) {
            long k = (long)det * nsamples + s;
            double gamma = O_HALFPI - alpha - rot_ac[s];
            double cg = O_COS(gamma);
            double sy = O_SIN(y_ac[s]), cy = O_COS(y_ac[s]);
            y_det[k] = O_ASIN(cr * sy + sr * cy * cg);
            double seta = O_SIN(rot_ac[s]), ceta = O_COS(rot_ac[s]);
            double cl = cy * ceta;
            double sl_seps = cy * seta * cdx - sy * sdx;
            double sl_ceps = sy * cdx + cy * seta * sdx;
            double pa_det = O_ATAN2(sl_seps, cl * cdy - sl_ceps * sdy);
            rot_det[k] = pa_det + polang;
            if (flag == 2) {
                x_det[k] = -pa_det;
            } else {
                double sg = O_SIN(gamma);
                double dxd = O_ATAN2(sr * sg, cr * cy - sr * sy * cg);
                x_det[k] = (flag == 0) ? x_ac[s] + dxd : x_ac[s] - dxd;
            }
        }
    }
}
