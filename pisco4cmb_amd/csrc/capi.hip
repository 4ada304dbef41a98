/*
 * capi.hip -- the extern "C" boundary (include/pisco.h): context, device
 * buffers, host <-> device staging and kernel launches.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "pisco.h"
#include "device_hpx.h"
#include "kernels.hpp"

using namespace pisco;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                              \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            return fail(PISCO_ERR_DEVICE, "%s failed: %s", #expr,                 \
                        hipGetErrorString(e_));                                   \
    } while (0)

#define CHK(expr)                                                                 \
    do {                                                                          \
        int rc_ = (expr);                                                         \
        if (rc_ != PISCO_OK)                                                      \
            return rc_;                                                           \
    } while (0)

constexpr int kMaxNside = 8192;   /* int32 pixel indices, as Healpix_Base<int> */

bool valid_nside(int n) { return n >= 1 && n <= kMaxNside; }

/* A grow-only device allocation.  While `frozen` is set (pisco_freeze_buffers:
 * a captured graph holds p as a kernel argument) a request that would move
 * the buffer fails instead of freeing it. */
template <typename T> struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;   /* elements */
    bool frozen = false;
    int ensure(size_t n)
    {
        if (n <= cap)
            return PISCO_OK;
        if (frozen)
            return fail(PISCO_ERR_STATE,
                        "a %zu-byte scratch buffer must grow to %zu bytes, but the context's "
                        "buffers are frozen (a captured graph refers to them); use another "
                        "context for larger calls, or unfreeze after the graph is gone",
                        cap * sizeof(T), n * sizeof(T));
        if (p)
            (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(PISCO_ERR_MEMORY, "hipMalloc of %zu bytes failed: %s",
                        n * sizeof(T), hipGetErrorString(e));
        }
        cap = n;
        return PISCO_OK;
    }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

} // namespace

int pisco::set_error(int code, const char *msg)
{
    g_err = msg;
    return code;
}

int pisco::pair_params(int nside_sky, const float *I, const float *Q, const float *U,
                       int nside_beam, long npix_beam, const float *beam8, long nsamples,
                       const float *ra, const float *dec, const float *pa, PairParams &p)
{
    if (!valid_nside(nside_sky) || !valid_nside(nside_beam))
        return fail(PISCO_ERR_INVALID, "NSIDE out of range [1, %d]", kMaxNside);
    long npb_max = 12L * nside_beam * nside_beam;
    if (npix_beam < 1 || npix_beam >= npb_max)
        return fail(PISCO_ERR_INVALID,
                    "beam npix %ld must be in [1, 12*nside_beam^2 - 1]", npix_beam);
    if (nsamples < 0)
        return fail(PISCO_ERR_INVALID, "negative nsamples");
    memset(&p, 0, sizeof p);
    p.hs = hpx_make(nside_sky);
    p.rings = nullptr;
    p.I = I;
    p.Q = Q;
    p.U = U;
    p.hb = hpx_make(nside_beam);
    p.npix_beam = (int)npix_beam;
    p.rho_max = pisco_beam_rho_max(nside_beam, npix_beam);
    p.cos_rmax = pl_cos(p.rho_max < kPi ? p.rho_max : kPi);
    p.fc = fast_consts(p.hb);
    p.fc64 = fast_consts64(p.hb);
    p.beam = beam8;
    p.ra = ra;
    p.dec = dec;
    p.pa = pa;
    p.s_begin = 0;
    p.s_end = nsamples;
    return PISCO_OK;
}

/* Measured on MI355X over 1e6 CES samples: 8 lanes took 0.59x the time of 64
 * at 135 pixels per disc and 0.88x at 540; 16 lanes 0.98x at 1,500; 64 lanes
 * were fastest at 6,000. */
int pisco::forward_lanes_auto(const PairParams &p)
{
    const double s = pl_sin(0.5 * std::min(p.rho_max, kPi));
    const double dp = (double)p.hs.npix * s * s;   /* expected pixels per disc */
    return dp < 1000.0 ? 8 : dp < 3000.0 ? 16 : 64;
}

struct pisco_ctx {
    pisco_run_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;   /* host-API TOD copies, overlapped with the forward */

    bool has_sky = false, has_beam = false, has_scan = false;
    int nside_sky = 0;
    DevBuf<float> I, Q, U;
    DevBuf<float> V;              /* sky V when update_sky had one (has_v) */
    bool has_v = false;
    int nside_beam = 0;
    long npix_beam = 0;
    double rho_max = 0.0;
    DevBuf<float> beam;           /* [npix][8]: A0 A1 A2 B0 B1 B2 A3 B3 */
    long nsamples = 0;
    DevBuf<float> ra, dec, pa;
    DevBuf<float> tod;            /* 2 * nsamples */
    DevBuf<float> cot;            /* 2 * nsamples cotangents */
    DevBuf<double> acc;           /* FP64 accumulators */
    DevBuf<float> fout;           /* float staging of adjoint outputs */
    std::map<int, RingRec *> ring_tabs;
    int ncu = 256;                /* compute units of the device */
    /* tuning (pisco_set_tuning) */
    int adjmap_chunk = -1;        /* samples per LDS window; 0 = direct atomics, -1 = auto */
    int adjmap_cap = 6144;        /* LDS window pixels */
    int adjmap_rw = 1024;         /* LDS window rings */
    int adjmap_blocks = 0;        /* persistent blocks; 0 = one block per window (work item) */
    int adjmap_support = 1;       /* size LDS windows by the beam's non-zero rows */
    int batch_group = 4;          /* pairs per wave in batched forwards (1: per-pair items) */
    int batch_adj_group = 4;      /* pairs per wave in batched sky adjoints */
    int beam_interp = 0;          /* 1: beam interpolated from get_interpol neighbours */
    int adjmap_det = -1;          /* sky adjoint sums: 1 integer (deterministic), 0 float,
                                     -1 integer for nearest-pixel beams, float with beam
                                     interpolation */
    int forward_lanes = 0;        /* lanes per sample of the forward: 64, 32, 16, 8; 0 = auto */
    int adjmap_lanes = 0;         /* lanes per sample of the sky adjoint: 64, 16, 8; 0 = auto */
    int batch_pair_major = -1;    /* batched launches pair by pair: 1 yes, 0 no, -1 auto */
    int batch_pair_range = 0;     /* pairs per batched launch: 4k; 0 = auto; -1 = all */
    int adjbeam_bands = 1;        /* beam adjoint in LDS bands (1) or global atomics (0) */
    int adjbeam_chunk = 1024;     /* samples per work item of the banded beam adjoint */
    int adjbeam_rows = 0;         /* beam rows per band; 0 = as many as fit the LDS */
    int adjbeam_lanes = 8;        /* lanes per sample of the banded beam adjoint: 8 or 64 */
    int adjbeam_records = 1;      /* beam adjoint through visit records (1) or the bands (0) */
    int adjbeam_pool_mb = 8192;   /* memory of the records per chunk of samples, MiB */
    int v_term = 0;               /* 1: host-API TOD adds V A3, V B3 as the reference's GPU kernel */
    DevBuf<unsigned> fxwork;      /* deterministic mode: max row sum, max |g| */
    DevBuf<double> fxscale;       /* deterministic mode: the scale (device) */
    DevBuf<int> support;          /* last non-zero beam row (device) */
    DevBuf<float> sky4;           /* the forward's packed sky: I, Q, U, 0 per pixel */
    DevBuf<float4> recs;          /* beam adjoint: visit records (16 bytes each) */
    DevBuf<long> recidx;          /* beam adjoint: record offsets per sample and band */
    DevBuf<float> recvmax;        /* beam adjoint: bounds of the record values */
    DevBuf<double> recpart;       /* beam adjoint: partial band sums */
    bool frozen = false;          /* pisco_freeze_buffers */

    template <typename F> void each_buf(F f)
    {
        f(I); f(Q); f(U); f(V); f(beam); f(ra); f(dec); f(pa); f(tod); f(cot); f(acc); f(fout);
        f(fxwork); f(fxscale); f(support); f(sky4); f(recs); f(recidx);
        f(recvmax); f(recpart);
    }
};

namespace {

/* Integer (deterministic) sums in the sky adjoint: the default -1 takes them
 * wherever the beam is nearest-pixel.  They also measured faster: 50.0
 * against 51.0 ms per 1e6 samples at the 5-degree, NSIDE 512 shape, 2.81
 * against 3.19 ms at 1.5 degrees, NSIDE 256, and 56.9 against 60.0 ms for
 * the batched 32-pair configuration-4 adjoint with groups of four pairs. */
bool det_single(const pisco_ctx *c)
{
    return c->adjmap_det > 0 || (c->adjmap_det < 0 && !c->beam_interp);
}

/* Stream of a pisco_dev_* call: the caller's hipStream_t; NULL is HIP's
 * null (legacy default) stream, which is also torch's default stream. */
hipStream_t pick(pisco_ctx *c, void *s)
{
    (void)c;
    return (hipStream_t)s;
}

int ring_table(pisco_ctx *c, int nside, const RingRec **out)
{
    auto it = c->ring_tabs.find(nside);
    if (it != c->ring_tabs.end()) {
        *out = it->second;
        return PISCO_OK;
    }
    RingRec *t = nullptr;
    HIPCHK(hipMalloc((void **)&t, sizeof(RingRec) * 4 * (size_t)nside));
    HIPCHK(launch_ring_table(hpx_make(nside), t, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ring_tabs[nside] = t;
    *out = t;
    return PISCO_OK;
}

double cos_radius(double r) { return pl_cos(r < kPi ? r : kPi); }

/* The reference's memory estimate (gpuconvolver.cu:41-64), one stream. */
size_t reference_mem_estimate(const pisco_run_config &cfg)
{
    size_t ptg = (size_t)cfg.ptgPerConv;
    size_t result = ptg * sizeof(float);
    size_t ptgbuf = 4 * ptg * sizeof(float);
    size_t pixbuf = ptg * (size_t)cfg.pixelsPerDisc * sizeof(int);
    size_t npixbuf = ptg * sizeof(int);
    return result + ptgbuf + npixbuf + pixbuf;
}

int setup_pair(pisco_ctx *c, int nside_sky, const float *I, const float *Q,
               const float *U, int nside_beam, long npix_beam, const float *beam8,
               long nsamples, const float *ra, const float *dec, const float *pa,
               PairParams &p)
{
    CHK(pair_params(nside_sky, I, Q, U, nside_beam, npix_beam, beam8, nsamples, ra, dec, pa,
                    p));
    const RingRec *rt = nullptr;
    CHK(ring_table(c, nside_sky, &rt));
    p.rings = rt;
    return PISCO_OK;
}

/* Expected sky pixels in a disc of radius rho_max: npix (1 - cos rho) / 2. */
double disc_pixels(const PairParams &p)
{
    const double s = pl_sin(0.5 * std::min(p.rho_max, kPi));
    return (double)p.hs.npix * s * s;
}

/* Lanes per sample of the nearest-pixel forward: the tuning value, or
 * (auto) forward_lanes_auto. */
int fwd_lanes(const pisco_ctx *c, const PairParams &p)
{
    if (c->forward_lanes > 0)
        return c->forward_lanes;
    return forward_lanes_auto(p);
}

/* Lanes per sample of the single-pair sky adjoint: the tuning value, or
 * (auto) from the expected disc size.  8 lanes took 0.80x the time of 64 at
 * 135 pixels per disc and 0.95x at 540; at 1,500 pixels 16 lanes equalled
 * 64 and 8 took 1.23x. */
int adj_lanes(const pisco_ctx *c, const PairParams &p)
{
    if (c->adjmap_lanes > 0)
        return c->adjmap_lanes;
    return disc_pixels(p) < 1000.0 ? 8 : 64;
}

/* Launches op over samples p.s_begin .. p.s_end in chunks of ptgPerConv (or
 * one launch). */
int run_pair(pisco_ctx *c, int op, PairParams p, hipStream_t st, bool sky_packed = false,
             const float *V = nullptr)
{
    const bool fwd = op == OP_FWD || op == OP_FWD_V;
    if (op == OP_FWD_V && c->beam_interp)
        return fail(PISCO_ERR_INVALID, "the V term is not available with beam interpolation");
    if (fwd && !c->beam_interp && sky_packed) {
        p.sky4 = (const float4 *)c->sky4.p;   /* packed by the caller on this stream */
    } else if (fwd && !c->beam_interp) {
        /* the nearest-pixel forward reads the sky packed as one float4 per
         * pixel (I, Q, U and, with the V term, V): one load per visit instead
         * of three */
        const long npix = p.hs.npix;
        CHK(c->sky4.ensure(4 * (size_t)npix));
        HIPCHK(launch_pack_sky(p.I, p.Q, p.U, npix, (float4 *)c->sky4.p, st,
                               op == OP_FWD_V ? V : nullptr));
        p.sky4 = (const float4 *)c->sky4.p;
    }
    const int lanes = (fwd && !c->beam_interp) ? fwd_lanes(c, p) : 64;
    const long n = p.s_end, s0 = p.s_begin;
    long chunk = (c->cfg.ptgPerConv > 0) ? c->cfg.ptgPerConv : n - s0;
    if (chunk < (1L << 20))
        chunk = std::max<long>(chunk, std::min<long>(n - s0, 1L << 20));
    for (long s = s0; s < n; s += chunk) {
        p.s_begin = s;
        p.s_end = std::min(n, s + chunk);
        if (c->beam_interp)
            HIPCHK(launch_pair_interp(op, p, st));
        else if (lanes < 64)
            HIPCHK(launch_fwd_sub(lanes, p, st, op == OP_FWD_V));
        else
            HIPCHK(launch_pair(op, p, st));
    }
    return PISCO_OK;
}

/* Deterministic sky adjoint: the scale (device) for the given beams and
 * cotangents, count = the number of contributions a pixel can receive at
 * most (samples x pairs). */
int det_scale(pisco_ctx *c, const float *const *beams, const long *npix, int nbeams,
              const float *g_a, const float *g_b, long ng, long count, hipStream_t st,
              const double **scale)
{
    CHK(c->fxwork.ensure(2));
    CHK(c->fxscale.ensure(1));
    HIPCHK(launch_fix_scale(beams, npix, nbeams, g_a, g_b, ng, count, c->fxwork.p, c->fxscale.p,
                            st));
    *scale = c->fxscale.p;
    return PISCO_OK;
}

/* Samples per LDS window of a single-pair sky adjoint over discs of radius
 * rho with `lanes` lanes per sample: the tuning value, or (auto) the largest
 * power of two in [16, 512] that keeps a window at about 96,000 expected
 * visits, and at least one sample per lane group of every wave of the block
 * (16 x 64 / lanes).  Each window costs three block-wide passes besides its
 * visits; at 6,000 pixels per disc 16 samples per window measured fastest,
 * at 135 pixels 256 samples took 0.64x the time of 16 (one sample per wave)
 * and, with 8 lanes per sample, 0.97x the time of 128; at 540 pixels with 8
 * lanes, 32 samples took 2.0x the time of 128. */
int adj_chunk(const pisco_ctx *c, int nside_sky, double rho, int lanes)
{
    if (c->adjmap_chunk > 0)
        return c->adjmap_chunk;
    const double s = pl_sin(0.5 * std::min(rho, kPi));
    const double disc = 12.0 * nside_sky * nside_sky * s * s;   /* npix (1 - cos rho) / 2 */
    int chunk = 16;
    while (chunk < 512 && 2.0 * chunk * disc <= 96000.0)
        chunk *= 2;
    const int fill = 16 * (64 / std::max(lanes, 1));   /* lane groups of a 16-wave block */
    return std::max(chunk, std::min(fill, 512));
}

/* The window capacity that fits the LDS next to the per-wave tables. */
int fit_cap(int cap, int rw, int group, bool det)
{
    while (cap > 256 && adjmap_lds_bytes(cap, rw, group, det) > 163840)
        cap -= 256;
    return cap;
}

/* The batched launches compute the nearest-pixel convolution only. */
int no_interp(pisco_ctx *c)
{
    if (c->beam_interp)
        return fail(PISCO_ERR_INVALID,
                    "beam interpolation is available for single-pair launches only");
    return PISCO_OK;
}

int need_state(pisco_ctx *c, bool beam_needed, bool sky_needed)
{
    if (!c)
        return fail(PISCO_ERR_INVALID, "NULL context");
    if (!c->has_scan)
        return fail(PISCO_ERR_STATE, "update_scan must be called before exec");
    if (beam_needed && !c->has_beam)
        return fail(PISCO_ERR_STATE, "update_beam must be called before exec");
    if (sky_needed && !c->has_sky)
        return fail(PISCO_ERR_STATE, "update_sky must be called before exec");
    return PISCO_OK;
}

} // namespace

extern "C" {

const char *pisco_last_error(void) { return g_err.c_str(); }

int pisco_version(char *buf, int len)
{
    int rt = 0;
    if (hipRuntimeGetVersion(&rt) != hipSuccess)
        rt = 0;
    if (buf && len > 0)
        snprintf(buf, (size_t)len, "pisco4cmb_amd 0.1 gfx950 hip-runtime %d", rt);
    return PISCO_OK;
}

double pisco_beam_rho_max(int nside_beam, long npix)
{
    if (!valid_nside(nside_beam) || npix < 0 || npix >= 12L * nside_beam * nside_beam)
        return -1.0;
    double th, ph;
    pix2ang(hpx_make(nside_beam), (int)npix, th, ph);
    return th;
}

long pisco_band_annulus(int nside_sky, int nside_beam, long npix_beam, int rows, int band,
                        float ra, float dec, int cover, long cap, int *pix)
{
    if (!valid_nside(nside_sky) || !valid_nside(nside_beam) || npix_beam < 1 ||
        npix_beam >= 12L * nside_beam * nside_beam || rows < 1 || band < 0 ||
        (long)band * rows >= npix_beam || cap < 0 || (cap > 0 && !pix))
        return -1;
    const Hpx hs = hpx_make(nside_sky), hb = hpx_make(nside_beam);
    const double rho_max = pisco_beam_rho_max(nside_beam, npix_beam);
    const int lo = band * rows;
    const int nrow = (int)std::min<long>(npix_beam - lo, rows);
    bool inner;
    double r_in, c_in, r_out, c_out;
    band_annulus_radii(hb, rho_max, cos_radius(rho_max), lo, nrow, inner, r_in, c_in, r_out,
                       c_out);
    const Centre c = centre_make(ra, dec, 0.0f);
    Disc dout, din;
    disc_setup(hs, kHalfPi - c.dec, c.ra, r_out, c_out, dout);
    if (inner)
        disc_setup(hs, kHalfPi - c.dec, c.ra, r_in, c_in, din);
    else
        din = dout;
    long n = 0;
    for (int r = dout.r_first; r <= dout.r_last; r++) {
        int p0[4], len[4];
        const int k = cover ? annulus_ring<true>(hs, dout, din, inner, r, p0, len)
                            : annulus_ring<false>(hs, dout, din, inner, r, p0, len);
        for (int q = 0; q < k; q++)
            for (int v = 0; v < len[q]; v++) {
                if (n >= cap)
                    return -2;
                pix[n++] = p0[q] + v;
            }
    }
    return n;
}

long pisco_ang2pix(int nside, double theta, double phi)
{
    if (!valid_nside(nside) || !(theta >= 0.0 && theta <= kPi) || !(fabs(phi) < 1e300))
        return -1;
    return ang2pix(hpx_make(nside), theta, phi);
}

int pisco_pix2ang(int nside, long pix, double *theta, double *phi)
{
    if (!valid_nside(nside) || pix < 0 || pix >= 12L * nside * nside || !theta || !phi)
        return fail(PISCO_ERR_INVALID, "bad argument");
    pix2ang(hpx_make(nside), (int)pix, *theta, *phi);
    return PISCO_OK;
}

int pisco_set_tuning(pisco_ctx *c, const char *key, long value)
{
    if (!c || !key)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    std::string k(key);
    if (k == "adjmap_chunk") {
        if (value < -1 || value > 4096)
            return fail(PISCO_ERR_INVALID, "adjmap_chunk must be -1 (auto) or in [0, 4096]");
        c->adjmap_chunk = (int)value;
    } else if (k == "adjmap_cap" || k == "adjmap_rw") {
        int cap = k == "adjmap_cap" ? (int)value : c->adjmap_cap;
        int rw = k == "adjmap_rw" ? (int)value : c->adjmap_rw;
        if (value < 1 || adjmap_lds_bytes(cap, rw, 0, false) > 163840)
            return fail(PISCO_ERR_INVALID, "LDS window does not fit 160 KiB");
        c->adjmap_cap = cap;
        c->adjmap_rw = rw;
    } else if (k == "adjmap_support") {
        if (value != 0 && value != 1)
            return fail(PISCO_ERR_INVALID, "adjmap_support must be 0 or 1");
        c->adjmap_support = (int)value;
    } else if (k == "batch_group") {
        if (value != 1 && value != 2 && value != 4)
            return fail(PISCO_ERR_INVALID, "batch_group must be 1, 2 or 4");
        c->batch_group = (int)value;
    } else if (k == "batch_adj_group") {
        if (value != 1 && value != 2 && value != 4)
            return fail(PISCO_ERR_INVALID, "batch_adj_group must be 1, 2 or 4");
        c->batch_adj_group = (int)value;
    } else if (k == "adjmap_deterministic") {
        if (value < -1 || value > 1)
            return fail(PISCO_ERR_INVALID, "adjmap_deterministic must be -1 (auto), 0 or 1");
        c->adjmap_det = (int)value;
    } else if (k == "beam_interpolation") {
        if (value != 0 && value != 1)
            return fail(PISCO_ERR_INVALID, "beam_interpolation must be 0 or 1");
        c->beam_interp = (int)value;
    } else if (k == "forward_lanes") {
        if (value != 0 && value != 8 && value != 16 && value != 32 && value != 64)
            return fail(PISCO_ERR_INVALID, "forward_lanes must be 0 (auto), 8, 16, 32 or 64");
        c->forward_lanes = (int)value;
    } else if (k == "batch_pair_range") {
        if (value < -1 || value > 65536 || (value > 0 && value % 4))
            return fail(PISCO_ERR_INVALID,
                        "batch_pair_range must be -1 (all), 0 (auto) or a multiple of 4");
        c->batch_pair_range = (int)value;
    } else if (k == "adjbeam_bands") {
        if (value != 0 && value != 1)
            return fail(PISCO_ERR_INVALID, "adjbeam_bands must be 0 or 1");
        c->adjbeam_bands = (int)value;
    } else if (k == "adjbeam_lanes") {
        if (value != 8 && value != 64)
            return fail(PISCO_ERR_INVALID, "adjbeam_lanes must be 8 or 64");
        c->adjbeam_lanes = (int)value;
    } else if (k == "adjbeam_rows") {
        if (value < 0 || value > (1L << 30))
            return fail(PISCO_ERR_INVALID, "adjbeam_rows out of range");
        c->adjbeam_rows = (int)value;
    } else if (k == "adjbeam_records") {
        if (value != 0 && value != 1)
            return fail(PISCO_ERR_INVALID, "adjbeam_records must be 0 or 1");
        c->adjbeam_records = (int)value;
    } else if (k == "adjbeam_pool_mb") {
        if (value < 1 || value > (1L << 20))
            return fail(PISCO_ERR_INVALID, "adjbeam_pool_mb must be in [1, 2^20]");
        c->adjbeam_pool_mb = (int)value;
    } else if (k == "adjbeam_chunk") {
        if (value < 1 || value > (1L << 24))
            return fail(PISCO_ERR_INVALID, "adjbeam_chunk must be in [1, 2^24]");
        c->adjbeam_chunk = (int)value;
    } else if (k == "batch_pair_major") {
        if (value < -1 || value > 1)
            return fail(PISCO_ERR_INVALID, "batch_pair_major must be -1 (auto), 0 or 1");
        c->batch_pair_major = (int)value;
    } else if (k == "adjmap_lanes") {
        if (value != 0 && value != 8 && value != 16 && value != 64)
            return fail(PISCO_ERR_INVALID, "adjmap_lanes must be 0 (auto), 8, 16 or 64");
        c->adjmap_lanes = (int)value;
    } else if (k == "adjmap_blocks") {
        if (value < 0 || value > 65536)
            return fail(PISCO_ERR_INVALID, "adjmap_blocks out of range");
        c->adjmap_blocks = (int)value;
    } else if (k == "v_term") {
        if (value != 0 && value != 1)
            return fail(PISCO_ERR_INVALID, "v_term must be 0 or 1");
        c->v_term = (int)value;
    } else {
        return fail(PISCO_ERR_INVALID, "unknown tuning key '%s'", key);
    }
    return PISCO_OK;
}

int pisco_get_tuning(pisco_ctx *c, const char *key, long *value)
{
    if (!c || !key || !value)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    std::string k(key);
    if (k == "adjmap_chunk")
        *value = c->adjmap_chunk;
    else if (k == "forward_lanes")
        *value = c->forward_lanes;
    else if (k == "adjmap_lanes")
        *value = c->adjmap_lanes;
    else if (k == "batch_pair_major")
        *value = c->batch_pair_major;
    else if (k == "batch_pair_range")
        *value = c->batch_pair_range;
    else if (k == "adjbeam_bands")
        *value = c->adjbeam_bands;
    else if (k == "adjbeam_chunk")
        *value = c->adjbeam_chunk;
    else if (k == "adjbeam_rows")
        *value = c->adjbeam_rows;
    else if (k == "adjbeam_lanes")
        *value = c->adjbeam_lanes;
    else if (k == "adjbeam_records")
        *value = c->adjbeam_records;
    else if (k == "adjbeam_pool_mb")
        *value = c->adjbeam_pool_mb;

    else if (k == "adjmap_cap")
        *value = c->adjmap_cap;
    else if (k == "adjmap_rw")
        *value = c->adjmap_rw;
    else if (k == "adjmap_blocks")
        *value = c->adjmap_blocks;
    else if (k == "adjmap_support")
        *value = c->adjmap_support;
    else if (k == "batch_group")
        *value = c->batch_group;
    else if (k == "beam_interpolation")
        *value = c->beam_interp;
    else if (k == "adjmap_deterministic")
        *value = c->adjmap_det;
    else if (k == "batch_adj_group")
        *value = c->batch_adj_group;
    else if (k == "cu_count")
        *value = c->ncu;
    else if (k == "v_term")
        *value = c->v_term;
    else
        return fail(PISCO_ERR_INVALID, "unknown tuning key '%s'", key);
    return PISCO_OK;
}

int pisco_beam_gaussian_fields(int nside_beam, long npix, double fwhmx, double fwhmy,
                               double phi0, double *magEco, double *phsEco,
                               double *magEcx, double *phsEcx)
{
    if (!valid_nside(nside_beam) || npix < 0 || npix > 12L * nside_beam * nside_beam)
        return fail(PISCO_ERR_INVALID, "bad beam geometry");
    if (!magEco || !phsEco || !magEcx || !phsEcx)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    if (!(fwhmx > 0) || !(fwhmy > 0))
        return fail(PISCO_ERR_INVALID, "FWHM must be positive");
    Hpx h = hpx_make(nside_beam);
    const double deg2rad = M_PI / 180.0;
    /* FWHM = 2 sqrt(2 ln 2) sigma */
    double sx = (deg2rad * fwhmx) / 2.35482;
    double sy = (deg2rad * fwhmy) / 2.35482;
    double p0 = deg2rad * phi0;
    double ca = cos(p0), sa = sin(p0), s2 = sin(2 * p0);
    double qa = (ca * ca) / (2 * sx * sx) + (sa * sa) / (2 * sy * sy);
    double qb = -(s2) / (4 * sx * sx) + (s2) / (4 * sy * sy);
    double qc = (sa * sa) / (2 * sx * sx) + (ca * ca) / (2 * sy * sy);
    for (long i = 0; i < npix; i++) {
        double rho, sig;
        pix2ang(h, (int)i, rho, sig);
        double cs = cos(sig), sn = sin(sig);
        double val = exp(-(qa * cs * cs + 2.0 * qb * cs * sn + qc * sn * sn) * rho * rho);
        magEco[i] = sqrt(val);
        magEcx[i] = 0.0;
        phsEco[i] = 0.0;
        phsEcx[i] = 0.0;
    }
    return PISCO_OK;
}

int pisco_beam_from_fields(long npix, const double *magEco, const double *phsEco,
                           const double *magEcx, const double *phsEcx, float *I, float *Q,
                           float *U, float *V)
{
    if (npix < 0 || !magEco || !phsEco || !magEcx || !phsEcx || !I || !Q || !U || !V)
        return fail(PISCO_ERR_INVALID, "bad argument");
    for (long i = 0; i < npix; i++) {
        /* E_co, E_cx as complex numbers (re, im) */
        double cr = magEco[i] * cos(phsEco[i]), ci = magEco[i] * sin(phsEco[i]);
        double xr = magEcx[i] * cos(phsEcx[i]), xi = magEcx[i] * sin(phsEcx[i]);
        double nco = cr * cr + ci * ci, ncx = xr * xr + xi * xi;
        double pr = cr * xr - ci * (-xi);     /* Re(E_co conj(E_cx)) */
        double pim = cr * (-xi) + ci * xr;    /* Im(E_co conj(E_cx)) */
        I[i] = (float)(nco + ncx);
        Q[i] = (float)(nco - ncx);
        U[i] = (float)(2 * pr);
        V[i] = (float)(-2 * pim);
    }
    return PISCO_OK;
}

int pisco_polbeam_build(long npix, int nside_beam, int nside_sky, double epsilon, char dets,
                        const float *const tilde[8], float *const out[8])
{
    if (npix < 0 || !tilde || !out || !valid_nside(nside_beam) || !valid_nside(nside_sky))
        return fail(PISCO_ERR_INVALID, "bad argument");
    if (dets != 'a' && dets != 'b' && dets != 'p')
        return fail(PISCO_ERR_INVALID, "enabledDets must be 'a', 'b' or 'p'");
    for (int k = 0; k < 8; k++)
        if (!tilde[k] || !out[k])
            return fail(PISCO_ERR_INVALID, "NULL map %d", k);
    const float *Ia = tilde[0], *Qa = tilde[1], *Ua = tilde[2], *Va = tilde[3];
    const float *Ib = tilde[4], *Qb = tilde[5], *Ub = tilde[6], *Vb = tilde[7];
    double sumIa = 0, sumIb = 0;   /* solid angles */
    for (long i = 0; i < npix; i++) {
        if (dets == 'a' || dets == 'p')
            sumIa += Ia[i];
        if (dets == 'b' || dets == 'p')
            sumIb += Ib[i];
    }
    double ocmp = nside_sky / (double)nside_beam;
    ocmp = ocmp * ocmp;
    double aNorm = sumIa * ocmp, bNorm = sumIb * ocmp;
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
    return PISCO_OK;
}

int pisco_create(const pisco_run_config *cfg, pisco_ctx **out)
{
    if (!cfg || !out)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    *out = nullptr;
    pisco_run_config c = *cfg;
    c.nStreams = 1;   /* as the reference (gpuconvolver.cu:39) */
    if (c.ptgPerConv <= 0)
        return fail(PISCO_ERR_INVALID, "ptgPerConv must be positive");
    if (c.pixelsPerDisc < 0)
        return fail(PISCO_ERR_INVALID, "pixelsPerDisc must be non-negative");
    size_t req = reference_mem_estimate(c);
    if (!(req < c.maxMemUsage))
        return fail(PISCO_ERR_MEMORY,
                    "GPUConvolver has an invalid configuration: requested memory %zu "
                    "bytes is not below maxMemUsage %zu bytes",
                    req, c.maxMemUsage);
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (c.deviceId < 0 || c.deviceId >= ndev)
        return fail(PISCO_ERR_INVALID, "deviceId %d out of range (%d devices)",
                    c.deviceId, ndev);
    HIPCHK(hipSetDevice(c.deviceId));
    pisco_ctx *ctx = new pisco_ctx();
    ctx->cfg = c;
    ctx->device = c.deviceId;
    if (hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, c.deviceId) !=
            hipSuccess ||
        ctx->ncu <= 0)
        ctx->ncu = 256;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(PISCO_ERR_DEVICE, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return PISCO_OK;
}

void pisco_destroy(pisco_ctx *ctx)
{
    if (!ctx)
        return;
    /* teardown: errors are not reportable here */
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->each_buf([](auto &b) { b.release(); });
    for (auto &kv : ctx->ring_tabs)
        (void)hipFree(kv.second);
    if (ctx->copy_stream) {
        (void)hipStreamSynchronize(ctx->copy_stream);
        (void)hipStreamDestroy(ctx->copy_stream);
    }
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int pisco_freeze_buffers(pisco_ctx *c, int frozen)
{
    if (!c)
        return fail(PISCO_ERR_INVALID, "NULL context");
    c->frozen = frozen != 0;
    c->each_buf([&](auto &b) { b.frozen = c->frozen; });
    return PISCO_OK;
}

int pisco_update_sky(pisco_ctx *c, int nside, const float *I, const float *Q,
                     const float *U, const float *V)
{
    /* V is kept for the V-term option (v_term); the CPU Convolver does not
     * read it (convolver.cpp:157-164) */
    if (!c || !I || !Q || !U)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    if (!valid_nside(nside))
        return fail(PISCO_ERR_INVALID, "sky NSIDE %d out of range [1, %d]", nside,
                    kMaxNside);
    HIPCHK(hipSetDevice(c->device));
    size_t n = 12UL * nside * nside;
    CHK(c->I.ensure(n));
    CHK(c->Q.ensure(n));
    CHK(c->U.ensure(n));
    HIPCHK(hipMemcpyAsync(c->I.p, I, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->Q.p, Q, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->U.p, U, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    c->has_v = V != nullptr;
    if (V) {
        CHK(c->V.ensure(n));
        HIPCHK(hipMemcpyAsync(c->V.p, V, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    const RingRec *rt;
    CHK(ring_table(c, nside, &rt));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->nside_sky = nside;
    c->has_sky = true;
    return PISCO_OK;
}

int pisco_update_beam(pisco_ctx *c, int nside_beam, long npix, const float *const a[4],
                      const float *const b[4])
{
    if (!c || !a || !b)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    for (int k = 0; k < 3; k++)
        if (!a[k] || !b[k])
            return fail(PISCO_ERR_INVALID, "beam maps 0..2 of a and b are required");
    if (!valid_nside(nside_beam))
        return fail(PISCO_ERR_INVALID, "beam NSIDE %d out of range", nside_beam);
    if (npix < 1 || npix >= 12L * nside_beam * nside_beam)
        return fail(PISCO_ERR_INVALID, "beam npix %ld out of range", npix);
    HIPCHK(hipSetDevice(c->device));
    CHK(c->beam.ensure(8 * (size_t)npix));
    /* host transpose into rows A0 A1 A2 B0 B1 B2 A3 B3 (update_beam packs
     * float4 rows per detector, gpuconvolver.cu:302-340) */
    float *h = nullptr;
    HIPCHK(hipHostMalloc((void **)&h, sizeof(float) * 8 * (size_t)npix, 0));
    for (long i = 0; i < npix; i++) {
        float *o = h + 8 * i;
        o[0] = a[0][i]; o[1] = a[1][i]; o[2] = a[2][i];
        o[3] = b[0][i]; o[4] = b[1][i]; o[5] = b[2][i];
        o[6] = a[3] ? a[3][i] : 0.f; o[7] = b[3] ? b[3][i] : 0.f;
    }
    hipError_t e = hipMemcpyAsync(c->beam.p, h, sizeof(float) * 8 * (size_t)npix,
                                  hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    (void)hipHostFree(h);
    HIPCHK(e);
    c->nside_beam = nside_beam;
    c->npix_beam = npix;
    c->rho_max = pisco_beam_rho_max(nside_beam, npix);
    c->has_beam = true;
    return PISCO_OK;
}

int pisco_update_scan(pisco_ctx *c, long nsamples, const float *ra, const float *dec,
                      const float *pa)
{
    if (!c || (nsamples > 0 && (!ra || !dec || !pa)))
        return fail(PISCO_ERR_INVALID, "NULL argument");
    if (nsamples < 0)
        return fail(PISCO_ERR_INVALID, "negative nsamples");
    HIPCHK(hipSetDevice(c->device));
    size_t n = (size_t)nsamples;
    CHK(c->ra.ensure(n));
    CHK(c->dec.ensure(n));
    CHK(c->pa.ensure(n));
    if (n) {
        HIPCHK(hipMemcpyAsync(c->ra.p, ra, n * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->dec.p, dec, n * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->pa.p, pa, n * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    c->nsamples = nsamples;
    c->has_scan = true;
    return PISCO_OK;
}

int pisco_exec_convolution(pisco_ctx *c, char polflag, float *tod_a, float *tod_b)
{
    if (!c)
        return fail(PISCO_ERR_INVALID, "NULL context");
    if (polflag != 'a' && polflag != 'b' && polflag != 'p')
        return fail(PISCO_ERR_INVALID, "valid polFlags are 'a', 'b' and 'p'");
    if ((polflag != 'b' && !tod_a) || (polflag != 'a' && !tod_b))
        return fail(PISCO_ERR_INVALID, "NULL output for the requested polFlag");
    CHK(need_state(c, true, true));
    HIPCHK(hipSetDevice(c->device));
    long n = c->nsamples;
    if (n == 0)
        return PISCO_OK;
    CHK(c->tod.ensure(2 * (size_t)n));
    PairParams p;
    CHK(setup_pair(c, c->nside_sky, c->I.p, c->Q.p, c->U.p, c->nside_beam, c->npix_beam,
                   (const float *)c->beam.p, n, c->ra.p, c->dec.p, c->pa.p, p));
    p.tod_a = c->tod.p;
    p.tod_b = c->tod.p + n;
    /* The forward in up to 4 launches, each followed by an event; the copies
     * of a launch's TOD to the host run on a second stream behind its event,
     * so they overlap the launches after it.  A sample's TOD does not depend
     * on how the samples are split into launches. */
    if (!c->copy_stream)
        HIPCHK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    const int K = n >= (1L << 18) ? 4 : 1;
    const long part = (n + K - 1) / K;
    hipEvent_t ev[4];
    int nev = 0;
    int rc = PISCO_OK;
    for (int k = 0; k < K && rc == PISCO_OK; k++) {
        PairParams q = p;
        q.s_begin = k * part;
        q.s_end = std::min(n, (k + 1) * part);
        if (q.s_begin >= q.s_end)
            break;
        rc = run_pair(c, c->v_term ? OP_FWD_V : OP_FWD, q, c->stream, k > 0,
                      c->has_v ? c->V.p : nullptr);
        if (rc == PISCO_OK) {
            hipError_t e = hipEventCreateWithFlags(&ev[nev], hipEventDisableTiming);
            if (e == hipSuccess)
                e = hipEventRecord(ev[nev], c->stream);
            if (e != hipSuccess)
                rc = fail(PISCO_ERR_DEVICE, "event failed: %s", hipGetErrorString(e));
            else
                nev++;
        }
    }
    for (int k = 0; k < nev && rc == PISCO_OK; k++) {
        const long b = k * part, e_ = std::min(n, (k + 1) * part);
        hipError_t e = hipStreamWaitEvent(c->copy_stream, ev[k], 0);
        if (e == hipSuccess && polflag != 'b')
            e = hipMemcpyAsync(tod_a + b, c->tod.p + b, (e_ - b) * 4, hipMemcpyDeviceToHost,
                               c->copy_stream);
        if (e == hipSuccess && polflag != 'a')
            e = hipMemcpyAsync(tod_b + b, c->tod.p + n + b, (e_ - b) * 4, hipMemcpyDeviceToHost,
                               c->copy_stream);
        if (e != hipSuccess)
            rc = fail(PISCO_ERR_DEVICE, "TOD copy failed: %s", hipGetErrorString(e));
    }
    hipError_t es = hipStreamSynchronize(c->copy_stream);
    hipError_t ek = hipStreamSynchronize(c->stream);
    for (int k = 0; k < nev; k++)
        (void)hipEventDestroy(ev[k]);
    if (rc != PISCO_OK)
        return rc;
    HIPCHK(es);
    HIPCHK(ek);
    return PISCO_OK;
}

int pisco_exec_adjoint_map(pisco_ctx *c, const float *g_a, const float *g_b, float *gI,
                           float *gQ, float *gU)
{
    if (!c || !gI || !gQ || !gU)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    CHK(need_state(c, true, true));
    HIPCHK(hipSetDevice(c->device));
    long n = c->nsamples;
    size_t npix = 12UL * c->nside_sky * c->nside_sky;
    CHK(c->cot.ensure(2 * (size_t)n + 1));
    if (n) {
        if (g_a)
            HIPCHK(hipMemcpyAsync(c->cot.p, g_a, n * 4, hipMemcpyHostToDevice, c->stream));
        if (g_b)
            HIPCHK(hipMemcpyAsync(c->cot.p + n, g_b, n * 4, hipMemcpyHostToDevice, c->stream));
    }
    CHK(c->fout.ensure(3 * npix));
    CHK(pisco_dev_adjoint_map(c, c->stream, c->nside_sky, c->nside_beam, c->npix_beam,
                              (const float *)c->beam.p, n, c->ra.p, c->dec.p, c->pa.p,
                              g_a ? c->cot.p : nullptr, g_b ? c->cot.p + n : nullptr,
                              c->fout.p, c->fout.p + npix, c->fout.p + 2 * npix));
    HIPCHK(hipMemcpyAsync(gI, c->fout.p, npix * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(gQ, c->fout.p + npix, npix * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(gU, c->fout.p + 2 * npix, npix * 4, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return PISCO_OK;
}

int pisco_exec_adjoint_beam(pisco_ctx *c, const float *g_a, const float *g_b,
                            float *const ga[4], float *const gb[4])
{
    if (!c || !ga || !gb)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    CHK(need_state(c, true, true));
    HIPCHK(hipSetDevice(c->device));
    long n = c->nsamples;
    long nb = c->npix_beam;
    CHK(c->cot.ensure(2 * (size_t)n + 1));
    if (n) {
        if (g_a)
            HIPCHK(hipMemcpyAsync(c->cot.p, g_a, n * 4, hipMemcpyHostToDevice, c->stream));
        if (g_b)
            HIPCHK(hipMemcpyAsync(c->cot.p + n, g_b, n * 4, hipMemcpyHostToDevice, c->stream));
    }
    /* with the V term, A3 and B3 receive g V: the beam adjoint's A0 and B0
     * columns (g I) of a second pass with V in place of I */
    const bool vpass = c->v_term && c->has_v;
    CHK(c->fout.ensure((vpass ? 16 : 8) * (size_t)nb));
    CHK(pisco_dev_adjoint_beam(c, c->stream, c->nside_sky, c->I.p, c->Q.p, c->U.p,
                               c->nside_beam, nb, n, c->ra.p, c->dec.p, c->pa.p,
                               g_a ? c->cot.p : nullptr, g_b ? c->cot.p + n : nullptr,
                               c->fout.p));
    if (vpass)
        CHK(pisco_dev_adjoint_beam(c, c->stream, c->nside_sky, c->V.p, c->V.p, c->V.p,
                                   c->nside_beam, nb, n, c->ra.p, c->dec.p, c->pa.p,
                                   g_a ? c->cot.p : nullptr, g_b ? c->cot.p + n : nullptr,
                                   c->fout.p + 8 * nb));
    /* [nb][8] -> 8 maps on the host */
    float *h = nullptr;
    HIPCHK(hipHostMalloc((void **)&h, sizeof(float) * (vpass ? 16 : 8) * (size_t)nb, 0));
    hipError_t e = hipMemcpyAsync(h, c->fout.p, sizeof(float) * (vpass ? 16 : 8) * (size_t)nb,
                                  hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) {
        static const int ia[4] = {0, 1, 2, 6}, ib[4] = {3, 4, 5, 7};
        for (long i = 0; i < nb; i++)
            for (int k = 0; k < 4; k++) {
                if (ga[k]) ga[k][i] = h[8 * i + ia[k]];
                if (gb[k]) gb[k][i] = h[8 * i + ib[k]];
            }
        if (vpass) {
            const float *hv = h + 8 * nb;
            for (long i = 0; i < nb; i++) {
                if (ga[3]) ga[3][i] = hv[8 * i + 0];
                if (gb[3]) gb[3][i] = hv[8 * i + 3];
            }
        }
    }
    (void)hipHostFree(h);
    HIPCHK(e);
    return PISCO_OK;
}

/* ---- device-resident entry points -------------------------------------- */

int pisco_dev_pack_beam(pisco_ctx *c, void *stream, long npix, const float *const maps[8],
                        float *beam8)
{
    if (!c || !maps || !beam8)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    for (int k = 0; k < 8; k++)
        if (k != 3 && k != 7 && !maps[k])
            return fail(PISCO_ERR_INVALID, "beam map %d is NULL", k);
    if (npix < 1)
        return fail(PISCO_ERR_INVALID, "npix must be positive");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(launch_pack_beam(npix, maps, beam8, pick(c, stream)));
    return PISCO_OK;
}

int pisco_dev_forward(pisco_ctx *c, void *stream, int nside_sky, const float *I,
                      const float *Q, const float *U, int nside_beam, long npix_beam,
                      const float *beam8, long nsamples, const float *ra, const float *dec,
                      const float *pa, float *tod_a, float *tod_b)
{
    if (!c || !I || !Q || !U || !beam8 || (nsamples > 0 && (!ra || !dec || !pa)))
        return fail(PISCO_ERR_INVALID, "NULL argument");
    HIPCHK(hipSetDevice(c->device));
    PairParams p;
    CHK(setup_pair(c, nside_sky, I, Q, U, nside_beam, npix_beam, beam8, nsamples, ra, dec,
                   pa, p));
    p.tod_a = tod_a;
    p.tod_b = tod_b;
    return run_pair(c, OP_FWD, p, pick(c, stream));
}

int pisco_dev_forward_v(pisco_ctx *c, void *stream, int nside_sky, const float *I,
                        const float *Q, const float *U, const float *V, int nside_beam,
                        long npix_beam, const float *beam8, long nsamples, const float *ra,
                        const float *dec, const float *pa, float *tod_a, float *tod_b)
{
    if (!c || !I || !Q || !U || !V || !beam8 || (nsamples > 0 && (!ra || !dec || !pa)))
        return fail(PISCO_ERR_INVALID, "NULL argument");
    HIPCHK(hipSetDevice(c->device));
    PairParams p;
    CHK(setup_pair(c, nside_sky, I, Q, U, nside_beam, npix_beam, beam8, nsamples, ra, dec,
                   pa, p));
    p.tod_a = tod_a;
    p.tod_b = tod_b;
    return run_pair(c, OP_FWD_V, p, pick(c, stream), false, V);
}

int pisco_dev_adjoint_map(pisco_ctx *c, void *stream, int nside_sky, int nside_beam,
                          long npix_beam, const float *beam8, long nsamples,
                          const float *ra, const float *dec, const float *pa,
                          const float *g_a, const float *g_b, float *gI, float *gQ,
                          float *gU)
{
    if (!c || !beam8 || !gI || !gQ || !gU || (nsamples > 0 && (!ra || !dec || !pa)))
        return fail(PISCO_ERR_INVALID, "NULL argument");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = pick(c, stream);
    PairParams p;
    CHK(setup_pair(c, nside_sky, nullptr, nullptr, nullptr, nside_beam, npix_beam, beam8,
                   nsamples, ra, dec, pa, p));
    long npix = 12L * nside_sky * nside_sky;
    CHK(c->acc.ensure(3 * (size_t)npix));
    HIPCHK(launch_zero_f64(c->acc.p, 3 * npix, st));
    p.g_a = g_a;
    p.g_b = g_b;
    p.gI = c->acc.p;
    p.gQ = c->acc.p + npix;
    p.gU = c->acc.p + 2 * npix;
    if (c->adjmap_det > 0 && c->beam_interp)
        return fail(PISCO_ERR_INVALID,
                    "the deterministic sky adjoint is the nearest-pixel one: "
                    "beam_interpolation must be 0");
    const bool det = det_single(c);
    const double *scale = nullptr;
    if (det) {
        const float *bl[1] = {beam8};
        const long nl[1] = {npix_beam};
        CHK(det_scale(c, bl, nl, 1, g_a, g_b, nsamples, nsamples, st, &scale));
    }
    if (g_a || g_b) {
        if ((c->adjmap_chunk != 0 || det) && !c->beam_interp) {
            const int nb = c->adjmap_blocks;   /* 0: one block per window */
            const int chunk = adj_chunk(c, nside_sky, p.rho_max, adj_lanes(c, p));
            const int cap = fit_cap(c->adjmap_cap, c->adjmap_rw, 0, det);
            if (c->adjmap_support) {
                CHK(c->support.ensure(1));
                HIPCHK(launch_beam_support(beam8, npix_beam, c->support.p, st));
                p.beam_last = c->support.p;
            }
            hipError_t e = launch_adjmap_lds(p, chunk, cap, c->adjmap_rw, nb, st, nullptr, 0.0, 1,
                                             scale, adj_lanes(c, p));
            if (e != hipSuccess) {
                char d[256];
                adjmap_describe(d, sizeof d, cap, c->adjmap_rw);
                return fail(PISCO_ERR_DEVICE, "sky adjoint launch failed: %s (%s)",
                            hipGetErrorString(e), d);
            }
        } else {
            CHK(run_pair(c, OP_ADJ_MAP, p, st));
        }
    }
    if (det) {
        HIPCHK(launch_fix_to_f32(p.gI, scale, gI, npix, st));
        HIPCHK(launch_fix_to_f32(p.gQ, scale, gQ, npix, st));
        HIPCHK(launch_fix_to_f32(p.gU, scale, gU, npix, st));
    } else {
        HIPCHK(launch_f64_to_f32(p.gI, gI, npix, st));
        HIPCHK(launch_f64_to_f32(p.gQ, gQ, npix, st));
        HIPCHK(launch_f64_to_f32(p.gU, gU, npix, st));
    }
    return PISCO_OK;
}

int pisco_dev_adjoint_beam(pisco_ctx *c, void *stream, int nside_sky, const float *I,
                           const float *Q, const float *U, int nside_beam, long npix_beam,
                           long nsamples, const float *ra, const float *dec,
                           const float *pa, const float *g_a, const float *g_b,
                           float *gbeam8)
{
    if (!c || !I || !Q || !U || !gbeam8 || (nsamples > 0 && (!ra || !dec || !pa)))
        return fail(PISCO_ERR_INVALID, "NULL argument");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = pick(c, stream);
    PairParams p;
    /* beam values are not read by the beam adjoint; pass a valid pointer */
    CHK(setup_pair(c, nside_sky, I, Q, U, nside_beam, npix_beam, gbeam8, nsamples, ra,
                   dec, pa, p));
    CHK(c->acc.ensure(8 * (size_t)npix_beam));
    HIPCHK(launch_zero_f64(c->acc.p, 8 * npix_beam, st));
    p.g_a = g_a;
    p.g_b = g_b;
    p.gbeam = c->acc.p;
    if ((g_a || g_b) && c->adjbeam_bands && !c->beam_interp) {
        const long npix = p.hs.npix;
        CHK(c->sky4.ensure(4 * (size_t)npix));
        HIPCHK(launch_pack_sky(I, Q, U, npix, (float4 *)c->sky4.p, st));
        p.sky4 = (const float4 *)c->sky4.p;
        RecPlan pl;
        hipError_t e;
        /* the records' memory: adjbeam_pool_mb, at most a quarter of the free
         * device memory (beside what the context already holds for them) */
        size_t pool = (size_t)c->adjbeam_pool_mb << 20;
        size_t mfree = 0, mtotal = 0;
        if (hipMemGetInfo(&mfree, &mtotal) == hipSuccess) {
            const size_t held = c->recs.cap * sizeof(float4) + c->recidx.cap * sizeof(long);
            pool = std::min(pool, std::max<size_t>((mfree + held) / 4, (size_t)64 << 20));
        }
        if (c->adjbeam_records && adjbeam_rec_plan(p, c->adjbeam_rows, pool, c->ncu, pl)) {
            CHK(c->recs.ensure(pl.nrec));
            CHK(c->recidx.ensure(pl.nidx));
            CHK(c->recvmax.ensure(2));
            CHK(c->recpart.ensure(pl.npart));
            e = launch_adjbeam_records(p, pl, c->recs.p, c->recidx.p, c->recvmax.p,
                                       c->recpart.p, st);
        } else {
            e = launch_adjbeam_bands(p, c->adjbeam_chunk, c->adjbeam_rows, c->adjmap_blocks,
                                     c->ncu, c->adjbeam_lanes, st);
        }
        if (e != hipSuccess)
            return fail(PISCO_ERR_DEVICE, "beam adjoint launch failed: %s",
                        hipGetErrorString(e));
    } else if (g_a || g_b) {
        CHK(run_pair(c, OP_ADJ_BEAM, p, st));
    }
    HIPCHK(launch_f64_to_f32(c->acc.p, gbeam8, 8 * npix_beam, st));
    return PISCO_OK;
}

int pisco_dev_pix2ang(pisco_ctx *c, void *stream, int nside, long n, const int *pix,
                      double *theta, double *phi)
{
    if (!c || !valid_nside(nside) || n < 0)
        return fail(PISCO_ERR_INVALID, "bad argument");
    if (n == 0)
        return PISCO_OK;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(launch_pix2ang(hpx_make(nside), n, pix, theta, phi, pick(c, stream)));
    return PISCO_OK;
}

int pisco_dev_ang2pix(pisco_ctx *c, void *stream, int nside, long n, const double *theta,
                      const double *phi, int *pix)
{
    if (!c || !valid_nside(nside) || n < 0)
        return fail(PISCO_ERR_INVALID, "bad argument");
    if (n == 0)
        return PISCO_OK;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(launch_ang2pix(hpx_make(nside), n, theta, phi, pix, pick(c, stream)));
    return PISCO_OK;
}

int pisco_dev_get_interpol(pisco_ctx *c, void *stream, int nside, long n,
                           const double *theta, const double *phi, int *pix, double *wgt)
{
    if (!c || !valid_nside(nside) || n < 0)
        return fail(PISCO_ERR_INVALID, "bad argument");
    if (n == 0)
        return PISCO_OK;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(launch_interpol(hpx_make(nside), n, theta, phi, pix, wgt, pick(c, stream)));
    return PISCO_OK;
}

int pisco_dev_query_disc(pisco_ctx *c, void *stream, int nside, long n, const float *ra,
                         const float *dec, double radius, const long *offsets, int *counts,
                         int *pixels)
{
    if (!c || !valid_nside(nside) || n < 0 || !(radius >= 0))
        return fail(PISCO_ERR_INVALID, "bad argument");
    if (pixels && !offsets)
        return fail(PISCO_ERR_INVALID, "pixels requires offsets");
    if (n == 0)
        return PISCO_OK;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(launch_disc(hpx_make(nside), n, ra, dec, radius, cos_radius(radius), offsets,
                       counts, pixels, pick(c, stream)));
    return PISCO_OK;
}

int pisco_dev_match(pisco_ctx *c, void *stream, int nside_sky, int nside_beam, long n,
                    const float *ra, const float *dec, const float *pa, const long *offsets,
                    const int *pixels, int *bpix, double *chi)
{
    return pisco_dev_match_mode(c, stream, 0, nside_sky, nside_beam, n, ra, dec, pa, offsets,
                                pixels, bpix, chi);
}

int pisco_dev_match_mode(pisco_ctx *c, void *stream, int mode, int nside_sky, int nside_beam,
                         long n, const float *ra, const float *dec, const float *pa,
                         const long *offsets, const int *pixels, int *bpix, double *chi)
{
    if (!c || !valid_nside(nside_sky) || !valid_nside(nside_beam) || n < 0)
        return fail(PISCO_ERR_INVALID, "bad argument");
    if (n == 0)
        return PISCO_OK;
    HIPCHK(hipSetDevice(c->device));
    const RingRec *rt;
    CHK(ring_table(c, nside_sky, &rt));
    if (mode < 0 || mode > 5)
        return fail(PISCO_ERR_INVALID, "match mode must be 0 to 5");
    HIPCHK(launch_match(hpx_make(nside_sky), rt, hpx_make(nside_beam), mode, n, ra, dec, pa,
                        offsets, pixels, bpix, chi, pick(c, stream)));
    return PISCO_OK;
}

int pisco_dev_libm(pisco_ctx *c, void *stream, int op, long n, const double *x,
                   const double *y, double *out)
{
    if (!c || op < 0 || op > 8 || n < 0)
        return fail(PISCO_ERR_INVALID, "bad argument");
    if (n == 0)
        return PISCO_OK;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(launch_libm(op, n, x, y, out, pick(c, stream)));
    return PISCO_OK;
}

} /* extern "C" */

/* ---- batched detector pairs ---------------------------------------------- */

struct pisco_batch {
    int device = 0;
    int npairs = 0;
    int npad = 0;                 /* table length: npairs rounded up to 4 */
    BatchPair *table = nullptr;   /* device, sorted by rho_max descending */
    std::vector<BatchPair> host;  /* same, on the host */
    double rho_support = 0.0;     /* union over pairs of the non-zero beam support */
};

int pisco_batch_create(pisco_ctx *c, int npairs, const pisco_batch_pair *pairs,
                       pisco_batch **out)
{
    if (!c || !pairs || !out || npairs < 1)
        return fail(PISCO_ERR_INVALID, "bad argument");
    *out = nullptr;
    std::vector<BatchPair> t((size_t)npairs);
    for (int k = 0; k < npairs; k++) {
        const pisco_batch_pair &q = pairs[k];
        if (!valid_nside(q.nside_beam))
            return fail(PISCO_ERR_INVALID, "pair %d: NSIDE out of range", k);
        long npb_max = 12L * q.nside_beam * q.nside_beam;
        if (q.npix_beam < 1 || q.npix_beam >= npb_max || !q.beam8)
            return fail(PISCO_ERR_INVALID, "pair %d: bad beam", k);
        BatchPair &b = t[(size_t)k];
        memset(&b, 0, sizeof b);
        b.beam = q.beam8;
        b.npix = (int)q.npix_beam;
        b.nside = q.nside_beam;
        FastConsts fc = fast_consts(hpx_make(q.nside_beam));
        b.nside_f = fc.nside;
        b.marg_k = fc.marg_k;
        b.rho_max = pisco_beam_rho_max(q.nside_beam, q.npix_beam);
        b.cos_rmax = cos_radius(b.rho_max);
        const double sh = pl_sin(0.5 * b.rho_max);
        b.omc_cut = (float)(2.0 * sh * sh);   /* 1 - cos(rho_max) without cancellation */
        b.pa_off = q.pa_offset;
        b.out = k;
    }
    std::stable_sort(t.begin(), t.end(),
                     [](const BatchPair &a, const BatchPair &b) { return a.rho_max > b.rho_max; });
    const int npad = (npairs + 3) / 4 * 4;
    while ((int)t.size() < npad) {   /* inert pairs: never members, no output */
        BatchPair b = t[0];
        b.npix = 0;
        b.omc_cut = -1.0f;
        b.rho_max = 0.0;
        b.cos_rmax = 1.0;
        b.out = -1;
        t.push_back(b);
    }
    HIPCHK(hipSetDevice(c->device));
    BatchPair *d = nullptr;
    HIPCHK(hipMalloc((void **)&d, sizeof(BatchPair) * t.size()));
    hipError_t e = hipMemcpy(d, t.data(), sizeof(BatchPair) * t.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(PISCO_ERR_DEVICE, "hipMemcpy failed: %s", hipGetErrorString(e));
    }
    /* union of the pairs' non-zero beam supports (sizes the sky adjoint's
     * LDS windows; see k_adjmap_lds): theta(ring of the last non-zero row + 2)
     * capped at the pair's rho_max */
    double support = 0.0;
    {
        int *last = nullptr;
        e = hipMalloc((void **)&last, sizeof(int) * (size_t)npairs);
        std::vector<int> hl((size_t)npairs, -1);
        for (int k = 0; k < npairs && e == hipSuccess; k++)
            e = launch_beam_support(pairs[k].beam8, pairs[k].npix_beam, last + k, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(hl.data(), last, sizeof(int) * (size_t)npairs,
                               hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        if (last)
            (void)hipFree(last);
        if (e != hipSuccess) {
            (void)hipFree(d);
            return fail(PISCO_ERR_DEVICE, "beam support: %s", hipGetErrorString(e));
        }
        for (int k = 0; k < npairs; k++) {
            const Hpx hb = hpx_make(pairs[k].nside_beam);
            const double rmax = pisco_beam_rho_max(pairs[k].nside_beam, pairs[k].npix_beam);
            const int rq = (hl[(size_t)k] < 0 ? 0 : pix2ring(hb, hl[(size_t)k])) + 2;
            double th = rq < 4 * hb.nside ? ring_theta(hb, rq) : rmax;
            support = std::max(support, std::min(th, rmax));
        }
    }
    pisco_batch *bt = new pisco_batch();
    bt->device = c->device;
    bt->npairs = npairs;
    bt->npad = npad;
    bt->table = d;
    bt->host = std::move(t);
    bt->rho_support = support;
    *out = bt;
    return PISCO_OK;
}

void pisco_batch_destroy(pisco_batch *b)
{
    if (!b)
        return;
    (void)hipSetDevice(b->device);
    (void)hipFree(b->table);
    delete b;
}

namespace {

int setup_batch(pisco_ctx *c, const pisco_batch *b, int nside_sky, const float *I,
                const float *Q, const float *U, long nsamples, const float *ra,
                const float *dec, const float *pa, BatchParams &p)
{
    const RingRec *rt = nullptr;
    CHK(ring_table(c, nside_sky, &rt));
    memset(&p, 0, sizeof p);
    p.hs = hpx_make(nside_sky);
    p.rings = rt;
    p.I = I;
    p.Q = Q;
    p.U = U;
    p.pairs = b->table;
    p.npairs = b->npad;
    p.ngroups = b->npad;
    p.fc = fast_consts(hpx_make(1));     /* only the rho > 1e-6 band is read */
    p.fc64 = fast_consts64(hpx_make(1)); /* only the crho thresholds are read */
    p.ra = ra;
    p.dec = dec;
    p.pa = pa;
    p.nsamples = nsamples;
    p.s_begin = 0;
    p.s_end = nsamples;
    return PISCO_OK;
}

/* The pair ranges of a batched launch: starts of consecutive ranges of the
 * pair table (sorted by disc radius, largest first), each a multiple of 4
 * pairs, ending with npad.  The tuning value: -1 one range, k > 0 ranges of k
 * pairs; 0 (auto): ranges whose packed beams (32 B per row) add up to at most
 * one XCD's 4 MiB L2, with at least 4 pairs each.  On 32 pairs with beams of
 * 0.05-1.1 MB, ranges of 8 pairs took the forward + sky adjoint from 134.5 to
 * 109.1 ms, almost all of it in the adjoint (ranges of 4: 111.1 ms; of 16:
 * 122.6 ms). */
std::vector<int> batch_ranges(const pisco_ctx *c, const pisco_batch *b)
{
    std::vector<int> cut{0};
    if (c->batch_pair_range > 0) {
        for (int r = c->batch_pair_range; r < b->npad; r += c->batch_pair_range)
            cut.push_back(r);
    } else if (c->batch_pair_range == 0) {
        double bytes = 0.0;
        int start = 0;
        for (int k = 0; k < b->npad; k += 4) {
            double blk = 0.0;
            for (int j = k; j < k + 4 && j < b->npad; j++)
                if (b->host[j].out >= 0)
                    blk += 32.0 * (double)b->host[j].npix;
            if (k > start && bytes + blk > 4.0 * 1024 * 1024) {
                cut.push_back(k);
                start = k;
                bytes = 0.0;
            }
            bytes += blk;
        }
    }
    cut.push_back(b->npad);
    return cut;
}

/* The deterministic mode's scale for a batch: every pair's beam, the
 * cotangents of all pairs (2 npairs nsamples values), at most npairs
 * nsamples contributions per pixel. */
int batch_det_scale(pisco_ctx *c, hipStream_t st, const pisco_batch *b, long nsamples,
                    const float *g, const double **scale)
{
    std::vector<const float *> bl;
    std::vector<long> nl;
    for (int k = 0; k < b->npad; k++) {
        bl.push_back(b->host[k].beam);
        nl.push_back(b->host[k].out < 0 ? 0 : b->host[k].npix);
    }
    return det_scale(c, bl.data(), nl.data(), (int)bl.size(), g, nullptr,
                     2L * b->npairs * nsamples, (long)b->npairs * nsamples, st, scale);
}

/* Whether a batched launch runs pair by pair (the single-pair kernels, one
 * pair after another) instead of all pairs per chunk of samples: the tuning
 * value, or (auto) when a pair's packed beam (32 B per row) exceeds one XCD's
 * 4 MiB L2.  Pairs interleaved per chunk gather from every beam at once; for
 * 5-degree beams at NSIDE 4096 (12.2 MB each) that took 2.4x (forward) and
 * 2.0x (sky adjoint) the time of running the pairs one after another. */
bool pair_major(const pisco_ctx *c, const pisco_batch *b)
{
    if (c->batch_pair_major >= 0)
        return c->batch_pair_major == 1;
    for (const BatchPair &q : b->host)
        if (q.out >= 0 && 32.0 * (double)q.npix > 4.0 * 1024 * 1024)
            return true;
    return false;
}

/* The forward of every pair of a batch, pair by pair, with the single-pair
 * kernels: pair q reads its pointing at offset q.out * ptg_stride (0: one
 * pointing shared by all pairs), with pa + q.pa_off rounded to float as the
 * batched kernels form it, and writes TOD rows 2 q.out (a) and 2 q.out + 1
 * (b). */
int batch_forward_pairwise(pisco_ctx *c, hipStream_t st, const pisco_batch *b, int nside_sky,
                           const float *I, const float *Q, const float *U, long nsamples,
                           const float *ra, const float *dec, const float *pa, long ptg_stride,
                           float *tod)
{
    CHK(c->cot.ensure(std::max<long>(nsamples, 1)));
    const long npix = 12L * nside_sky * nside_sky;
    CHK(c->sky4.ensure(4 * (size_t)npix));   /* packed once for all pairs */
    HIPCHK(launch_pack_sky(I, Q, U, npix, (float4 *)c->sky4.p, st));
    for (const BatchPair &q : b->host) {
        if (q.out < 0 || nsamples == 0)
            continue;
        const long po = (long)q.out * ptg_stride;
        HIPCHK(launch_add_offset(pa + po, q.pa_off, c->cot.p, nsamples, st));
        PairParams p;
        CHK(setup_pair(c, nside_sky, I, Q, U, q.nside, q.npix, q.beam, nsamples, ra + po,
                       dec + po, c->cot.p, p));
        p.tod_a = tod + 2L * q.out * nsamples;
        p.tod_b = p.tod_a + nsamples;
        CHK(run_pair(c, OP_FWD, p, st, true));
    }
    return PISCO_OK;
}

/* The sky adjoint of every pair of a batch, pair by pair (single-pair LDS
 * adjoints), all adding into the context's FP64 gradient (cleared by the
 * caller); pointing as in batch_forward_pairwise; cotangent rows 2 q.out and
 * 2 q.out + 1 of g; scale: the deterministic mode's scale or NULL. */
int batch_adjoint_pairwise(pisco_ctx *c, hipStream_t st, const pisco_batch *b, int nside_sky,
                           long nsamples, const float *ra, const float *dec, const float *pa,
                           long ptg_stride, const float *g, const double *scale)
{
    const long npix = 12L * nside_sky * nside_sky;
    const int nb = c->adjmap_blocks;   /* 0: one block per window */
    const int cap = fit_cap(c->adjmap_cap, c->adjmap_rw, 0, scale != nullptr);
    CHK(c->cot.ensure(std::max<long>(nsamples, 1)));
    CHK(c->support.ensure(1));
    for (const BatchPair &q : b->host) {
        if (q.out < 0 || nsamples == 0)
            continue;
        const long po = (long)q.out * ptg_stride;
        HIPCHK(launch_add_offset(pa + po, q.pa_off, c->cot.p, nsamples, st));
        PairParams p;
        CHK(setup_pair(c, nside_sky, nullptr, nullptr, nullptr, q.nside, q.npix, q.beam,
                       nsamples, ra + po, dec + po, c->cot.p, p));
        p.g_a = g + 2L * q.out * nsamples;
        p.g_b = p.g_a + nsamples;
        p.gI = c->acc.p;
        p.gQ = c->acc.p + npix;
        p.gU = c->acc.p + 2 * npix;
        if (c->adjmap_support) {
            HIPCHK(launch_beam_support(q.beam, q.npix, c->support.p, st));
            p.beam_last = c->support.p;
        }
        const int lanes = adj_lanes(c, p);
        hipError_t e = launch_adjmap_lds(p, adj_chunk(c, nside_sky, p.rho_max, lanes), cap,
                                         c->adjmap_rw, nb, st, nullptr, 0.0, 1, scale, lanes);
        if (e != hipSuccess)
            return fail(PISCO_ERR_DEVICE, "per-pair sky adjoint launch failed: %s",
                        hipGetErrorString(e));
    }
    return PISCO_OK;
}

} // namespace

int pisco_dev_batched_forward(pisco_ctx *c, void *stream, const pisco_batch *b, int nside_sky,
                              const float *I, const float *Q, const float *U, long nsamples,
                              const float *ra, const float *dec, const float *pa, float *tod)
{
    if (!c || !b || !I || !Q || !U || !tod || (nsamples > 0 && (!ra || !dec || !pa)))
        return fail(PISCO_ERR_INVALID, "NULL argument");
    if (!valid_nside(nside_sky) || nsamples < 0)
        return fail(PISCO_ERR_INVALID, "bad argument");
    if (b->device != c->device)
        return fail(PISCO_ERR_INVALID, "batch and context are on different devices");
    CHK(no_interp(c));
    HIPCHK(hipSetDevice(c->device));
    if (pair_major(c, b))
        return batch_forward_pairwise(c, pick(c, stream), b, nside_sky, I, Q, U, nsamples, ra,
                                      dec, pa, 0, tod);
    BatchParams p;
    CHK(setup_batch(c, b, nside_sky, I, Q, U, nsamples, ra, dec, pa, p));
    const int G = c->batch_group;
    p.tod = tod;
    hipStream_t st = pick(c, stream);
    CHK(c->sky4.ensure(4 * (size_t)p.hs.npix));
    HIPCHK(launch_pack_sky(I, Q, U, p.hs.npix, (float4 *)c->sky4.p, st));
    p.sky4 = (const float4 *)c->sky4.p;
    long chunk = (c->cfg.ptgPerConv > 0) ? c->cfg.ptgPerConv : nsamples;
    chunk = std::max<long>(chunk, std::min<long>(nsamples, 1L << 20));
    const std::vector<int> cut = batch_ranges(c, b);
    for (size_t r = 0; r + 1 < cut.size(); r++) {   /* ranges of the sorted pair table */
        const int r0 = cut[r], r1 = cut[r + 1];
        p.pairs = b->table + r0;
        p.npairs = r1 - r0;
        p.ngroups = (r1 - r0) / G;
        for (long s0 = 0; s0 < nsamples; s0 += chunk) {
            p.s_begin = s0;
            p.s_end = std::min(nsamples, s0 + chunk);
            HIPCHK(launch_batch_forward(G, p, st));
        }
    }
    return PISCO_OK;
}

int pisco_dev_batched_forward_pp(pisco_ctx *c, void *stream, const pisco_batch *b,
                                 int nside_sky, const float *I, const float *Q, const float *U,
                                 long nsamples, const float *ra, const float *dec,
                                 const float *pa, float *tod)
{
    if (!c || !b || !I || !Q || !U || !tod || (nsamples > 0 && (!ra || !dec || !pa)))
        return fail(PISCO_ERR_INVALID, "NULL argument");
    if (!valid_nside(nside_sky) || nsamples < 0)
        return fail(PISCO_ERR_INVALID, "bad argument");
    if (b->device != c->device)
        return fail(PISCO_ERR_INVALID, "batch and context are on different devices");
    CHK(no_interp(c));
    HIPCHK(hipSetDevice(c->device));
    if (pair_major(c, b))
        return batch_forward_pairwise(c, pick(c, stream), b, nside_sky, I, Q, U, nsamples, ra,
                                      dec, pa, nsamples, tod);
    BatchParams p;
    CHK(setup_batch(c, b, nside_sky, I, Q, U, nsamples, ra, dec, pa, p));
    p.ptg_stride = nsamples;
    p.tod = tod;
    hipStream_t st = pick(c, stream);
    CHK(c->sky4.ensure(4 * (size_t)p.hs.npix));
    HIPCHK(launch_pack_sky(I, Q, U, p.hs.npix, (float4 *)c->sky4.p, st));
    p.sky4 = (const float4 *)c->sky4.p;
    long chunk = (c->cfg.ptgPerConv > 0) ? c->cfg.ptgPerConv : nsamples;
    chunk = std::max<long>(chunk, std::min<long>(nsamples, 1L << 20));
    for (long s0 = 0; s0 < nsamples; s0 += chunk) {
        p.s_begin = s0;
        p.s_end = std::min(nsamples, s0 + chunk);
        HIPCHK(launch_batch_forward(1, p, st));   /* per-pair items: no shared geometry */
    }
    return PISCO_OK;
}

int pisco_dev_batched_adjoint_map_pp(pisco_ctx *c, void *stream, const pisco_batch *b,
                                     int nside_sky, long nsamples, const float *ra,
                                     const float *dec, const float *pa, const float *g,
                                     float *gI, float *gQ, float *gU)
{
    if (!c || !b || !g || !gI || !gQ || !gU || (nsamples > 0 && (!ra || !dec || !pa)))
        return fail(PISCO_ERR_INVALID, "NULL argument");
    if (!valid_nside(nside_sky) || nsamples < 0)
        return fail(PISCO_ERR_INVALID, "bad argument");
    if (b->device != c->device)
        return fail(PISCO_ERR_INVALID, "batch and context are on different devices");
    CHK(no_interp(c));
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = pick(c, stream);
    const long npix = 12L * nside_sky * nside_sky;
    CHK(c->acc.ensure(3 * (size_t)npix));
    CHK(c->cot.ensure(std::max<long>(nsamples, 1)));
    CHK(c->support.ensure(1));
    HIPCHK(launch_zero_f64(c->acc.p, 3 * npix, st));
    const bool det = det_single(c);
    const double *scale = nullptr;
    if (det)
        CHK(batch_det_scale(c, st, b, nsamples, g, &scale));
    /* one single-pair sky adjoint per pair, each with its own pointing and
     * pa + pa_offset rounded to float (what the forward uses), all adding into
     * one FP64 gradient */
    CHK(batch_adjoint_pairwise(c, st, b, nside_sky, nsamples, ra, dec, pa, nsamples, g, scale));
    float *out[3] = {gI, gQ, gU};
    for (int k = 0; k < 3; k++) {
        if (det)
            HIPCHK(launch_fix_to_f32(c->acc.p + k * npix, scale, out[k], npix, st));
        else
            HIPCHK(launch_f64_to_f32(c->acc.p + k * npix, out[k], npix, st));
    }
    return PISCO_OK;
}

int pisco_dev_batched_adjoint_map(pisco_ctx *c, void *stream, const pisco_batch *b,
                                  int nside_sky, long nsamples, const float *ra,
                                  const float *dec, const float *pa, const float *g, float *gI,
                                  float *gQ, float *gU)
{
    if (!c || !b || !g || !gI || !gQ || !gU || (nsamples > 0 && (!ra || !dec || !pa)))
        return fail(PISCO_ERR_INVALID, "NULL argument");
    if (!valid_nside(nside_sky) || nsamples < 0)
        return fail(PISCO_ERR_INVALID, "bad argument");
    if (b->device != c->device)
        return fail(PISCO_ERR_INVALID, "batch and context are on different devices");
    CHK(no_interp(c));
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = pick(c, stream);
    BatchParams bp;
    CHK(setup_batch(c, b, nside_sky, nullptr, nullptr, nullptr, nsamples, ra, dec, pa, bp));
    const long npix = 12L * nside_sky * nside_sky;
    CHK(c->acc.ensure(3 * (size_t)npix));
    HIPCHK(launch_zero_f64(c->acc.p, 3 * npix, st));
    float *out[3] = {gI, gQ, gU};
    if (pair_major(c, b)) {
        const double *scale = nullptr;
        if (det_single(c))
            CHK(batch_det_scale(c, st, b, nsamples, g, &scale));
        CHK(batch_adjoint_pairwise(c, st, b, nside_sky, nsamples, ra, dec, pa, 0, g, scale));
        for (int k = 0; k < 3; k++) {
            if (scale)
                HIPCHK(launch_fix_to_f32(c->acc.p + k * npix, scale, out[k], npix, st));
            else
                HIPCHK(launch_f64_to_f32(c->acc.p + k * npix, out[k], npix, st));
        }
        return PISCO_OK;
    }
    bp.g = g;
    bp.gI = c->acc.p;
    bp.gQ = c->acc.p + npix;
    bp.gU = c->acc.p + 2 * npix;
    const int nb = c->adjmap_blocks;   /* 0: one block per window */
    const int chunk = c->adjmap_chunk > 0 ? c->adjmap_chunk : 16;
    const bool det = det_single(c);
    const int grp = c->batch_adj_group;
    /* the window capacity that fits next to the group's per-wave tables */
    const int cap = fit_cap(c->adjmap_cap, c->adjmap_rw, grp, det);
    const double *scale = nullptr;
    if (det)
        CHK(batch_det_scale(c, st, b, nsamples, g, &scale));
    const std::vector<int> cut = batch_ranges(c, b);
    for (size_t r = 0; r + 1 < cut.size(); r++) {   /* ranges of the sorted pair table */
        const int r0 = cut[r], r1 = cut[r + 1];
        bp.pairs = b->table + r0;
        bp.npairs = r1 - r0;
        bp.ngroups = (r1 - r0) / grp;
        /* the range's union disc (its first, largest pair) drives the window
         * passes */
        const BatchPair &q0 = b->host[r0];
        PairParams p;
        memset(&p, 0, sizeof p);
        p.hs = bp.hs;
        p.rings = bp.rings;
        p.hb = hpx_make(q0.nside);
        p.npix_beam = q0.npix;
        p.rho_max = q0.rho_max;
        p.cos_rmax = q0.cos_rmax;
        p.ra = ra;
        p.dec = dec;
        p.pa = pa;
        p.gI = bp.gI;
        p.gQ = bp.gQ;
        p.gU = bp.gU;
        p.s_begin = 0;
        p.s_end = nsamples;
        hipError_t e = launch_adjmap_lds(p, chunk, cap, c->adjmap_rw, nb, st, &bp,
                                         std::min(b->rho_support, q0.rho_max), grp, scale);
        if (e != hipSuccess)
            return fail(PISCO_ERR_DEVICE, "batched sky adjoint launch failed: %s",
                        hipGetErrorString(e));
    }
    for (int k = 0; k < 3; k++) {
        if (det)
            HIPCHK(launch_fix_to_f32(c->acc.p + k * npix, scale, out[k], npix, st));
        else
            HIPCHK(launch_f64_to_f32(c->acc.p + k * npix, out[k], npix, st));
    }
    return PISCO_OK;
}

/* ======================================================================== */
/* Binned map-maker (SkyMap)                                                */
/* ======================================================================== */

struct pisco_skymap {
    int device = 0;
    hipStream_t stream = nullptr;
    int nside = 0;
    long npix = 0;
    long chunk = 1L << 24;               /* samples per accumulation pass */
    DevBuf<double> acc;                  /* MM_PLANES x npix */
    DevBuf<int> count, off, tsum;        /* per pixel / per scan tile */
    DevBuf<unsigned> keys;               /* 3 x chunk: pixels and the sort's buffers */
    DevBuf<int> vals;                    /* 2 x chunk */
    DevBuf<double> terms;                /* 2 x chunk: cos 2psi, sin 2psi */
    DevBuf<float> tdata;                 /* chunk */
    DevBuf<int> hist;                    /* 256 x tiles, its scan (+1), scan tiles */
    DevBuf<int> ppix;                    /* pixels of a shared pointing */
    DevBuf<unsigned long long> counters; /* [0] skipped, [1] singular */
    DevBuf<float> stage;                 /* host-API inputs */
    DevBuf<float> out;                   /* host-API outputs */
    DevBuf<int> hits;
};

namespace {

int map_device(pisco_skymap *m)
{
    HIPCHK(hipSetDevice(m->device));
    return PISCO_OK;
}

int map_accumulate(pisco_skymap *m, hipStream_t st, long nsamples, const float *ra,
                   const float *dec, const float *pa, long ptg_stride, int ndet,
                   const float *pol, const float *data)
{
    const long total = nsamples * (long)ndet;
    const long chunk = std::min<long>(m->chunk, total);
    const long nh = 256L * map_tiles((int)chunk);
    CHK(m->keys.ensure(3 * chunk));
    CHK(m->vals.ensure(2 * chunk));
    CHK(m->terms.ensure(2 * chunk));
    CHK(m->tdata.ensure(chunk));
    CHK(m->hist.ensure(2 * nh + 1 + map_tiles((int)nh)));
    MapScratch s;
    for (int k = 0; k < 3; k++)
        s.keys[k] = m->keys.p + k * chunk;
    s.vals[0] = nullptr;
    s.vals[1] = m->vals.p;
    s.vals[2] = m->vals.p + chunk;
    s.tcs = (double2 *)m->terms.p;
    s.td = m->tdata.p;
    s.count = m->count.p;
    s.off = m->off.p;
    s.tsum = m->tsum.p;
    s.hist = m->hist.p;
    s.hoff = m->hist.p + nh;
    s.htsum = m->hist.p + 2 * nh + 1;
    s.skipped = m->counters.p;
    s.ppix = nullptr;
    Hpx h = hpx_make(m->nside);
    if (ptg_stride == 0 && ndet > 1) {
        CHK(m->ppix.ensure(nsamples));
        HIPCHK(launch_map_ptg_pixel(h, ra, dec, nsamples, m->ppix.p, st));
        s.ppix = m->ppix.p;
    }
    for (long t0 = 0; t0 < total; t0 += chunk) {
        MapChunk c;
        c.ra = ra;
        c.dec = dec;
        c.pa = pa;
        c.pol = pol;
        c.data = data;
        c.nsamples = nsamples;
        c.ptg_stride = ptg_stride;
        c.t0 = t0;
        c.len = (int)std::min(chunk, total - t0);
        hipError_t e = launch_map_accumulate(h, c, s, m->acc.p, st);
        if (e != hipSuccess)
            return fail(PISCO_ERR_DEVICE, "map accumulation launch failed: %s",
                        hipGetErrorString(e));
    }
    return PISCO_OK;
}

int map_check(pisco_skymap *m, long nsamples, int ndet)
{
    if (!m)
        return fail(PISCO_ERR_INVALID, "NULL map");
    if (nsamples < 0 || ndet < 0)
        return fail(PISCO_ERR_INVALID, "negative nsamples or ndet");
    return PISCO_OK;
}

} // namespace

extern "C" {

int pisco_skymap_create(int device, int nside, pisco_skymap **out)
{
    if (!out)
        return fail(PISCO_ERR_INVALID, "NULL out");
    *out = nullptr;
    if (!valid_nside(nside))
        return fail(PISCO_ERR_INVALID, "nside %d out of range [1, %d]", nside, kMaxNside);
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(PISCO_ERR_INVALID, "device %d not present (%d devices)", device, ndev);
    pisco_skymap *m = new pisco_skymap;
    m->device = device;
    m->nside = nside;
    m->npix = 12L * nside * nside;
    int rc = map_device(m);
    if (rc == PISCO_OK && hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        rc = fail(PISCO_ERR_DEVICE, "hipStreamCreate failed");
    if (rc == PISCO_OK)
        rc = m->acc.ensure(MM_PLANES * m->npix);
    if (rc == PISCO_OK)
        rc = m->count.ensure(m->npix);
    if (rc == PISCO_OK)
        rc = m->off.ensure(m->npix + 1);
    if (rc == PISCO_OK)
        rc = m->tsum.ensure(map_tiles((int)m->npix));
    if (rc == PISCO_OK)
        rc = m->counters.ensure(2);
    if (rc == PISCO_OK)
        rc = pisco_skymap_reset(m);
    if (rc != PISCO_OK) {
        std::string keep = g_err;
        pisco_skymap_destroy(m);
        g_err = keep;
        return rc;
    }
    *out = m;
    return PISCO_OK;
}

void pisco_skymap_destroy(pisco_skymap *m)
{
    if (!m)
        return;
    (void)hipSetDevice(m->device);
    if (m->stream)
        (void)hipStreamSynchronize(m->stream);
    m->acc.release();
    m->count.release();
    m->off.release();
    m->tsum.release();
    m->keys.release();
    m->vals.release();
    m->terms.release();
    m->tdata.release();
    m->hist.release();
    m->ppix.release();
    m->counters.release();
    m->stage.release();
    m->out.release();
    m->hits.release();
    if (m->stream)
        (void)hipStreamDestroy(m->stream);
    delete m;
}

int pisco_skymap_reset(pisco_skymap *m)
{
    if (!m)
        return fail(PISCO_ERR_INVALID, "NULL map");
    CHK(map_device(m));
    HIPCHK(launch_zero_f64(m->acc.p, MM_PLANES * m->npix, m->stream));
    HIPCHK(launch_fill_u64(m->counters.p, 2, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return PISCO_OK;
}

int pisco_skymap_set_chunk(pisco_skymap *m, long samples)
{
    if (!m)
        return fail(PISCO_ERR_INVALID, "NULL map");
    if (samples < 1 || samples > (1L << 30))
        return fail(PISCO_ERR_INVALID, "chunk must be in [1, 2^30] samples");
    m->chunk = samples;
    return PISCO_OK;
}

int pisco_dev_skymap_accumulate(pisco_skymap *m, void *stream, long nsamples, const float *ra,
                                const float *dec, const float *pa, int shared_pointing,
                                int ndet, const float *pol, const float *data)
{
    CHK(map_check(m, nsamples, ndet));
    if (nsamples == 0 || ndet == 0)
        return PISCO_OK;
    if (!ra || !dec || !pa || !pol || !data)
        return fail(PISCO_ERR_INVALID, "NULL input array");
    CHK(map_device(m));
    return map_accumulate(m, (hipStream_t)stream, nsamples, ra, dec, pa,
                          shared_pointing ? 0 : nsamples, ndet, pol, data);
}

int pisco_skymap_accumulate(pisco_skymap *m, long nsamples, const float *ra, const float *dec,
                            const float *pa, int ndet, const float *pol, const float *data)
{
    CHK(map_check(m, nsamples, ndet));
    if (nsamples == 0 || ndet == 0)
        return PISCO_OK;
    if (!ra || !dec || !pa || !pol || !data)
        return fail(PISCO_ERR_INVALID, "NULL input array");
    CHK(map_device(m));
    const long n = nsamples * (long)ndet;
    CHK(m->stage.ensure(4 * n + ndet));
    float *d = m->stage.p;
    unsigned long long before = 0, after = 0;
    HIPCHK(hipMemcpyAsync(&before, m->counters.p, sizeof before, hipMemcpyDeviceToHost,
                          m->stream));
    HIPCHK(hipMemcpyAsync(d, ra, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(d + n, dec, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(d + 2 * n, pa, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(d + 3 * n, data, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(d + 4 * n, pol, ndet * sizeof(float), hipMemcpyHostToDevice,
                          m->stream));
    CHK(map_accumulate(m, m->stream, nsamples, d, d + n, d + 2 * n, nsamples, ndet, d + 4 * n,
                       d + 3 * n));
    HIPCHK(hipMemcpyAsync(&after, m->counters.p, sizeof after, hipMemcpyDeviceToHost,
                          m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    if (after != before)
        return fail(PISCO_ERR_INVALID,
                    "%llu samples have theta = pi/2 - dec outside [0, pi] or a non-finite ra; "
                    "they were skipped",
                    after - before);
    return PISCO_OK;
}

int pisco_skymap_skipped(pisco_skymap *m, long *count)
{
    if (!m || !count)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    CHK(map_device(m));
    unsigned long long v = 0;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(&v, m->counters.p, sizeof v, hipMemcpyDeviceToHost));
    *count = (long)v;
    return PISCO_OK;
}

int pisco_dev_skymap_solve(pisco_skymap *m, void *stream, float *I, float *Q, float *U,
                           int *hits, long *singular)
{
    if (!m || !I || !Q || !U || !hits)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    CHK(map_device(m));
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(launch_fill_u64(m->counters.p + 1, 1, st));
    HIPCHK(launch_map_solve((int)m->npix, m->acc.p, I, Q, U, hits, m->counters.p + 1, st));
    if (singular) {
        unsigned long long v = 0;
        HIPCHK(hipMemcpyAsync(&v, m->counters.p + 1, sizeof v, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        *singular = (long)v;
    }
    return PISCO_OK;
}

int pisco_skymap_solve(pisco_skymap *m, float *I, float *Q, float *U, int *hits,
                       long *singular)
{
    if (!m || !I || !Q || !U || !hits)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    CHK(map_device(m));
    const long n = m->npix;
    CHK(m->out.ensure(3 * n));
    CHK(m->hits.ensure(n));
    CHK(pisco_dev_skymap_solve(m, m->stream, m->out.p, m->out.p + n, m->out.p + 2 * n, m->hits.p,
                               singular));
    HIPCHK(hipMemcpyAsync(I, m->out.p, n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipMemcpyAsync(Q, m->out.p + n, n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipMemcpyAsync(U, m->out.p + 2 * n, n * sizeof(float), hipMemcpyDeviceToHost,
                          m->stream));
    HIPCHK(hipMemcpyAsync(hits, m->hits.p, n * sizeof(int), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return PISCO_OK;
}

int pisco_skymap_get_normal(pisco_skymap *m, double *AtA, double *AtD)
{
    if (!m || !AtA || !AtD)
        return fail(PISCO_ERR_INVALID, "NULL argument");
    CHK(map_device(m));
    const long n = m->npix;
    std::vector<double> h((size_t)(MM_PLANES * n));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(h.data(), m->acc.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    const double *P[MM_PLANES];
    for (int k = 0; k < MM_PLANES; k++)
        P[k] = h.data() + k * n;
    for (long p = 0; p < n; p++) {
        double *a = AtA + 9 * p;
        a[0] = P[MM_N][p];
        a[1] = a[3] = P[MM_C][p];
        a[2] = a[6] = P[MM_S][p];
        a[4] = P[MM_CC][p];
        a[5] = a[7] = P[MM_CS][p];
        a[8] = P[MM_SS][p];
        AtD[3 * p + 0] = P[MM_D][p];
        AtD[3 * p + 1] = P[MM_CD][p];
        AtD[3 * p + 2] = P[MM_SD][p];
    }
    return PISCO_OK;
}

} /* extern "C" */

/* ======================================================================== */
/* Detector pointing (BPoint::focal_plane_to_sky)                           */
/* ======================================================================== */

namespace {

int fp_check(int ndets, long nsamples, int flag)
{
    if (ndets < 0 || nsamples < 0)
        return fail(PISCO_ERR_INVALID, "negative ndets or nsamples");
    if (flag < 0 || flag > 2)
        return fail(PISCO_ERR_INVALID, "flag must be 0 (az/el), 1 (ra/dec) or 2 (source)");
    return PISCO_OK;
}

} // namespace

extern "C" {

int pisco_dev_focal_plane_to_sky(pisco_ctx *c, void *stream, int ndets, const double *det_dx,
                                 const double *det_dy, const double *det_pol_angle,
                                 long nsamples, const double *x_ac, const double *y_ac,
                                 const double *rot_ac, int flag, double *x_det, double *y_det,
                                 double *rot_det, float *x32, float *y32, float *rot32)
{
    if (!c)
        return fail(PISCO_ERR_INVALID, "NULL context");
    CHK(fp_check(ndets, nsamples, flag));
    if (ndets == 0 || nsamples == 0)
        return PISCO_OK;
    if (!det_dx || !det_dy || !det_pol_angle || !x_ac || !y_ac || !rot_ac)
        return fail(PISCO_ERR_INVALID, "NULL input array");
    HIPCHK(hipSetDevice(c->device));
    FocalPlaneParams q;
    q.ndets = ndets;
    q.dx = det_dx;
    q.dy = det_dy;
    q.pol = det_pol_angle;
    q.n = nsamples;
    q.x_ac = x_ac;
    q.y_ac = y_ac;
    q.rot_ac = rot_ac;
    q.flag = flag;
    q.x = x_det;
    q.y = y_det;
    q.rot = rot_det;
    q.x32 = x32;
    q.y32 = y32;
    q.rot32 = rot32;
    hipError_t e = launch_focal_plane(q, pick(c, stream));
    if (e != hipSuccess)
        return fail(PISCO_ERR_DEVICE, "focal-plane launch failed: %s", hipGetErrorString(e));
    return PISCO_OK;
}

int pisco_focal_plane_to_sky(pisco_ctx *c, int ndets, const double *det_dx,
                             const double *det_dy, const double *det_pol_angle, long nsamples,
                             const double *x_ac, const double *y_ac, const double *rot_ac,
                             int flag, double *x_det, double *y_det, double *rot_det)
{
    if (!c)
        return fail(PISCO_ERR_INVALID, "NULL context");
    CHK(fp_check(ndets, nsamples, flag));
    if (ndets == 0 || nsamples == 0)
        return PISCO_OK;
    if (!det_dx || !det_dy || !det_pol_angle || !x_ac || !y_ac || !rot_ac || !x_det ||
        !y_det || !rot_det)
        return fail(PISCO_ERR_INVALID, "NULL array");
    HIPCHK(hipSetDevice(c->device));
    const long nd = ndets, n = nsamples, no = nd * n;
    DevBuf<double> buf;
    CHK(buf.ensure(3 * nd + 3 * n + 3 * no));
    double *d = buf.p;
    hipStream_t st = c->stream;
    int rc = PISCO_OK;
    const double *src[6] = {det_dx, det_dy, det_pol_angle, x_ac, y_ac, rot_ac};
    const long len[6] = {nd, nd, nd, n, n, n};
    double *dst[6] = {d, d + nd, d + 2 * nd, d + 3 * nd, d + 3 * nd + n, d + 3 * nd + 2 * n};
    double *out = d + 3 * nd + 3 * n;
    for (int k = 0; k < 6 && rc == PISCO_OK; k++)
        if (hipMemcpyAsync(dst[k], src[k], len[k] * sizeof(double), hipMemcpyHostToDevice, st) !=
            hipSuccess)
            rc = fail(PISCO_ERR_DEVICE, "host-to-device copy failed");
    if (rc == PISCO_OK)
        rc = pisco_dev_focal_plane_to_sky(c, st, ndets, dst[0], dst[1], dst[2], nsamples, dst[3],
                                          dst[4], dst[5], flag, out, out + no, out + 2 * no,
                                          nullptr, nullptr, nullptr);
    double *hout[3] = {x_det, y_det, rot_det};
    for (int k = 0; k < 3 && rc == PISCO_OK; k++)
        if (hipMemcpyAsync(hout[k], out + k * no, no * sizeof(double), hipMemcpyDeviceToHost,
                           st) != hipSuccess)
            rc = fail(PISCO_ERR_DEVICE, "device-to-host copy failed");
    if (rc == PISCO_OK && hipStreamSynchronize(st) != hipSuccess)
        rc = fail(PISCO_ERR_DEVICE, "stream synchronisation failed");
    (void)hipStreamSynchronize(st);
    buf.release();
    return rc;
}

} /* extern "C" */
