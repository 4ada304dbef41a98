/*
 * kernels.hpp -- internal interface between the C-ABI layer and the kernels.
 */
#ifndef PISCO_KERNELS_HPP
#define PISCO_KERNELS_HPP

#include <hip/hip_runtime.h>

#include "device_hpx.h"

namespace pisco {

/* OP_FWD_V: the forward with the V terms of the reference's GPU kernel
 * (V A3, V B3, cudaConv2.cu:204,214,217-218); p.sky4 then carries V. */
enum { OP_FWD = 0, OP_ADJ_MAP = 1, OP_ADJ_BEAM = 2, OP_FWD_V = 3 };

/* One launch over samples [s_begin, s_end) of a detector pair a/b that
 * share pointing and beam geometry. */
struct PairParams {
    /* sky: RING map of NSIDE hs.nside, SoA float I,Q,U */
    Hpx hs;
    const RingRec *rings;     /* per-ring constants, index 1..4N-1 */
    const float *I, *Q, *U;
    const float4 *sky4;       /* forward: I, Q, U, V (0 without the V term) per pixel */
    /* beam: first npix_beam RING pixels of a map of NSIDE hb.nside, rows of
     * eight floats A0 A1 A2 B0 B1 B2 A3 B3 (what the CPU forward reads comes
     * first: one 16-byte and one 8-byte load per visit) */
    Hpx hb;
    int npix_beam;
    double rho_max, cos_rmax;
    FastConsts fc;            /* single-precision mapping (sky adjoint, beam adjoint) */
    FastConsts64 fc64;        /* double-precision mapping (forward) */
    const float *beam;
    /* device int: index of the last beam row whose A0 A1 A2 B0 B1 B2 are not
     * all zero (-1: none), or NULL.  Only sizes the sky adjoint's LDS
     * windows; every visit of the rho_max disc is still computed. */
    const int *beam_last;
    /* pointing (float radians; pa counter-clockwise as Scan holds it) */
    const float *ra, *dec, *pa;
    long s_begin, s_end;
    /* forward output */
    float *tod_a, *tod_b;
    /* adjoint inputs (TOD cotangents) and outputs (FP64 accumulators) */
    const float *g_a, *g_b;
    double *gI, *gQ, *gU;     /* [npix sky] each */
    double *gbeam;            /* [npix_beam][8], same row layout as beam */
};

/* One detector pair of a batched launch: its own beam (rows A0 A1 A2 B0 B1
 * B2 A3 B3 of a map of NSIDE nside), its disc radius and the offset its
 * position angle has from the shared pointing.  The device table is sorted by
 * rho_max, largest first. */
struct BatchPair {
    const float *beam;
    int npix;               /* beam pixels */
    int nside;              /* beam NSIDE */
    float nside_f, marg_k;  /* FastConsts of this beam NSIDE */
    float omc_cut;          /* 1 - cos(rho_max) */
    float pa_off;           /* pa of the pair = (float)(pa + pa_off) */
    double rho_max, cos_rmax;
    int out;                /* position of the pair in the caller's order */
    int pad;
};

/* A batched launch over samples [s_begin, s_end) and pairs [0, npairs). */
struct BatchParams {
    Hpx hs;
    const RingRec *rings;
    const float *I, *Q, *U;
    const float4 *sky4;       /* forward: I, Q, U, 0 per pixel */
    const BatchPair *pairs;   /* device, sorted by rho_max descending */
    int npairs, ngroups;      /* ngroups = ceil(npairs / G) */
    FastConsts fc;            /* rho > 1e-6 band (omc_lo, omc_hi) */
    FastConsts64 fc64;        /* crho thresholds of the double-precision form */
    const float *ra, *dec, *pa;
    long ptg_stride;          /* 0: shared pointing; else pair `out` at out * ptg_stride */
    long s_begin, s_end, nsamples;
    float *tod;               /* [out][2][nsamples] */
    const float *g;           /* cotangents, same layout */
    double *gI, *gQ, *gU;
};

/* Binned map-maker.  Samples of a chunk are global indices t in
 * [t0, t0 + len) of the detector-major arrays (t = det * nsamples + s);
 * pointing of sample t is at det * ptg_stride + s (ptg_stride 0: every
 * detector shares one pointing). */
struct MapChunk {
    const float *ra, *dec, *pa;
    const float *pol;         /* [ndet] detector polarisation angles */
    const float *data;        /* [ndet][nsamples] */
    long nsamples, ptg_stride, t0;
    int len;
};
/* FP64 planes of the normal equations, npix each: hits, sum c, sum s,
 * sum cc, sum cs, sum ss (AtA) and sum d, sum cd, sum sd (AtD) */
enum { MM_N, MM_C, MM_S, MM_CC, MM_CS, MM_SS, MM_D, MM_CD, MM_SD, MM_PLANES };
struct MapScratch {
    unsigned *keys[3];        /* [chunk] each: pixels (0) and the sort's two buffers */
    int *vals[3];             /* [chunk] (vals[0] unused) */
    double2 *tcs;             /* [chunk] (cos 2psi, sin 2psi) in sorted order */
    float *td;                /* [chunk] data in sorted order */
    int *count;               /* [npix] */
    int *off;                 /* [npix + 1] */
    int *tsum;                /* [map_tiles(npix)] */
    int *hist, *hoff;         /* [256 * map_tiles(chunk)] (+1 for hoff) */
    int *htsum;               /* [map_tiles(256 * map_tiles(chunk))] */
    const int *ppix;          /* [nsamples] pixels of a shared pointing, or NULL */
    unsigned long long *skipped;
};
/* scan tiles of an array of n counts (4096 per tile) */
int map_tiles(int n);
/* radix passes of 8 bits that order the pixels of a map of npix pixels */
int map_sort_passes(int npix);
hipError_t launch_map_accumulate(const Hpx &h, const MapChunk &c, const MapScratch &s,
                                 double *acc, hipStream_t st);
/* ppix[s] = pixel of pointing s (-1: skipped), for a pointing shared by
 * every detector */
hipError_t launch_map_ptg_pixel(const Hpx &h, const float *ra, const float *dec, long n,
                                int *ppix, hipStream_t st);
hipError_t launch_map_solve(int npix, const double *acc, float *I, float *Q, float *U,
                            int *hits, unsigned long long *singular, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long *x, long n, hipStream_t st);

/* Detector pointing from array-centre pointing (focalplane.hip).  Outputs
 * [ndets][n]; any output may be NULL. */
struct FocalPlaneParams {
    int ndets;
    const double *dx, *dy, *pol;          /* [ndets] */
    long n;
    const double *x_ac, *y_ac, *rot_ac;   /* [n] */
    int flag;                             /* 0 az/el, 1 ra/dec, 2 source-relative */
    double *x, *y, *rot;
    float *x32, *y32, *rot32;
};
hipError_t launch_focal_plane(const FocalPlaneParams &q, hipStream_t st);

hipError_t launch_batch_forward(int group, const BatchParams &p, hipStream_t st);

hipError_t launch_pair(int op, const PairParams &p, hipStream_t st);
/* The beam-interpolation mode of the same operations (k_pair_interp). */
hipError_t launch_pair_interp(int op, const PairParams &p, hipStream_t st);
/* Forward with 64/sub samples per wave, sub = 8, 16 or 32 lanes each (small
 * discs); same results as launch_pair(OP_FWD), or OP_FWD_V with vterm. */
hipError_t launch_fwd_sub(int sub, const PairParams &p, hipStream_t st, bool vterm = false);
/* Beam adjoint in LDS bands of beam rows (work items: band x chunk of
 * samples); p.gbeam accumulates in FP64, p.sky4 holds the packed sky.  rows:
 * rows per band (0 or more than fits: as many as fit the LDS). */
hipError_t launch_adjbeam_bands(const PairParams &p, int chunk, int rows, int nblocks, int ncu,
                                int lanes, hipStream_t st);
int adjbeam_band_rows();

/* Beam adjoint through visit records: each visit mapped twice (count, then
 * write its record in the order of its band of beam rows), bands summed in
 * LDS; p.gbeam accumulates in FP64, p.sky4 holds the packed sky. */
constexpr int kRecMaxBands = 512;    /* more bands: the banded kernels */
struct BeamRec;
struct RecParams {
    int rows, nband;
    float inv_rows;
    BeamRec *brec;        /* records in band order, at most chunk x bound */
    long nrec_cap;        /* records brec holds (below 2^31) */
    long *bcnt;           /* [nband][chunk] records per (band, sample), then their offsets */
    long chunk;           /* samples per chunk (the row length of bcnt) */
    long *bbase;          /* [nband + 1] records per band, then the bands' starts */
    const float *vmax;    /* device: max |g|, max over sky pixels of max(|I|, |Q| + |U|) */
    double *part;         /* [nsplit][npix_beam][6] partial sums */
    int nsplit;
};
struct RecPlan {
    int rows, nband, nsplit;
    long chunk;           /* samples per chunk */
    long bound;           /* visits per sample, at most */
    size_t nrec;          /* records (16 bytes) */
    size_t nidx;          /* longs: bcnt, bbase */
    size_t npart;         /* doubles of the partial sums */
};
/* rows: rows per band (0: as many as the band sums' LDS holds); pool_bytes:
 * the memory budget of the records and offsets, which sets the samples per
 * chunk.  False when the beam needs more than kRecMaxBands bands. */
bool adjbeam_rec_plan(const PairParams &p, int rows, size_t pool_bytes, int ncu, RecPlan &pl);
/* vmax: two device floats of scratch. */
hipError_t launch_adjbeam_records(const PairParams &p, const RecPlan &pl, void *recs, long *idx,
                                  float *vmax, double *part, hipStream_t st);
/* Sky adjoint with LDS-privatised windows over chunks of `chunk` samples;
 * `cap` window pixels and `rw` rings per chunk; nblocks persistent blocks. */
/* fx_scale (device, or NULL): deterministic mode with the scale of
 * launch_fix_scale; the global accumulators then hold int64 bits.  lanes
 * (single pair only): 64, or 8 or 16 lanes per sample with 64/lanes samples
 * per wave (small discs) */
hipError_t launch_adjmap_lds(const PairParams &p, int chunk, int cap, int rw, int nblocks,
                             hipStream_t st, const BatchParams *batch = nullptr,
                             double rho_w = 0.0, int group = 1,
                             const double *fx_scale = nullptr, int lanes = 64);
/* group: 0 one pair, 1 per-pair batch items, 2 or 4 pairs per wave; det: the
 * deterministic mode's integer windows */
size_t adjmap_lds_bytes(int cap, int rw, int group = 0, bool det = false);
/* Deterministic sky adjoint: *scale (device) from the beams (beam8 rows)
 * and the cotangents (g_a, g_b of nsamples floats, either may be NULL);
 * count = the most contributions a pixel can receive; work holds 2 unsigned
 * (device). */
hipError_t launch_fix_scale(const float *const *beams, const long *npix, int nbeams,
                            const float *g_a, const float *g_b, long nsamples, long count,
                            unsigned *work, double *scale, hipStream_t st);
/* y = int64 bits of x / scale */
hipError_t launch_fix_to_f32(const double *x, const double *scale, float *y, long n,
                             hipStream_t st);
/* Launch resources of the sky-adjoint kernel, for error messages. */
int adjmap_describe(char *buf, size_t n, int cap, int rw);
/* *last = index of the last row of beam8 [npix][8] with a non-zero value in
 * columns 0..5 (-1 if none). */
hipError_t launch_beam_support(const float *beam8, long npix, int *last, hipStream_t st);
hipError_t launch_ring_table(const Hpx &h, RingRec *out, hipStream_t st);
hipError_t launch_zero_f64(double *x, long n, hipStream_t st);
/* sky4[p] = (I[p], Q[p], U[p], V[p]), V[p] = 0 if V is NULL */
hipError_t launch_pack_sky(const float *I, const float *Q, const float *U, long npix,
                           float4 *sky4, hipStream_t st, const float *V = nullptr);
hipError_t launch_f64_to_f32(const double *x, float *y, long n, hipStream_t st);
/* y = x + off (float), n elements */
hipError_t launch_add_offset(const float *x, float off, float *y, long n, hipStream_t st);
hipError_t launch_pack_beam(long npix, const float *const in[8], float *out, hipStream_t st);

hipError_t launch_pix2ang(const Hpx &h, long n, const int *pix, double *th, double *ph,
                          hipStream_t st);
hipError_t launch_ang2pix(const Hpx &h, long n, const double *th, const double *ph,
                          int *pix, hipStream_t st);
hipError_t launch_interpol(const Hpx &h, long n, const double *th, const double *ph,
                           int *pix, double *wgt, hipStream_t st);
hipError_t launch_disc(const Hpx &h, long n, const float *ra, const float *dec,
                       double radius, double cosr, const long *offsets, int *counts,
                       int *pixels, hipStream_t st);
hipError_t launch_match(const Hpx &hs, const RingRec *rings, const Hpx &hb, int mode, long n,
                        const float *ra, const float *dec, const float *pa,
                        const long *offsets, const int *pixels, int *bpix, double *chi,
                        hipStream_t st);
FastConsts fast_consts(const Hpx &hb);
FastConsts64 fast_consts64(const Hpx &hb);
hipError_t launch_libm(int op, long n, const double *x, const double *y, double *o,
                       hipStream_t st);

/* Records msg as the calling thread's pisco_last_error() and returns code
 * (host code outside capi.hip). */
int set_error(int code, const char *msg);

/* A single-pair launch's parameters from the arguments of pisco_dev_forward
 * and the adjoints, checked (PISCO_ERR_INVALID with the message set
 * otherwise); p.rings is left NULL for the caller to fill. */
int pair_params(int nside_sky, const float *I, const float *Q, const float *U, int nside_beam,
                long npix_beam, const float *beam8, long nsamples, const float *ra,
                const float *dec, const float *pa, PairParams &p);
/* Lanes per sample of the nearest-pixel forward from the expected disc size
 * npix (1 - cos rho_max) / 2: 8 below 1,000 pixels, 16 below 3,000, else 64. */
int forward_lanes_auto(const PairParams &p);

} /* namespace pisco */

#endif
