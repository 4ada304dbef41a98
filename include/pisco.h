/*
 * pisco.h -- C-ABI of the MI355X (gfx950) pixel-space beam convolution.
 *
 * Drop-in boundary for PISCO's GPUConvolver path.  Every entry point takes
 * plain pointers and sizes and returns an int status (PISCO_OK == 0); on
 * failure pisco_last_error() describes the problem.  Nothing here calls
 * exit(): where the reference printed and exited (CUDA_ERROR_CHECK,
 * src/GpuConvolver/gpuconvolver.cu:8-32; launchers, cudaConv2.cu:262-268) or
 * threw std::invalid_argument (gpuconvolver.cu:74-78, convolver.cpp:66-70),
 * these functions return an error code.
 *
 * Conventions shared with the reference:
 *   - sky maps are RING-ordered HEALPix maps of 12*nside^2 float pixels
 *     (Sky, src/Sky/sky.hpp:15-20);
 *   - beams are the first npix RING pixels of a HEALPix map of NSIDE
 *     nside_beam centred on the north pole, four maps (I,Q,U,V response) per
 *     detector, detectors a and b (PolBeam::aBeams/bBeams,
 *     src/Polbeam/polbeam.hpp:51-52); the disc radius is
 *     rhoMax = theta(pixel npix) (polbeam.cpp:35-36);
 *   - pointing is float ra, dec, pa in radians with pa counter-clockwise
 *     (Scan, src/Scan/scan.hpp:14-18); the kernels use psi = -pa
 *     (convolver.cpp:49-51, gpuconvolver.cu:375-377);
 *   - TOD follows the CPU Convolver (convolver.cpp:157-164): the V maps are
 *     not used.
 * Departures (deliberate): every sample is computed (the reference drops the
 * nsamples % ptgPerConv tail, gpuconvolver.cu:386), and buffers are
 * re-allocated when an update grows them (gpuconvolver.cu:279-290,354-367).
 *
 * Host-buffer entry points copy their inputs; callers may free them after the
 * call returns.  pisco_dev_* entry points take device pointers and a
 * hipStream_t (NULL = HIP's null stream) and are asynchronous with respect to
 * the host.  Calls on one context share its workspace and must not overlap
 * in time on different streams.
 */
#ifndef PISCO_H
#define PISCO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PISCO_OK 0
#define PISCO_ERR_INVALID 1      /* bad argument or configuration */
#define PISCO_ERR_STATE 2        /* call order (update_* before exec) */
#define PISCO_ERR_MEMORY 3       /* allocation failed or over maxMemUsage */
#define PISCO_ERR_DEVICE 4       /* HIP runtime error */

/* Mirrors CUDACONV::RunConfig (src/GpuConvolver/cuda/cudaConv2.h:13-26).
 * maxMemUsage is checked against the reference's own estimate
 * (gpuconvolver.cu:41-64); ptgPerConv is the scan chunk of one launch;
 * gridSize / blockSize / pixelsPerDisc are accepted and, as in the
 * reference, not used to size anything (the kernels size their own
 * work from the exact discs). */
typedef struct pisco_run_config {
    int nStreams;
    size_t maxMemUsage;
    int deviceId;
    int gridSizeX;
    int gridSizeY;
    int blockSizeX;
    int blockSizeY;
    int ptgPerConv;
    int pixelsPerDisc;
} pisco_run_config;

typedef struct pisco_ctx pisco_ctx;

/* Text of the last error of the calling thread ("" if none). */
const char *pisco_last_error(void);

/* Library / device information: writes up to len bytes. */
int pisco_version(char *buf, int len);

/* GPUConvolver::GPUConvolver(RunConfig) + configure_execution()
 * (gpuconvolver.cu:34-87): validates the configuration, selects
 * cfg->deviceId and creates the context's stream. */
int pisco_create(const pisco_run_config *cfg, pisco_ctx **out);

/* GPUConvolver::~GPUConvolver (gpuconvolver.cu:172-221). NULL is ignored. */
void pisco_destroy(pisco_ctx *ctx);

/* No counterpart in the reference (GPUConvolver re-allocates nothing after
 * its first update, gpuconvolver.cu:279-290,354-367): frozen != 0 pins the
 * context's device workspace, so that a graph captured from calls on it
 * keeps valid pointers.  While frozen, a call that would need a larger
 * workspace fails with PISCO_ERR_STATE instead of re-allocating it;
 * frozen == 0 lifts the pin. */
int pisco_freeze_buffers(pisco_ctx *ctx, int frozen);

/* GPUConvolver::update_sky(Sky*) (gpuconvolver.cu:223-269).  V may be NULL. */
int pisco_update_sky(pisco_ctx *ctx, int nside, const float *I, const float *Q,
                     const float *U, const float *V);

/* GPUConvolver::update_beam(PolBeam*) (gpuconvolver.cu:271-345): a[k], b[k]
 * are PolBeam::aBeams[k], bBeams[k] (k = 0..3, npix floats each; a[3]/b[3]
 * may be NULL). */
int pisco_update_beam(pisco_ctx *ctx, int nside_beam, long npix,
                      const float *const a[4], const float *const b[4]);

/* GPUConvolver::update_scan(Scan*) (gpuconvolver.cu:347-380). */
int pisco_update_scan(pisco_ctx *ctx, long nsamples, const float *ra,
                      const float *dec, const float *pa);

/* GPUConvolver::exec_convolution(data_a, data_b) (gpuconvolver.cu:382-505)
 * with the polFlag of Convolver::exec_convolution (convolver.cpp:22-72):
 * 'a' writes tod_a, 'b' writes tod_b, 'p' writes both; the GPUConvolver
 * form (no flag) is polflag 'p'. */
int pisco_exec_convolution(pisco_ctx *ctx, char polflag, float *tod_a,
                           float *tod_b);

/* Adjoint with respect to the sky (no counterpart in the reference: the
 * transpose of convolver.cpp:157-164): gI,gQ,gU (12*nside^2 floats each)
 * = A^T (g_a, g_b).  g_a or g_b may be NULL (treated as zero). */
int pisco_exec_adjoint_map(pisco_ctx *ctx, const float *g_a, const float *g_b,
                           float *gI, float *gQ, float *gU);

/* Adjoint with respect to the beam maps (no counterpart in the reference):
 * ga[k], gb[k] receive d<g,TOD>/d aBeams[k], bBeams[k] (npix floats each;
 * k = 3 is zero because the forward does not read V, unless the v_term
 * tuning key is set).  Entries may be NULL. */
int pisco_exec_adjoint_beam(pisco_ctx *ctx, const float *g_a, const float *g_b,
                            float *const ga[4], float *const gb[4]);

/* Launch-shape tuning of a context (defaults are the measured best on the
 * NSIDE 512, 5-degree workload; except for beam_interpolation, results do
 * not depend on them).  Keys: "adjmap_chunk" (samples per LDS window of the sky adjoint;
 * 0 selects direct global atomics; -1, the default, picks 16 to 128 from the
 * disc size for single-pair adjoints and 16 for batched ones), "adjmap_cap" (window pixels),
 * "adjmap_rw" (window rings), "adjmap_blocks" (blocks of the adjoint launches,
 * each looping over windows; 0, the default, one block per window), "adjmap_support" (1: windows cover only the discs of the beam's
 * non-zero rows, 0: the full rho_max discs; results are the same either way),
 * "batch_group" (batched forward: 1 = one wave per (pair, sample), the
 * single-pair computation; 2 or 4 = pairs per wave sharing each visit's
 * geometry), "batch_adj_group" (batched sky adjoint: 1, 2 or 4 pairs per
 * wave, as batch_group), "forward_lanes" (single-pair forward: 64 = one
 * sample per wave; 32, 16 or 8 lanes per sample with 64/lanes samples per
 * wave; 0, the default, picks 8 below 1,000 expected pixels per disc, 16
 * below 3,000, else 64), "adjmap_lanes" (single-pair sky adjoint: 64, 16 or
 * 8; 0, the default, picks 8 below 1,000 pixels per disc, else 64),
 * "batch_pair_major" (batched launches: 1 runs the pairs one after another
 * with the single-pair kernels, 0 all pairs per chunk of samples; -1, the
 * default, picks 1 when a pair's packed beam exceeds 4 MiB),
 * "batch_pair_range" (batched launches over consecutive ranges of the pair
 * table, sorted by disc radius: k > 0 (a multiple of 4) pairs per range, -1
 * one range of all pairs; 0, the default, ranges whose packed beams add up
 * to at most 4 MiB), "adjbeam_bands" (beam adjoint: 1, the default, sums in
 * LDS bands of beam rows; 0 global FP64 atomics), "adjbeam_records" (with
 * adjbeam_bands 1: 1, the default, maps each visit twice -- counting, then
 * writing a 16-byte record per visit in the order of its band of beam rows
 * -- and sums the records band by band in LDS as 64-bit fixed-point
 * integers, which makes the result independent of the order of the
 * additions (beams of up to 512 bands); 0 walks each sample's annulus
 * once per band), "adjbeam_pool_mb" (memory of the records, which sets the
 * samples per chunk, MiB; default 8192, and at most a quarter of the free
 * device memory),
 * "adjbeam_lanes" (annulus walks: 8, the default, lanes per sample
 * with 8 samples per wave, or 64),
 * "adjbeam_chunk" (samples per annulus work item, default 1024),
 * "adjbeam_rows" (rows per band; 0, the default, as many as fit the LDS),
 * "adjmap_deterministic" (sky adjoint sums: 1 rounds every contribution
 * once to a multiple of 2^-k and sums 64-bit integers, so the output bits
 * do not depend on the order of the additions, the launch shape or the
 * block count (batched launches: for a given grouping of the pairs, each
 * visit's sum over a group formed in pair order); 0 float sums; -1, the
 * default, integer sums for nearest-pixel beams, where they also measured
 * faster, and float sums with beam interpolation; 1 is refused with beam
 * interpolation),
 * "v_term" (host-buffer calls: 1 adds the reference GPU kernel's V A3 and
 * V B3 (cudaConv2.cu:192,204,214,217-218) to pisco_exec_convolution's TOD,
 * with update_sky's V (zero if it was NULL), and fills the a[3]/b[3]
 * gradients of pisco_exec_adjoint_beam; 0, the default, follows the CPU
 * Convolver, which reads no V; changes results);
 * pisco_get_tuning also reads "cu_count". */
int pisco_set_tuning(pisco_ctx *ctx, const char *key, long value);
int pisco_get_tuning(pisco_ctx *ctx, const char *key, long *value);

/* ---- device-resident entry points (device pointers, asynchronous) ------- */

/* Packs eight float beam maps (a0..a3, b0..b3; a3/b3 may be NULL) into the
 * kernels' row layout beam8: npix rows of A0 A1 A2 B0 B1 B2 A3 B3
 * (8*npix floats, device). */
int pisco_dev_pack_beam(pisco_ctx *ctx, void *stream, long npix,
                        const float *const maps[8], float *beam8);

/* Forward: tod_a/tod_b (either may be NULL) for nsamples samples. */
int pisco_dev_forward(pisco_ctx *ctx, void *stream, int nside_sky,
                      const float *I, const float *Q, const float *U,
                      int nside_beam, long npix_beam, const float *beam8,
                      long nsamples, const float *ra, const float *dec,
                      const float *pa, float *tod_a, float *tod_b);

/* No counterpart in the reference: pisco_dev_forward for callers that
 * cannot allocate device memory during the call (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It takes no
 * context and allocates nothing: its workspace (the sky's ring table and
 * packed pixels) is a fixed device array of 48.1 MiB that each device the
 * library runs on holds for the life of the process, so the sky NSIDE is at
 * most 512 (a larger one is refused with PISCO_ERR_INVALID; pisco_dev_forward
 * takes any).  Runs on the current device and the given stream with the
 * default launch shapes; results equal pisco_dev_forward's bit for bit.
 * Every call on a device shares that device's workspace: calls that overlap
 * in time on one device (other streams, other host threads) race. */
int pisco_dev_forward_noalloc(void *stream, int nside_sky, const float *I, const float *Q,
                              const float *U, int nside_beam, long npix_beam,
                              const float *beam8, long nsamples, const float *ra,
                              const float *dec, const float *pa, float *tod_a, float *tod_b);

/* Forward with the V terms of the reference's GPU kernel: d_a gains V A3
 * and d_b V B3 (cudaConv2.cu:192,204,214,217-218), appended last to the
 * per-visit float expression of convolver.cpp:157-164; V (device,
 * 12*nside_sky^2 floats) is required.  The beam rows' A3 and B3 columns are
 * read.  Not available with beam interpolation (PISCO_ERR_INVALID). */
int pisco_dev_forward_v(pisco_ctx *ctx, void *stream, int nside_sky, const float *I,
                        const float *Q, const float *U, const float *V, int nside_beam,
                        long npix_beam, const float *beam8, long nsamples, const float *ra,
                        const float *dec, const float *pa, float *tod_a, float *tod_b);

/* Adjoint w.r.t. the sky: gI,gQ,gU (device, 12*nside_sky^2 floats each). */
int pisco_dev_adjoint_map(pisco_ctx *ctx, void *stream, int nside_sky,
                          int nside_beam, long npix_beam, const float *beam8,
                          long nsamples, const float *ra, const float *dec,
                          const float *pa, const float *g_a, const float *g_b,
                          float *gI, float *gQ, float *gU);

/* Adjoint w.r.t. the beams: gbeam8 (device, npix_beam rows of 8 floats in
 * the beam8 layout; the A3/B3 entries are zero). */
int pisco_dev_adjoint_beam(pisco_ctx *ctx, void *stream, int nside_sky,
                           const float *I, const float *Q, const float *U,
                           int nside_beam, long npix_beam, long nsamples,
                           const float *ra, const float *dec, const float *pa,
                           const float *g_a, const float *g_b, float *gbeam8);

/* ---- batched detector pairs (device-resident) --------------------------- */

/* One detector pair of a batched launch: its beam (beam8 layout, device), its
 * beam map geometry and the offset of its position angle from the shared
 * pointing: the pair's pa is (float)(pa + pa_offset), i.e. what its own Scan
 * would hold.  The disc radius is rhoMax of its beam (polbeam.cpp:35-36). */
typedef struct pisco_batch_pair {
    int nside_beam;
    long npix_beam;
    const float *beam8;
    float pa_offset;
} pisco_batch_pair;

typedef struct pisco_batch pisco_batch;

/* A detector table on the context's device (no counterpart in the reference,
 * whose GPUConvolver holds one pair: gpuconvolver.hpp:17-82).  The table is
 * copied; the beams are referenced (they must stay valid while in use). */
int pisco_batch_create(pisco_ctx *ctx, int npairs, const pisco_batch_pair *pairs,
                       pisco_batch **out);
void pisco_batch_destroy(pisco_batch *batch);

/* Forward of every pair of the batch in one launch: tod (device) receives
 * npairs x 2 x nsamples floats, tod[(2k + 0) nsamples + i] = detector a of
 * pair k at sample i, [(2k + 1) nsamples + i] = detector b.  Each pair's TOD
 * is what pisco_dev_forward gives for that pair with pa + pa_offset. */
int pisco_dev_batched_forward(pisco_ctx *ctx, void *stream, const pisco_batch *batch,
                              int nside_sky, const float *I, const float *Q,
                              const float *U, long nsamples, const float *ra,
                              const float *dec, const float *pa, float *tod);

/* Sky adjoint of the batched forward: gI,gQ,gU (device, 12*nside^2 floats
 * each) = sum over pairs of A_k^T g_k, with g (device) in the layout of
 * pisco_dev_batched_forward's tod. */
int pisco_dev_batched_adjoint_map(pisco_ctx *ctx, void *stream, const pisco_batch *batch,
                                  int nside_sky, long nsamples, const float *ra,
                                  const float *dec, const float *pa, const float *g,
                                  float *gI, float *gQ, float *gU);

/* The batched forward with a pointing per pair: ra, dec and pa hold
 * npairs x nsamples floats, pair k (in the caller's order) at k * nsamples,
 * e.g. the float outputs of pisco_dev_focal_plane_to_sky for each pair's
 * focal-plane position.  pa_offset is still added.  One launch over
 * (pair, sample) items; pairs do not share visit geometry. */
int pisco_dev_batched_forward_pp(pisco_ctx *ctx, void *stream, const pisco_batch *batch,
                                 int nside_sky, const float *I, const float *Q,
                                 const float *U, long nsamples, const float *ra,
                                 const float *dec, const float *pa, float *tod);
/* Sky adjoint of pisco_dev_batched_forward_pp: one LDS-windowed sky adjoint
 * per pair, adding into one FP64 gradient. */
int pisco_dev_batched_adjoint_map_pp(pisco_ctx *ctx, void *stream, const pisco_batch *batch,
                                     int nside_sky, long nsamples, const float *ra,
                                     const float *dec, const float *pa, const float *g,
                                     float *gI, float *gQ, float *gU);

/* ---- binned map-maker (SkyMap) ------------------------------------------ */

/* The normal equations AtA (3x3) and AtD (3) of every pixel of a RING map,
 * held on one device; SkyMap (src/SkyMap/skymap.hpp:13-55). */
typedef struct pisco_skymap pisco_skymap;

/* SkyMap::SkyMap(int nside) (skymap.cpp:21-39).  The accumulators start at
 * zero (the reference's zeroing, skymap.cpp:26-38, is commented out and its
 * sums start from uninitialised memory). */
int pisco_skymap_create(int device, int nside, pisco_skymap **out);
void pisco_skymap_destroy(pisco_skymap *map);
/* Zero the accumulators (a fresh SkyMap). */
int pisco_skymap_reset(pisco_skymap *map);
/* Samples per accumulation pass (default 2^24; the sums do not depend on
 * it).  Scratch is about 41 bytes per sample of a pass. */
int pisco_skymap_set_chunk(pisco_skymap *map, long samples);

/* SkyMap::accumulate_data (skymap.cpp:68-132), host buffers.  ra, dec, pa
 * (float radians, pa counter-clockwise) and data hold ndet * nsamples floats,
 * detector-major; det_pol_angles holds ndet.  Adds onto the current sums in
 * the reference's order (detector, then sample), so the sums equal its
 * sequential FP64 sums bit for bit.  A sample whose theta = pi/2 - dec falls
 * outside [0, pi] (Healpix_Base::ang2pix's planck_assert) or whose ra is not
 * finite is skipped; every other sample is accumulated and the call then
 * returns PISCO_ERR_INVALID. */
int pisco_skymap_accumulate(pisco_skymap *map, long nsamples, const float *ra,
                            const float *dec, const float *pa, int ndet,
                            const float *det_pol_angles, const float *data);
/* Device-resident form on `stream` (NULL: HIP's null stream).  All pointers
 * are device memory.  shared_pointing = 1: ra, dec and pa hold nsamples
 * floats used by every detector (the sums of passing them replicated).
 * Skipped samples are counted (pisco_skymap_skipped), not reported.  Calls
 * on one map share its scratch: they must not overlap in time. */
int pisco_dev_skymap_accumulate(pisco_skymap *map, void *stream, long nsamples,
                                const float *ra, const float *dec, const float *pa,
                                int shared_pointing, int ndet, const float *det_pol_angles,
                                const float *data);
/* Samples skipped since the map was created or reset (synchronises the
 * device). */
int pisco_skymap_skipped(pisco_skymap *map, long *count);

/* SkyMap::solve_map (skymap.cpp:134-231) into host arrays of 12 nside^2:
 * pixels with at least 3 hits get the solution of AtA x = AtD by LAPACK
 * DGESV's arithmetic (Netlib DGETF2 + DGETRS); a singular system gives
 * I, Q, U = AtD and hits 0; fewer than 3 hits give 0, 0, 0 and the hits.
 * singular (nullable) receives the number of singular pixels, which the
 * reference prints one by one. */
int pisco_skymap_solve(pisco_skymap *map, float *I, float *Q, float *U, int *hits,
                       long *singular);
/* Device-resident form: I, Q, U, hits are device arrays; singular as above
 * (synchronises the stream when non-NULL). */
int pisco_dev_skymap_solve(pisco_skymap *map, void *stream, float *I, float *Q, float *U,
                           int *hits, long *singular);
/* The sums in the reference's storage (skymap.hpp:42-43): AtA [npix][9]
 * row-major 3x3, AtD [npix][3], host arrays. */
int pisco_skymap_get_normal(pisco_skymap *map, double *AtA, double *AtD);

/* ---- detector pointing (BPoint::focal_plane_to_sky) ---------------------- */

/* BPoint::focal_plane_to_sky (src/Bpoint/bpoint.cpp:553-674) on the device:
 * detector offsets det_dx (right), det_dy (up) in the receiver frame seen
 * looking out, and polarisation angles det_pol_angle (ccw from the receiver
 * y axis), all [ndets] radians; array-centre coordinates x_ac, y_ac and
 * angle rot_ac [nsamples] radians.  flag 0: az/el (x_ac azimuth, rot_ac
 * boresight angle ccw from zenith), 1: ra/dec (rot_ac position angle ccw from
 * the pole), 2: source-relative.  Outputs [ndets][nsamples] detector-major,
 * each NULL or device memory: x_det, y_det, rot_det in FP64 and, rounded to
 * float, x32, y32, rot32 (with flag 1 these are a Scan's ra, dec, pa).  All
 * input pointers are device memory; FP64 throughout, the same bits as the
 * CPU restatement. */
int pisco_dev_focal_plane_to_sky(pisco_ctx *ctx, void *stream, int ndets, const double *det_dx,
                                 const double *det_dy, const double *det_pol_angle,
                                 long nsamples, const double *x_ac, const double *y_ac,
                                 const double *rot_ac, int flag, double *x_det, double *y_det,
                                 double *rot_det, float *x32, float *y32, float *rot32);
/* The same with host arrays (the reference's signature, BPoint.h:99-103;
 * outputs of ndets * nsamples doubles each). */
int pisco_focal_plane_to_sky(pisco_ctx *ctx, int ndets, const double *det_dx,
                             const double *det_dy, const double *det_pol_angle, long nsamples,
                             const double *x_ac, const double *y_ac, const double *rot_ac,
                             int flag, double *x_det, double *y_det, double *rot_det);

/* ---- text files of the reference's formats -------------------------------- */

/* Reads a text file of ncols whitespace-separated numbers per row, the
 * format numpy.savetxt writes for the reference's inputs (sky maps: rows of
 * I Q U V per pixel, scripts/create_cmb.py:96-97; tilde beams: rows of
 * I Q U V per beam pixel, scripts/fields2beams.py:175-176).  Column k of
 * row r goes to cols[k][r] (a NULL column is skipped); values are rounded
 * to float as strtof does.  Blank lines and lines starting with '#' are
 * skipped.  *nrows receives the number of rows; with capacity 0 nothing is
 * stored (a row count), otherwise a file of more than capacity rows is an
 * error. */
int pisco_read_txt_columns(const char *path, int ncols, long capacity, float *const cols[],
                           long *nrows);
/* Writes nrows rows of ncols columns as numpy.savetxt does ("%.18e",
 * space-separated, one row per line). */
int pisco_write_txt_columns(const char *path, int ncols, long nrows, const float *const cols[]);

/* ---- HEALPix / matching primitives on the device (for verification) ---- */

/* Healpix_Base::pix2ang for n pixels (device arrays). */
int pisco_dev_pix2ang(pisco_ctx *ctx, void *stream, int nside, long n,
                      const int *pix, double *theta, double *phi);
/* Healpix_Base::ang2pix(pointing) for n angles (device arrays). */
int pisco_dev_ang2pix(pisco_ctx *ctx, void *stream, int nside, long n,
                      const double *theta, const double *phi, int *pix);
/* Healpix_Base::get_interpol: pix[4n], wgt[4n] (device arrays). */
int pisco_dev_get_interpol(pisco_ctx *ctx, void *stream, int nside, long n,
                           const double *theta, const double *phi, int *pix,
                           double *wgt);
/* Non-inclusive query_disc around (pi/2 - dec, ra) of n pointings: with
 * pixels == NULL writes counts[n]; otherwise writes the pixel lists at
 * offsets[i] (device arrays). */
int pisco_dev_query_disc(pisco_ctx *ctx, void *stream, int nside, long n,
                         const float *ra, const float *dec, double radius,
                         const long *offsets, int *counts, int *pixels);
/* Sky pixel -> (beam pixel, chi) for the pixel lists of n pointings
 * (what kernel 1 of the reference computes, cudaConv2.cu:59-128). */
int pisco_dev_match(pisco_ctx *ctx, void *stream, int nside_sky,
                    int nside_beam, long n, const float *ra, const float *dec,
                    const float *pa, const long *offsets, const int *pixels,
                    int *bpix, double *chi);
/* As pisco_dev_match (mode 0).  Mode 1 evaluates what the adjoint kernels
 * use -- the certified single-precision form, with the exact FP64 mapping
 * wherever that form cannot certify a pixel (same beam pixels by
 * construction) -- and writes float cos(2chi), sin(2chi) packed into each
 * 8-byte chi element instead of chi.  Mode 2 is the single-precision form
 * alone: bpix = -2 where it defers to the exact mapping.  Mode 3 is the
 * certified double-precision form of the small-disc kernels (same output
 * layout as mode 1).  Mode 4 is what the one-wave-per-sample forward uses:
 * the double-precision form with a float atan2 for the position angle,
 * and mode 3 where that cannot certify a pixel; mode 5 is that form alone
 * (bpix = -2 where it declines). */
int pisco_dev_match_mode(pisco_ctx *ctx, void *stream, int mode, int nside_sky,
                         int nside_beam, long n, const float *ra,
                         const float *dec, const float *pa, const long *offsets,
                         const int *pixels, int *bpix, double *chi);
/* Elementwise FP64 primitives of the index path on the device (op: 0 sin,
 * 1 cos, 2 acos, 3 atan2(x,y), 4 sqrt, 5 x/y, 6 fmodulo(x,y), 7 sincos;
 * 8 the fixed-point conversion of the integer sums: rint((float)x * y) for
 * a power of two y and |x y| < 2^51, as an int64 stored in the output's
 * 8 bytes). */
int pisco_dev_libm(pisco_ctx *ctx, void *stream, int op, long n,
                   const double *x, const double *y, double *out);

/* ---- host-side helpers (no device needed) ------------------------------ */

/* rhoMax of a beam of npix pixels: theta of pixel npix at NSIDE nside_beam
 * (PolBeam ctor, polbeam.cpp:35-36).  Returns -1 on bad arguments. */
double pisco_beam_rho_max(int nside_beam, long npix);

/* Healpix_Base::ang2pix(pointing(theta, phi)) on the host; used to size a
 * beam cap as scripts/fields2beams.py:102 does (npix = ang2pix(nside, R, 0)).
 * Returns -1 on bad arguments. */
long pisco_ang2pix(int nside, double theta, double phi);

/* Healpix_Base::pix2ang on the host.  Returns PISCO_OK or an error. */
int pisco_pix2ang(int nside, long pix, double *theta, double *phi);

/* The sky pixels (NSIDE nside_sky, RING) that the banded beam adjoint walks
 * for band `band` (beam rows [band*rows, min(npix_beam, (band+1)*rows)) of
 * a beam map of NSIDE nside_beam and npix_beam pixels) around the pointing
 * (ra, dec): the annulus runs of every sky ring, as the kernels form them
 * (cover != 0: the float-span form the kernels use, with an IEEE quotient
 * where the kernels use the hardware reciprocal estimate, within 1 ulp of
 * it and inside the cover's 4e-6 rad margin; 0: exact spans).  Writes
 * at most cap pixel indices to pix, in walk order, and returns their
 * number; -1 on bad arguments, -2 if more than cap.  Host only.  No
 * counterpart in the reference (its beam maps have no gradient); it exposes
 * the decomposition so that its coverage can be checked. */
long pisco_band_annulus(int nside_sky, int nside_beam, long npix_beam, int rows, int band,
                        float ra, float dec, int cover, long cap, int *pix);

/* E-field magnitude and phase of a perfectly co-polarised elliptical
 * Gaussian beam on the first npix pixels of a beam map
 * (PolBeam::make_unpol_gaussian_elliptical_beams, polbeam.cpp:290-372;
 * fwhm in degrees, phi0 in degrees). */
int pisco_beam_gaussian_fields(int nside_beam, long npix, double fwhmx,
                               double fwhmy, double phi0, double *magEco,
                               double *phsEco, double *magEcx, double *phsEcx);

/* tilde I,Q,U,V of one detector from its E fields
 * (PolBeam::beam_from_fields, polbeam.cpp:117-184). */
int pisco_beam_from_fields(long npix, const double *magEco, const double *phsEco,
                           const double *magEcx, const double *phsEcx, float *I,
                           float *Q, float *U, float *V);

/* Normalised, leakage-mixed beams (PolBeam::build, polbeam.cpp:186-288):
 * tilde = {Ia,Qa,Ua,Va,Ib,Qb,Ub,Vb}, out = {aBeams[0..3], bBeams[0..3]};
 * dets is 'a', 'b' or 'p' (PolBeam::enabledDets). */
int pisco_polbeam_build(long npix, int nside_beam, int nside_sky, double epsilon,
                        char dets, const float *const tilde[8],
                        float *const out[8]);

#ifdef __cplusplus
}
#endif

#endif /* PISCO_H */
