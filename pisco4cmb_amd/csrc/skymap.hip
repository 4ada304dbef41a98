/*
 * skymap.hip -- the binned map-maker on the device: SkyMap::accumulate_data
 * and SkyMap::solve_map (src/SkyMap/skymap.cpp:68-231).
 *
 * Accumulation.  Each pixel's normal equations are summed over its samples in
 * the order the reference visits them (detector-major, then sample), so the
 * FP64 sums equal the reference's sequential sums bit for bit:
 *   1. k_mm_pixel    pixel of every sample (Healpix_Base::ang2pix) and the
 *                    number of samples per pixel
 *   2. k_mm_tile_sums, k_mm_scan_sums, k_mm_tile_scan
 *                    exclusive prefix sums of the counts: where each pixel's
 *                    samples start in pixel order
 *   3. k_rs_hist, k_rs_scatter
 *                    a stable radix sort of the samples by pixel, so each
 *                    pixel's samples are listed in sample order
 *   4. k_mm_terms    cos 2psi, sin 2psi and the datum of every sorted sample
 *                    (all samples in parallel)
 *   5. k_mm_walk     one thread per pixel adds its samples' terms, in that
 *                    order, onto the pixel's sums: nine dependent FP64 adds
 *                    per sample
 * Only the 6 distinct entries of the symmetric AtA are kept: the reference's
 * AtA[1] and AtA[3] (and [2]/[6], [5]/[7]) receive the same terms in the
 * same order and are therefore equal.
 *
 * Solve.  One thread per pixel with the arithmetic of LAPACK's DGESV for
 * n = 3 (Netlib DGETF2 + DGETRS), the pivot search and row swaps written as
 * selects so the 3x3 system stays in registers.
 */
#include <hip/hip_runtime.h>

#include "device_hpx.h"
#include "kernels.hpp"
#include "pisco_libm.h"

namespace pisco {

namespace {

constexpr int kTileThreads = 256;
constexpr int kTilePer = 16;
constexpr int kTile = kTileThreads * kTilePer;   /* counts per scan tile */

inline int grid_for(long n, int bs) { return (int)((n + bs - 1) / bs); }

/* Exclusive scan of one int per thread over the block (blockDim.x a multiple
 * of 64, at most 1024); *total receives the block's sum. */
__device__ int block_excl_scan(int v, int *total)
{
    __shared__ int wsum[16];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int x = v;
    for (int d = 1; d < 64; d <<= 1) {
        int y = __shfl_up(x, d, 64);
        if (lane >= d)
            x += y;
    }
    if (lane == 63)
        wsum[w] = x;
    __syncthreads();
    int base = 0, all = 0;
    for (int k = 0; k < nw; k++) {
        base += (k < w) ? wsum[k] : 0;
        all += wsum[k];
    }
    __syncthreads();   /* wsum may be reused by a following call */
    *total = all;
    return base + x - v;
}

} // namespace

__global__ void k_mm_fill_i32(int *x, long n, int v)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = v;
}

/* Pixel of a pointing: ang2pix(theta = pi/2 - dec, phi = ra) as
 * skymap.cpp:91-96; -1 for a theta outside [0, pi] (planck_assert of
 * Healpix_Base::ang2pix) or a non-finite ra. */
__device__ inline int map_pixel(const Hpx &h, float ra, float dec)
{
    double theta = kHalfPi - (double)dec;
    double phi = ra;
    int p = -1;
    if (theta >= 0.0 && theta <= kPi && fabs(phi) < INFINITY) {
        p = ang2pix(h, theta, phi);
        if (p < 0 || p >= h.npix)
            p = -1;
    }
    return p;
}

/* Pixels of a pointing shared by every detector: ppix[s], s < n. */
__global__ void k_mm_ptg_pixel(Hpx h, const float *ra, const float *dec, long n, int *ppix)
{
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n)
        ppix[s] = map_pixel(h, ra[s], dec[s]);
}

/* Adds v to x[key] once per run of equal keys among the active lanes of a
 * wave (consecutive samples usually fall in one pixel). */
__device__ inline void wave_run_add(int key, bool active, int *x, unsigned long long *skipped)
{
    const int lane = threadIdx.x & 63;
    int prev = __shfl_up(key, 1, 64);
    bool prev_active = __shfl_up((int)active, 1, 64) != 0;
    bool head = active && (lane == 0 || !prev_active || prev != key);
    unsigned long long heads = __ballot(head);
    unsigned long long act = __ballot(active);
    if (!head)
        return;
    /* the run ends before the next head or the next inactive lane */
    unsigned long long after = (lane == 63) ? 0ull : (~0ull << (lane + 1));
    unsigned long long stop = (heads | ~act) & after;
    int end = stop ? __ffsll((long long)stop) - 1 : 64;
    int len = end - lane;
    if (key < 0)
        atomicAdd(skipped, (unsigned long long)len);
    else
        atomicAdd(&x[key], len);
}

/* Pixel key of sample j of the chunk (0xffffffff when skipped) and the
 * number of samples per pixel. */
__global__ void k_mm_pixel(Hpx h, MapChunk c, const int *ppix, unsigned *pix, int *count,
                           unsigned long long *skipped)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = j < c.len;
    int p = -1;
    if (active) {
        long t = c.t0 + j;
        long det = t / c.nsamples;
        long s = t - det * c.nsamples;
        if (ppix)
            p = ppix[s];
        else
            p = map_pixel(h, c.ra[det * c.ptg_stride + s], c.dec[det * c.ptg_stride + s]);
        pix[j] = p < 0 ? 0xffffffffu : (unsigned)p;
    }
    wave_run_add(p, active, count, skipped);
}

/* tsum[b] = sum of count[b * kTile, (b + 1) * kTile) */
__global__ __launch_bounds__(kTileThreads) void k_mm_tile_sums(const int *count, int n,
                                                               int *tsum)
{
    long base = (long)blockIdx.x * kTile;
    int s = 0;
    for (int k = 0; k < kTilePer; k++) {
        long i = base + k * kTileThreads + threadIdx.x;
        s += (i < n) ? count[i] : 0;
    }
    int total;
    (void)block_excl_scan(s, &total);
    if (threadIdx.x == 0)
        tsum[blockIdx.x] = total;
}

/* Exclusive scan of tsum[0, nt) in place, one block of 1024; off[n] = total. */
__global__ __launch_bounds__(1024) void k_mm_scan_sums(int *tsum, int nt, int *off, int n)
{
    int per = (nt + blockDim.x - 1) / blockDim.x;
    int b = threadIdx.x * per;
    int e = min(b + per, nt);
    int s = 0;
    for (int i = b; i < e; i++)
        s += tsum[i];
    int total;
    int run = block_excl_scan(s, &total);
    for (int i = b; i < e; i++) {
        int v = tsum[i];
        tsum[i] = run;
        run += v;
    }
    if (threadIdx.x == 0)
        off[n] = total;
}

/* off[i] (and cur[i] unless cur is NULL) = exclusive prefix sum of count at
 * i. */
__global__ __launch_bounds__(kTileThreads) void k_mm_tile_scan(const int *count, int n,
                                                               const int *tsum, int *off,
                                                               int *cur)
{
    __shared__ int tile[kTile];
    long base = (long)blockIdx.x * kTile;
    for (int k = 0; k < kTilePer; k++) {
        long i = base + k * kTileThreads + threadIdx.x;
        tile[k * kTileThreads + threadIdx.x] = (i < n) ? count[i] : 0;
    }
    __syncthreads();
    int *mine = tile + threadIdx.x * kTilePer;
    int s = 0;
    for (int k = 0; k < kTilePer; k++)
        s += mine[k];
    int total;
    int run = block_excl_scan(s, &total) + tsum[blockIdx.x];
    for (int k = 0; k < kTilePer; k++) {
        int v = mine[k];
        mine[k] = run;
        run += v;
    }
    __syncthreads();
    for (int k = 0; k < kTilePer; k++) {
        long i = base + k * kTileThreads + threadIdx.x;
        if (i < n) {
            int v = tile[k * kTileThreads + threadIdx.x];
            off[i] = v;
            if (cur)
                cur[i] = v;
        }
    }
}

/* ---- stable LSD radix sort of the chunk's samples by pixel --------------
 * Keys are pixels (0xffffffff for a skipped sample, which sorts last), values
 * the chunk-local sample index.  Eight bits per pass; within a tile of
 * kTile items an item's rank among equal digits follows item order (wave
 * ballots give the rank inside a wave, per-wave counts in LDS the rank across
 * the block's waves), so every pass is stable and the result lists each
 * pixel's samples in increasing sample index. */

/* hist[d * nb + b] = number of items of tile b whose digit is d */
__global__ __launch_bounds__(kTileThreads) void k_rs_hist(const unsigned *keys, int n, int shift,
                                                          int nb, int *hist)
{
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    long base = (long)blockIdx.x * kTile;
    for (int k = 0; k < kTilePer; k++) {
        long i = base + k * kTileThreads + threadIdx.x;
        if (i < n)
            atomicAdd(&h[(keys[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

/* Scatter of one pass: hoff = exclusive scan of hist (digit-major).  FIRST:
 * the values are the item indices themselves. */
template <bool FIRST>
__global__ __launch_bounds__(kTileThreads) void k_rs_scatter(const unsigned *kin, const int *vin,
                                                             int n, int shift, int nb,
                                                             const int *hoff, unsigned *kout,
                                                             int *vout)
{
    static_assert(kTileThreads == 256, "one thread per digit");
    __shared__ int run[256];
    __shared__ int wcnt[kTileThreads / 64][256];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    run[tid] = hoff[tid * nb + blockIdx.x];
    for (int q = 0; q < kTileThreads / 64; q++)
        wcnt[q][tid] = 0;
    __syncthreads();
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const long base = (long)blockIdx.x * kTile;
    for (int k = 0; k < kTilePer; k++) {
        long i = base + k * kTileThreads + tid;
        bool valid = i < n;
        unsigned key = valid ? kin[i] : 0u;
        int val = valid ? (FIRST ? (int)i : vin[i]) : 0;
        int d = (int)((key >> shift) & 255u);
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            unsigned long long bb = __ballot((d >> b) & 1);
            m &= ((d >> b) & 1) ? bb : ~bb;
        }
        int rank = __popcll(m & below);
        if (valid && rank == 0)
            wcnt[w][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            int pos = run[d] + rank;
            for (int q = 0; q < w; q++)
                pos += wcnt[q][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        int add = 0;
        for (int q = 0; q < kTileThreads / 64; q++) {
            add += wcnt[q][tid];
            wcnt[q][tid] = 0;
        }
        run[tid] += add;
        __syncthreads();
    }
}

/* Terms of the sorted samples (skymap.cpp:95-100): cos 2psi, sin 2psi and
 * the datum of sorted position i < *nvalid. */
__global__ void k_mm_terms(MapChunk c, const int *vals, const int *nvalid, double2 *tcs, float *td)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nvalid)
        return;
    long t = c.t0 + vals[i];
    long det = t / c.nsamples;
    long ip = det * c.ptg_stride + (t - det * c.nsamples);
    double psi = -((double)c.pol[det] + (double)c.pa[ip]);
    double s2, c2;
    pl_sincos(2.0 * psi, &s2, &c2);
    tcs[i] = make_double2(c2, s2);
    td[i] = c.data[t];
}

/* One thread per pixel adds its samples' terms, in sample order, onto the
 * pixel's nine sums (skymap.cpp:102-114).  acc: nine FP64 planes of npix. */
__global__ void k_mm_walk(int npix, const int *off, const double2 *tcs, const float *td,
                          double *acc)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix)
        return;
    int b = off[p], e = off[p + 1];
    if (b == e)
        return;
    const long np = npix;
    double n_ = acc[MM_N * np + p], sc = acc[MM_C * np + p], ss = acc[MM_S * np + p];
    double scc = acc[MM_CC * np + p], scs = acc[MM_CS * np + p], sss = acc[MM_SS * np + p];
    double sd = acc[MM_D * np + p], scd = acc[MM_CD * np + p], ssd = acc[MM_SD * np + p];
    /* Double-buffered: the terms of the next U samples are loaded while the
     * current U are added (U = 16 spills and is 2.7x slower).  The sums still take one sample at a time, in
     * order. */
    constexpr int U = 8;
    double2 cur[U], nxt[U];
    float dcur[U], dnxt[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
        int i = min(b + k, e - 1);
        cur[k] = tcs[i];
        dcur[k] = td[i];
    }
    for (int i0 = b; i0 < e; i0 += U) {
#pragma unroll
        for (int k = 0; k < U; k++) {
            int i = min(i0 + U + k, e - 1);
            nxt[k] = tcs[i];
            dnxt[k] = td[i];
        }
#pragma unroll
        for (int k = 0; k < U; k++) {
            if (i0 + k < e) {
                double c2 = cur[k].x, s2 = cur[k].y, d = dcur[k];
                sd += 1.0 * d;
                scd += c2 * d;
                ssd += s2 * d;
                n_ += 1.0;
                sc += c2;
                ss += s2;
                scc += c2 * c2;
                scs += c2 * s2;
                sss += s2 * s2;
            }
        }
#pragma unroll
        for (int k = 0; k < U; k++) {
            cur[k] = nxt[k];
            dcur[k] = dnxt[k];
        }
    }
    acc[MM_N * np + p] = n_;
    acc[MM_C * np + p] = sc;
    acc[MM_S * np + p] = ss;
    acc[MM_CC * np + p] = scc;
    acc[MM_CS * np + p] = scs;
    acc[MM_SS * np + p] = sss;
    acc[MM_D * np + p] = sd;
    acc[MM_CD * np + p] = scd;
    acc[MM_SD * np + p] = ssd;
}

namespace {

/* Swap rows r and q (q >= r) of the column-major 3x3 a when q == cand. */
__device__ inline void swap_rows_if(double a[9], int r, int cand, bool yes)
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double x = a[r + 3 * k], y = a[cand + 3 * k];
        a[r + 3 * k] = yes ? y : x;
        a[cand + 3 * k] = yes ? x : y;
    }
}

/* DGESV, n = 3, nrhs = 1 (Netlib DGETF2 then DGETRS).  a column-major; b is
 * overwritten with the solution unless a pivot is exactly zero (returns
 * non-zero, b untouched). */
__device__ inline int dgesv3(double a[9], double b[3])
{
    const double sfmin = 2.2250738585072014e-308;   /* DLAMCH('S') */
    int info = 0;
    int piv[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        /* IDAMAX over a(j:2, j): first index of the largest magnitude */
        int jp = j;
        double amax = fabs(a[j + 3 * j]);
#pragma unroll
        for (int i = j + 1; i < 3; i++) {
            double v = fabs(a[i + 3 * j]);
            bool gt = v > amax;
            jp = gt ? i : jp;
            amax = gt ? v : amax;
        }
        piv[j] = jp;
        double pv = a[j + 3 * j];
#pragma unroll
        for (int i = j + 1; i < 3; i++)
            pv = (jp == i) ? a[i + 3 * j] : pv;
        if (pv != 0.0) {
#pragma unroll
            for (int i = j + 1; i < 3; i++)
                swap_rows_if(a, j, i, jp == i);
            if (j < 2) {
                double pj = a[j + 3 * j];
                if (fabs(pj) >= sfmin) {
                    double r = 1.0 / pj;
#pragma unroll
                    for (int i = j + 1; i < 3; i++)
                        a[i + 3 * j] *= r;
                } else {
#pragma unroll
                    for (int i = j + 1; i < 3; i++)
                        a[i + 3 * j] = a[i + 3 * j] / pj;
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
        if (j < 2) {
#pragma unroll
            for (int k = j + 1; k < 3; k++) {
                double u = a[j + 3 * k];
                if (u != 0.0) {
                    double temp = -u;
#pragma unroll
                    for (int i = j + 1; i < 3; i++)
                        a[i + 3 * k] = a[i + 3 * k] + a[i + 3 * j] * temp;
                }
            }
        }
    }
    if (info != 0)
        return info;
    /* DLASWP on b */
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int q = i + 1; q < 3; q++) {
            bool yes = piv[i] == q;
            double x = b[i], y = b[q];
            b[i] = yes ? y : x;
            b[q] = yes ? x : y;
        }
    }
    /* DTRSM lower, unit diagonal */
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (b[k] != 0.0) {
#pragma unroll
            for (int i = k + 1; i < 3; i++)
                b[i] = b[i] - b[k] * a[i + 3 * k];
        }
    /* DTRSM upper, non-unit diagonal */
#pragma unroll
    for (int k = 2; k >= 0; k--)
        if (b[k] != 0.0) {
            b[k] = b[k] / a[k + 3 * k];
#pragma unroll
            for (int i = 0; i < k; i++)
                b[i] = b[i] - b[k] * a[i + 3 * k];
        }
    return 0;
}

} // namespace

/* skymap.cpp:157-224 */
__global__ void k_mm_solve(int npix, const double *acc, float *I, float *Q, float *U,
                           int *hits, unsigned long long *singular)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix)
        return;
    const long np = npix;
    double n_ = acc[MM_N * np + p];
    double ii = 0.0, qq = 0.0, uu = 0.0;
    int h = (int)n_;
    if (h >= 3) {
        double sc = acc[MM_C * np + p], ss = acc[MM_S * np + p];
        double a[9] = {n_, sc, ss, sc, acc[MM_CC * np + p], acc[MM_CS * np + p],
                       ss, acc[MM_CS * np + p], acc[MM_SS * np + p]};
        double b[3] = {acc[MM_D * np + p], acc[MM_CD * np + p], acc[MM_SD * np + p]};
        int info = dgesv3(a, b);
        ii = b[0];
        qq = b[1];
        uu = b[2];
        if (info != 0) {
            h = 0;
            atomicAdd(singular, 1ull);
        }
    }
    I[p] = (float)ii;
    Q[p] = (float)qq;
    U[p] = (float)uu;
    hits[p] = h;
}

int map_tiles(int npix) { return (npix + kTile - 1) / kTile; }

namespace {

void launch_scan(const int *count, int n, int *tsum, int *off, hipStream_t st)
{
    const int nt = map_tiles(n);
    hipLaunchKernelGGL(k_mm_tile_sums, dim3(nt), dim3(kTileThreads), 0, st, count, n, tsum);
    hipLaunchKernelGGL(k_mm_scan_sums, dim3(1), dim3(1024), 0, st, tsum, nt, off, n);
    hipLaunchKernelGGL(k_mm_tile_scan, dim3(nt), dim3(kTileThreads), 0, st, count, n, tsum, off,
                       (int *)nullptr);
}

} // namespace

int map_sort_passes(int npix)
{
    int p = 1;
    while (p < 4 && (unsigned long long)npix >= (1ull << (8 * p)))
        p++;
    return p;
}

hipError_t launch_map_accumulate(const Hpx &h, const MapChunk &c, const MapScratch &s,
                                 double *acc, hipStream_t st)
{
    if (c.len <= 0)
        return hipSuccess;
    const int npix = h.npix;
    const int n = c.len;
    const int nb = map_tiles(n);
    hipLaunchKernelGGL(k_mm_fill_i32, dim3(std::min(grid_for(npix, 256), 4096)), dim3(256), 0,
                       st, s.count, (long)npix, 0);
    hipLaunchKernelGGL(k_mm_pixel, dim3(grid_for(n, 256)), dim3(256), 0, st, h, c,
                       (const int *)s.ppix, s.keys[0], s.count, s.skipped);
    launch_scan(s.count, npix, s.tsum, s.off, st);
    /* radix passes: keys[0] (values implicit) -> buffers 1, 2, 1, ... */
    const int passes = map_sort_passes(npix);
    int src = 0;
    for (int pass = 0; pass < passes; pass++) {
        int dst = (pass & 1) ? 2 : 1;
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kTileThreads), 0, st,
                           (const unsigned *)s.keys[src], n, 8 * pass, nb, s.hist);
        launch_scan(s.hist, 256 * nb, s.htsum, s.hoff, st);
        if (pass == 0)
            hipLaunchKernelGGL(k_rs_scatter<true>, dim3(nb), dim3(kTileThreads), 0, st,
                               (const unsigned *)s.keys[src], (const int *)nullptr, n, 8 * pass,
                               nb, (const int *)s.hoff, s.keys[dst], s.vals[dst]);
        else
            hipLaunchKernelGGL(k_rs_scatter<false>, dim3(nb), dim3(kTileThreads), 0, st,
                               (const unsigned *)s.keys[src], (const int *)s.vals[src], n,
                               8 * pass, nb, (const int *)s.hoff, s.keys[dst], s.vals[dst]);
        src = dst;
    }
    hipLaunchKernelGGL(k_mm_terms, dim3(grid_for(n, 256)), dim3(256), 0, st, c,
                       (const int *)s.vals[src], (const int *)(s.off + npix), s.tcs, s.td);
    hipLaunchKernelGGL(k_mm_walk, dim3(grid_for(npix, 256)), dim3(256), 0, st, npix,
                       (const int *)s.off, (const double2 *)s.tcs, (const float *)s.td, acc);
    return hipGetLastError();
}

hipError_t launch_map_ptg_pixel(const Hpx &h, const float *ra, const float *dec, long n,
                                int *ppix, hipStream_t st)
{
    if (n > 0)
        hipLaunchKernelGGL(k_mm_ptg_pixel, dim3(grid_for(n, 256)), dim3(256), 0, st, h, ra, dec,
                           n, ppix);
    return hipGetLastError();
}

hipError_t launch_map_solve(int npix, const double *acc, float *I, float *Q, float *U,
                            int *hits, unsigned long long *singular, hipStream_t st)
{
    hipLaunchKernelGGL(k_mm_solve, dim3(grid_for(npix, 256)), dim3(256), 0, st, npix, acc, I, Q,
                       U, hits, singular);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *x, long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_mm_fill_i32, dim3(grid_for(2 * n, 256)), dim3(256), 0, st, (int *)x,
                       2 * n, 0);
    return hipGetLastError();
}

} /* namespace pisco */
