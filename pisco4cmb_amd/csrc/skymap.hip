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
 *                    exclusive prefix sums of the counts: one segment of the
 *                    index array per pixel
 *   3. k_mm_scatter  each sample's position into its pixel's segment
 *   4. k_mm_accum    one thread per pixel: sorts its segment by sample index
 *                    (insertion sort; the atomics leave it nearly sorted) and
 *                    adds the terms of its samples, in that order, onto the
 *                    pixel's sums
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

/* Pixel of sample j of the chunk: ang2pix(theta = pi/2 - dec, phi = ra) as
 * skymap.cpp:91-96; a theta outside [0, pi] (planck_assert of
 * Healpix_Base::ang2pix) or a non-finite pointing gives -1 and is counted in
 * *skipped. */
__global__ void k_mm_pixel(Hpx h, MapChunk c, int *pix, int *count,
                           unsigned long long *skipped)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= c.len)
        return;
    long t = c.t0 + j;
    long det = t / c.nsamples;
    long ip = det * c.ptg_stride + (t - det * c.nsamples);
    double theta = kHalfPi - (double)c.dec[ip];
    double phi = c.ra[ip];
    int p = -1;
    if (theta >= 0.0 && theta <= kPi && fabs(phi) < INFINITY) {
        p = ang2pix(h, theta, phi);
        if (p < 0 || p >= h.npix)
            p = -1;
    }
    if (p < 0)
        atomicAdd(skipped, 1ull);
    else
        atomicAdd(&count[p], 1);
    pix[j] = p;
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

/* off[i] = cur[i] = exclusive prefix sum of count at i. */
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
            cur[i] = v;
        }
    }
}

__global__ void k_mm_scatter(const int *pix, int len, int *cur, int *idx)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len)
        return;
    int p = pix[j];
    if (p < 0)
        return;
    int pos = atomicAdd(&cur[p], 1);
    idx[pos] = j;
}

/* acc: nine FP64 planes of npix (MapPlane order). */
__global__ void k_mm_accum(MapChunk c, int npix, const int *off, int *idx, double *acc)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix)
        return;
    int b = off[p], e = off[p + 1];
    if (b == e)
        return;
    for (int i = b + 1; i < e; i++) {
        int v = idx[i];
        int k = i - 1;
        while (k >= b) {
            int u = idx[k];
            if (u <= v)
                break;
            idx[k + 1] = u;
            k--;
        }
        idx[k + 1] = v;
    }
    const long np = npix;
    double n_ = acc[MM_N * np + p], sc = acc[MM_C * np + p], ss = acc[MM_S * np + p];
    double scc = acc[MM_CC * np + p], scs = acc[MM_CS * np + p], sss = acc[MM_SS * np + p];
    double sd = acc[MM_D * np + p], scd = acc[MM_CD * np + p], ssd = acc[MM_SD * np + p];
    for (int i = b; i < e; i++) {
        long t = c.t0 + idx[i];
        long det = t / c.nsamples;
        long ip = det * c.ptg_stride + (t - det * c.nsamples);
        /* skymap.cpp:95-114 */
        double psi = -((double)c.pol[det] + (double)c.pa[ip]);
        double d = c.data[t];
        double s2, c2;
        pl_sincos(2.0 * psi, &s2, &c2);
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

hipError_t launch_map_accumulate(const Hpx &h, const MapChunk &c, const MapScratch &s,
                                 double *acc, hipStream_t st)
{
    if (c.len <= 0)
        return hipSuccess;
    const int npix = h.npix;
    const int nt = map_tiles(npix);
    hipLaunchKernelGGL(k_mm_fill_i32, dim3(std::min(grid_for(npix, 256), 4096)), dim3(256), 0,
                       st, s.count, (long)npix, 0);
    hipLaunchKernelGGL(k_mm_pixel, dim3(grid_for(c.len, 256)), dim3(256), 0, st, h, c, s.pix,
                       s.count, s.skipped);
    hipLaunchKernelGGL(k_mm_tile_sums, dim3(nt), dim3(kTileThreads), 0, st, s.count, npix,
                       s.tsum);
    hipLaunchKernelGGL(k_mm_scan_sums, dim3(1), dim3(1024), 0, st, s.tsum, nt, s.off, npix);
    hipLaunchKernelGGL(k_mm_tile_scan, dim3(nt), dim3(kTileThreads), 0, st, s.count, npix,
                       s.tsum, s.off, s.cur);
    hipLaunchKernelGGL(k_mm_scatter, dim3(grid_for(c.len, 256)), dim3(256), 0, st, s.pix,
                       c.len, s.cur, s.idx);
    hipLaunchKernelGGL(k_mm_accum, dim3(grid_for(npix, 256)), dim3(256), 0, st, c, npix, s.off,
                       s.idx, acc);
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
