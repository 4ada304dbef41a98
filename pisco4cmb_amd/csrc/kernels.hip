/*
 * kernels.hip -- gfx950 kernels of the pixel-space beam convolution.
 *
 * Work mapping (CDNA4: 64-lane wavefronts, 4 waves per 256-thread block):
 *   one wavefront per pointing sample.  The wave first runs the ring part of
 *   query_disc with one ring per lane (64 rings per pass), turns the 0-2 pixel
 *   runs of each ring into a compact run table in LDS with a wave prefix sum,
 *   and then walks the flattened visit list lane-strided, so that consecutive
 *   lanes read consecutive sky pixels of a ring (coalesced) while the beam
 *   reads are gathers.  Each visit recomputes the FP64 beam-frame mapping
 *   (rho, sigma, chi -> beam pixel) exactly as the CPU path does; sums are
 *   reduced across the wave with shuffles and written by a single lane, so
 *   the forward is deterministic and needs no atomics.
 *
 * Blocks are remapped so that each of the 8 XCDs (each with its own L2) works
 * on one contiguous stretch of the scan: consecutive samples share most of
 * their disc pixels, and this keeps those reuses inside one L2.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>

#include "device_hpx.h"
#include "kernels.hpp"

namespace pisco {

constexpr int kWaveSize = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWaveSize * kWavesPerBlock;
constexpr int kMaxRuns = 2 * kWaveSize;   /* per ring pass: <= 2 runs per ring */

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* A wave-uniform double moved to scalar registers (one wave = one sample,
 * so the sample's constants need no vector registers). */
__device__ inline double uni(double x)
{
    long long v = __double_as_longlong(x);
    int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffffLL));
    int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ inline int wave_incl_scan(int x, int lane)
{
#pragma unroll
    for (int d = 1; d < kWaveSize; d <<= 1) {
        int y = __shfl_up(x, d, kWaveSize);
        if (lane >= d)
            x += y;
    }
    return x;
}

__device__ inline double wave_sum(double x)
{
#pragma unroll
    for (int m = kWaveSize / 2; m > 0; m >>= 1)
        x += __shfl_xor(x, m, kWaveSize);
    return x;
}

/* Logical block index such that XCD x (hardware dispatches blocks round-robin
 * over the 8 XCDs) receives a contiguous range of logical blocks. */
__device__ inline long xcd_remap(long b, long nblk)
{
    const long nx = 8;
    long q = nblk / nx, r = nblk % nx;
    long x = b % nx, k = b / nx;
    return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

/* Run table of one wave for one pass over (up to) 64 rings, and the wave's
 * queue of visits deferred to the exact FP64 mapping (sky pixel, ring). */
struct RunTable {
    int pref[kMaxRuns + 1];  /* visit offset of each run; pref[nrun] = total */
    int pix0[kMaxRuns];
    int ring[kMaxRuns];
    int qpix[2 * kWaveSize];
    int qring[2 * kWaveSize];
};

/* Builds the run table for rings r0 .. r0+63 (clipped to the disc); returns
 * the number of visits of the pass. */
__device__ inline int build_runs(const Hpx &hs, const Disc &d, int r0, int lane,
                                 RunTable &t)
{
    int r = r0 + lane;
    int p0[2] = {0, 0}, ln[2] = {0, 0};
    int ns = 0;
    if (r <= d.r_last)
        ns = disc_ring(hs, d, r, p0, ln);
    int cnt = ns;
    int tot = (ns > 0 ? ln[0] : 0) + (ns > 1 ? ln[1] : 0);
    int cnt_in = wave_incl_scan(cnt, lane);
    int tot_in = wave_incl_scan(tot, lane);
    int cbase = cnt_in - cnt;
    int off = tot_in - tot;
#pragma unroll
    for (int k = 0; k < 2; k++) {   /* constant indices keep p0, ln in registers */
        if (k < ns) {
            t.pix0[cbase + k] = p0[k];
            t.ring[cbase + k] = r;
            t.pref[cbase + k] = off;
            off += ln[k];
        }
    }
    int nrun = __shfl(cnt_in, kWaveSize - 1, kWaveSize);
    int total = __shfl(tot_in, kWaveSize - 1, kWaveSize);
    if (lane == 0)
        t.pref[nrun] = total;
    wave_sync();
    return total;
}

/* Beam pixel and float cos/sin(2 chi) of the visit of sky pixel pix (ring r,
 * ring record rr) from the sample with beam-centre constants c, sf; or
 * kNeedExact. */
__device__ inline int visit_beam(const PairParams &p, const Centre &c, const SampleF &sf,
                                 const RingRec &rr, int r, int pix, float &c2, float &s2)
{
    double ra_pix = ring_phi(p.hs, rr, r, pix);
    return beam_pixel_f32(p.fc, c, sf, ra_pix, rr.cdp, rr.sdp, c2, s2);
}

/* Wave-uniform double constants of one sample (double-precision mapping). */
struct SampleD {
    double c2psi, s2psi;   /* cos, sin 2 psi */
};

/* A visit whose beam pixel lies inside the beam, with what `use` needs. */
struct Pending {
    int pix, bp, r, sp;
    float c2, s2;
    float4 v0;        /* A0 A1 A2 B0 */
    float2 v1;        /* B1 B2 */
    float2 v2;        /* A3 B3 (forward with the V term only) */
    float sI, sQ, sU; /* sky of the visit (forward only) */
    float sV;         /* sky V of the visit (forward with the V term only) */
    bool live;
};

/* The TOD terms of one visit, a and b: the float expression of
 * convolver.cpp:157-164 and, with VT, the V terms the reference's GPU kernel
 * adds (V A3, V B3; cudaConv2.cu:204,214,217-218), appended last. */
template <bool VT>
__device__ inline void visit_tod(const Pending &e, float &ta, float &tb)
{
    const float A0 = e.v0.x, A1 = e.v0.y, A2 = e.v0.z;
    const float B0 = e.v0.w, B1 = e.v1.x, B2 = e.v1.y;
    const float c2 = e.c2, s2 = e.s2;
    if constexpr (VT) {
        ta = e.sI * (A0) + e.sQ * (A1 * c2 - A2 * s2) + e.sU * (A2 * c2 + A1 * s2) +
             e.sV * e.v2.x;
        tb = e.sI * (B0) + e.sQ * (-B1 * c2 + B2 * s2) + e.sU * (-B2 * c2 - B1 * s2) +
             e.sV * e.v2.y;
    } else {
        ta = e.sI * (A0) + e.sQ * (A1 * c2 - A2 * s2) + e.sU * (A2 * c2 + A1 * s2);
        tb = e.sI * (B0) + e.sQ * (-B1 * c2 + B2 * s2) + e.sU * (-B2 * c2 - B1 * s2);
    }
}

__device__ inline int lanes_below(unsigned long long m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

/* Maps queue entries base .. base+cnt-1 (cnt <= 64), one per lane, with the
 * certified double-precision form (beam_pixel_f64, whose own fallback is the
 * exact FP64 path; the beam pixel is the exact path's either way) and hands
 * the visits inside the beam to use().  Called with every lane of the wave
 * active.  The double-precision form costs about a quarter of the exact
 * path: draining through it measured 2% faster on the 5-degree, NSIDE 512
 * sky adjoint. */
template <bool LOAD_SKY, typename Use, bool VT = false>
__device__ inline void drain_exact(const PairParams &p, RunTable &t, int base, int cnt, int lane,
                                   const Centre &c, double c2psi, double s2psi, Use &use)
{
    wave_sync();
    if (lane < cnt) {
        const int pix = t.qpix[base + lane], r = t.qring[base + lane];
        const RingRec rr = p.rings[r];
        const double ra_pix = ring_phi(p.hs, rr, r, pix);
        ExactVisit e;
        e.bp = beam_pixel_f64(p.hb, p.fc64, c, c2psi, s2psi, ra_pix, rr.cdp, rr.sdp, e.c2, e.s2);
        if ((unsigned)e.bp < (unsigned)p.npix_beam) {
            Pending q;
            q.pix = pix;
            q.r = r;
            q.sp = rr.startpix;
            q.bp = e.bp;
            q.c2 = e.c2;
            q.s2 = e.s2;
            const float *row = p.beam + 8 * (long)e.bp;
            q.v0 = *(const float4 *)row;
            q.v1 = *(const float2 *)(row + 4);
            if constexpr (VT)
                q.v2 = *(const float2 *)(row + 6);
            if (LOAD_SKY) {
                const float4 sv = p.sky4[pix];
                q.sI = sv.x;
                q.sQ = sv.y;
                q.sU = sv.z;
                q.sV = sv.w;
            }
            q.live = true;
            use(q);
        }
    }
    /* Leave no load of this (rare) path outstanding: otherwise the merge of
     * the paths at the caller's loop head looks as if those loads might still
     * be pending, and the compiler would wait for the main path's in-flight
     * loads there too. */
    __builtin_amdgcn_s_waitcnt(0);
    wave_sync();
}

/* The walk over the visits of one ring pass: lane-strided, with every lane
 * running the same number of iterations (the deferred-visit queue is drained
 * by the whole wave).  Software pipelined by hand: the ring record of the
 * lane's next visit is loaded while the current one is computed, and the
 * beam/sky loads of a visit are issued at the end of its iteration and
 * consumed (by `use`) after the next visit's geometry.  The record is used
 * before the next loads are issued into it, so no register copy of a pending
 * load is needed.  In the forward every global load of the main path is
 * issued unconditionally (lanes without a visit load row 0 / pixel 0 and
 * discard it): a load skipped on some path makes the in-order load counter
 * ambiguous at the loop head, and the compiler then waits for all loads
 * there instead of only for the ring record (the adjoints, with fewer loads
 * per visit, measured faster with conditional loads).  Visits whose floors
 * the fast form cannot certify (the float mapping, or the FP64 form with its
 * float alpha) go to the queue (qn entries, wave-uniform); a full batch of 64
 * is drained at once.  use(q) is called once per visit whose beam pixel lies
 * inside the beam. */
template <bool LOAD_SKY, bool MAP64, typename Use, bool VT = false, bool LOAD_BEAM = true>
__device__ inline void walk_visits(const PairParams &p, RunTable &t, int total, int lane,
                                   const Centre &c, const SampleF &sf, const SampleD &sd,
                                   int &qn, Use &use)
{
    if (total <= 0)
        return;
    int s = 0;
    int v = lane;
    if (v < total)
        while (v >= t.pref[s + 1])
            s++;
    RingRec rr = p.rings[t.ring[s]];
    /* complete this first load here: a load still pending on the loop-entry
     * edge would make the compiler wait for the in-flight beam/sky loads at
     * every use of the ring record inside the loop */
    __builtin_amdgcn_s_waitcnt(0);
    Pending q;
    q.live = false;
    for (int v0 = 0; v0 < total; v0 += kWaveSize, v += kWaveSize) {
        const bool act = v < total;
        /* a lane past the end repeats the first visit of its run, discarded */
        const int pix = t.pix0[s] + (act ? v - t.pref[s] : 0);
        const int r = t.ring[s];
        float c2, s2;
        int bp, startpix;
        if constexpr (MAP64) {
            const RingRec cur = rr;
            const int vn = v + kWaveSize;
            if (vn < total)
                while (vn >= t.pref[s + 1])
                    s++;
            rr = p.rings[t.ring[s]];   /* prefetch of the next visit's record */
            startpix = cur.startpix;
            bp = beam_pixel_f64<true>(p.hb, p.fc64, c, sd.c2psi, sd.s2psi,
                                      ring_phi(p.hs, cur, r, pix), cur.cdp, cur.sdp, c2, s2);
        } else {
            /* the part of the geometry that reads the ring record, then the
             * prefetch of the next visit's record into the same registers */
            const VisitRing g = visit_ring_f32(c, ring_phi(p.hs, rr, r, pix), rr.cdp, rr.sdp);
            startpix = rr.startpix;
            const int vn = v + kWaveSize;
            if (vn < total)
                while (vn >= t.pref[s + 1])
                    s++;
            rr = p.rings[t.ring[s]];
            bp = beam_pixel_f32_tail(p.fc, sf, g, c2, s2);
        }
        bp = act ? bp : -1;
        {   /* visits whose floors the fast form cannot certify: queued */
            const unsigned long long m = __ballot(bp == kNeedExact);
            if (m) {
                if (bp == kNeedExact) {
                    const int pos = qn + lanes_below(m);
                    t.qpix[pos] = pix;
                    t.qring[pos] = r;
                }
                qn += __popcll(m);
            }
        }
        if (q.live)
            use(q);
        {
            /* drain here, where the previous visit is consumed and the next
             * loads are not yet issued: little is live across the exact
             * path's out-of-line call */
            if (qn >= kWaveSize) {
                drain_exact<LOAD_SKY, Use, VT>(p, t, qn - kWaveSize, kWaveSize, lane, c, sd.c2psi,
                                               sd.s2psi, use);
                qn -= kWaveSize;
            }
        }
        q.live = (unsigned)bp < (unsigned)p.npix_beam;   /* rejects -1 and kNeedExact */
        q.pix = pix;
        q.r = r;
        q.sp = startpix;
        q.bp = bp;
        q.c2 = c2;
        q.s2 = s2;
        if constexpr (LOAD_SKY) {
            /* forward: loads issued unconditionally (row 0 / pixel 0 when
             * the visit is dead), see above */
            if constexpr (LOAD_BEAM) {
                const float *row = p.beam + 8 * (long)(q.live ? bp : 0);
                q.v0 = *(const float4 *)row;
                q.v1 = *(const float2 *)(row + 4);
                if constexpr (VT)
                    q.v2 = *(const float2 *)(row + 6);
            }
            const int sp = q.live ? pix : 0;
            const float4 sv = p.sky4[sp];
            q.sI = sv.x;
            q.sQ = sv.y;
            q.sU = sv.z;
            q.sV = sv.w;
        } else if constexpr (LOAD_BEAM) {
            if (q.live) {
                /* adjoints: only live visits load (measured faster there) */
                const float *row = p.beam + 8 * (long)bp;
                q.v0 = *(const float4 *)row;
                q.v1 = *(const float2 *)(row + 4);
            }
        }
    }
    if (q.live)
        use(q);
}

/* Wave-uniform copy of a float (to scalar registers). */
__device__ inline float unif(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}

/* Per-sample setup shared by the visit kernels: beam-centre constants in
 * scalar registers and the sample's disc. */
__device__ inline void sample_setup(const PairParams &p, long i, float pa_i, Centre &c,
                                    SampleF &sf, SampleD &sd, Disc &d)
{
    c = centre_make(p.ra[i], p.dec[i], pa_i);
    c.ra = uni(c.ra);
    c.dec = uni(c.dec);
    c.psi = uni(c.psi);
    c.cdc = uni(c.cdc);
    c.sdc = uni(c.sdc);
    double s2psi, c2psi;
    pl_sincos(2.0 * c.psi, &s2psi, &c2psi);
    sd.c2psi = uni(c2psi);
    sd.s2psi = uni(s2psi);
    sf = sample_f32(c, c2psi, s2psi);
    sf.cdc = unif(sf.cdc);
    sf.sdc = unif(sf.sdc);
    sf.t0 = unif(sf.t0);
    sf.c2p = unif(sf.c2p);
    sf.s2p = unif(sf.s2p);
    disc_setup(p.hs, kHalfPi - c.dec, c.ra, p.rho_max, p.cos_rmax, d);
}

/* One sample of one detector pair (pa of the sample given): forward TOD,
 * or the direct-atomic adjoints. */
template <int OP>
__device__ inline void pair_sample(const PairParams &p, long i, float pa_i, RunTable &t,
                                   int lane)
{
    constexpr bool FWD = OP == OP_FWD || OP == OP_FWD_V;
    Centre c;
    SampleF sf;
    SampleD sd;
    Disc d;
    sample_setup(p, i, pa_i, c, sf, sd, d);
    double ga = 0.0, gb = 0.0;
    if (!FWD) {
        ga = uni(p.g_a ? (double)p.g_a[i] : 0.0);
        gb = uni(p.g_b ? (double)p.g_b[i] : 0.0);
    }
    double acc_a = 0.0, acc_b = 0.0;
    auto use = [&](const Pending &q) {
        const float A0 = q.v0.x, A1 = q.v0.y, A2 = q.v0.z;
        const float B0 = q.v0.w, B1 = q.v1.x, B2 = q.v1.y;
        const float c2 = q.c2, s2 = q.s2;
        const int pix = q.pix, bp = q.bp;
        if (FWD) {
            float ta, tb;
            visit_tod<OP == OP_FWD_V>(q, ta, tb);
            acc_a += ta;
            acc_b += tb;
        } else if (OP == OP_ADJ_MAP) {
            double cc = c2, ss = s2;
            double vI = ga * A0 + gb * B0;
            double vQ = ga * (A1 * cc - A2 * ss) + gb * (-B1 * cc + B2 * ss);
            double vU = ga * (A2 * cc + A1 * ss) + gb * (-B2 * cc - B1 * ss);
            unsafeAtomicAdd(&p.gI[pix], vI);
            unsafeAtomicAdd(&p.gQ[pix], vQ);
            unsafeAtomicAdd(&p.gU[pix], vU);
        } else {
            double cc = c2, ss = s2;
            double qq = p.Q[pix], u = p.U[pix], sI = p.I[pix];
            double *g = p.gbeam + 8 * (long)bp;
            unsafeAtomicAdd(&g[0], ga * sI);
            unsafeAtomicAdd(&g[1], ga * (qq * cc + u * ss));
            unsafeAtomicAdd(&g[2], ga * (-qq * ss + u * cc));
            unsafeAtomicAdd(&g[3], gb * sI);
            unsafeAtomicAdd(&g[4], gb * (-qq * cc - u * ss));
            unsafeAtomicAdd(&g[5], gb * (qq * ss - u * cc));
        }
    };

    int qn = 0;   /* deferred visits (wave-uniform) */
    for (int r0 = d.r_first; r0 <= d.r_last; r0 += kWaveSize) {
        int total = build_runs(p.hs, d, r0, lane, t);
        walk_visits<FWD, FWD, decltype(use), OP == OP_FWD_V>(p, t, total, lane, c, sf, sd, qn,
                                                             use);
        wave_sync();
    }
    if (qn > 0)
        drain_exact<FWD, decltype(use), OP == OP_FWD_V>(p, t, 0, qn, lane, c, sd.c2psi, sd.s2psi,
                                                       use);
    if (FWD) {
        acc_a = wave_sum(acc_a);
        acc_b = wave_sum(acc_b);
        if (lane == 0) {
            if (p.tod_a)
                p.tod_a[i] = (float)acc_a;
            if (p.tod_b)
                p.tod_b[i] = (float)acc_b;
        }
    }
}

/* 4 waves per SIMD (128 VGPRs): with the float alpha and its deferral queue
 * the forward needs about 100 VGPRs; under the 96 of 5 waves per SIMD it
 * spilled and ran 15% slower at the 5-degree, NSIDE 512 headline. */
template <int OP>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_pair(PairParams p)
{
    __shared__ RunTable tables[kWavesPerBlock];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWaveSize);
    const int lane = threadIdx.x % kWaveSize;
    long lb = xcd_remap(blockIdx.x, gridDim.x);
    long i = p.s_begin + lb * kWavesPerBlock + wave;
    if (i >= p.s_end)
        return;
    pair_sample<OP>(p, i, p.pa[i], tables[wave], lane);
}

/* ---- small-disc forward: several samples per wave -------------------------
 *
 * For discs of a few hundred pixels the per-sample part of k_pair<OP_FWD>
 * (beam-centre constants, disc set-up, the ring step, the final reduction)
 * runs on all 64 lanes for one sample and costs more than the visits.  Here
 * a wave holds G = 64/SUB samples of SUB lanes each: that part runs once for
 * the G samples, each lane group walks its own disc SUB visits at a time, and
 * a SUB-lane reduction gives each sample's TOD.  The mapping (certified FP64
 * form), the per-visit values and the FP64 accumulation are those of
 * k_pair<OP_FWD>. */
template <int SUB>
struct SubTable {
    int pref[2 * SUB + 1];   /* visit offset of each run; pref[nrun] = total */
    int pix0[2 * SUB];
    int ring[2 * SUB];
};

template <int SUB>
__device__ inline int seg_incl_scan(int x, int sl)
{
#pragma unroll
    for (int d = 1; d < SUB; d <<= 1) {
        int y = __shfl_up(x, d, SUB);
        if (sl >= d)
            x += y;
    }
    return x;
}

__device__ inline int wave_max_uniform(int x)
{
#pragma unroll
    for (int m = kWaveSize / 2; m > 0; m >>= 1)
        x = max(x, __shfl_xor(x, m, kWaveSize));
    return __builtin_amdgcn_readfirstlane(x);
}

/* Builds the lane group's run table for its rings d.r_first + ps*SUB ..
 * + SUB-1 (clipped to its nrings rings); returns the group's visit count. */
template <int SUB>
__device__ inline int build_sub(const Hpx &hs, const Disc &d, int nrings, int ps, int sl,
                                SubTable<SUB> &t)
{
    const int r = d.r_first + ps * SUB + sl;
    int p0[2] = {0, 0}, ln[2] = {0, 0};
    int ns = 0;
    if (ps * SUB + sl < nrings)
        ns = disc_ring(hs, d, r, p0, ln);
    const int tot = (ns > 0 ? ln[0] : 0) + (ns > 1 ? ln[1] : 0);
    const int cnt_in = seg_incl_scan<SUB>(ns, sl);
    const int tot_in = seg_incl_scan<SUB>(tot, sl);
    const int cbase = cnt_in - ns;
    int off = tot_in - tot;
#pragma unroll
    for (int k = 0; k < 2; k++) {   /* constant indices keep p0, ln in registers */
        if (k < ns) {
            t.pix0[cbase + k] = p0[k];
            t.ring[cbase + k] = r;
            t.pref[cbase + k] = off;
            off += ln[k];
        }
    }
    const int nrun = __shfl(cnt_in, SUB - 1, SUB);
    const int total = __shfl(tot_in, SUB - 1, SUB);
    if (sl == 0)
        t.pref[nrun] = total;
    wave_sync();
    return total;
}

/* Per-wave state of the forward's deferred visits in the several-samples-
 * per-wave kernels: the queue (sky pixel, ring, lane group) and, per lane
 * group, its sample's beam-centre constants and the FP64 sums of its
 * deferred visits' TOD. */
template <int SUB>
struct SubDefer {
    static constexpr int G = kWaveSize / SUB;
    int qpix[2 * kWaveSize];
    int qring[2 * kWaveSize];
    int qgrp[2 * kWaveSize];
    double ra[G], psi[G], cdc[G], sdc[G], c2psi[G], s2psi[G];
    double acc[G][2];
};

/* Maps the queued visits base .. base+cnt-1 (cnt <= 64, one per lane) with
 * the FP64-alpha form, each with its own lane group's sample, and adds their
 * TOD to that group's sums (whole wave). */
template <int SUB, bool VT>
__device__ inline void drain_sub(const PairParams &p, SubDefer<SUB> &q, int base, int cnt, int lane)
{
    wave_sync();
    if (lane < cnt) {
        const int pix = q.qpix[base + lane], r = q.qring[base + lane], g = q.qgrp[base + lane];
        Centre c;
        c.ra = q.ra[g];
        c.psi = q.psi[g];
        c.cdc = q.cdc[g];
        c.sdc = q.sdc[g];
        c.dec = 0.0;   /* not read by the mapping */
        const RingRec rr = p.rings[r];
        Pending e;
        e.bp = beam_pixel_f64(p.hb, p.fc64, c, q.c2psi[g], q.s2psi[g],
                              ring_phi(p.hs, rr, r, pix), rr.cdp, rr.sdp, e.c2, e.s2);
        if ((unsigned)e.bp < (unsigned)p.npix_beam) {
            const float *row = p.beam + 8 * (long)e.bp;
            e.v0 = *(const float4 *)row;
            e.v1 = *(const float2 *)(row + 4);
            if constexpr (VT)
                e.v2 = *(const float2 *)(row + 6);
            const float4 sv = p.sky4[pix];
            e.sI = sv.x;
            e.sQ = sv.y;
            e.sU = sv.z;
            e.sV = sv.w;
            float ta, tb;
            visit_tod<VT>(e, ta, tb);
            atomicAdd(&q.acc[g][0], (double)ta);
            atomicAdd(&q.acc[g][1], (double)tb);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);   /* as in drain_exact */
    wave_sync();
}

/* The walk of one ring pass of a lane group: its `total` visits, SUB at a
 * time, for wtot (the wave's largest total) iterations' worth, with the
 * software pipelining of walk_visits (next ring record prefetched; a visit's
 * loads consumed one iteration later) and the certified FP64 mapping.  In the
 * forward (LOAD_SKY) every load of the loop is issued unconditionally and the
 * mapping takes its alpha from a float atan2; visits whose floors that
 * cannot certify go to the wave's queue (dq, qn) and are drained 64 at a
 * time.  The adjoints load the beam row of live visits only.  A lane without
 * a visit computes the first visit of its run (or, in a group with no visit,
 * pixel 0 of ring 1) and discards it.  use(q) is called once per visit whose
 * beam pixel lies inside the beam and is not deferred. */
template <int SUB, bool LOAD_SKY, typename Use, bool VT = false>
__device__ inline void walk_sub(const PairParams &p, const SubTable<SUB> &t, int total,
                                int wtot, int sl, const Centre &c, double c2psi, double s2psi,
                                Use &use, SubDefer<SUB> *dq = nullptr, int *qn = nullptr)
{
    if (wtot <= 0)
        return;
    const bool any = total > 0;
    int s = 0;
    int v = sl;
    if (v < total)
        while (v >= t.pref[s + 1])
            s++;
    RingRec rr = p.rings[any ? t.ring[s] : 1];
    __builtin_amdgcn_s_waitcnt(0);   /* as in walk_visits */
    Pending q;
    q.live = false;
    for (int v0 = 0; v0 < wtot; v0 += SUB, v += SUB) {
        const bool act = v < total;
        const int r = any ? t.ring[s] : 1;
        const int pix = any ? t.pix0[s] + (act ? v - t.pref[s] : 0) : 0;
        const RingRec cur = rr;
        const int vn = v + SUB;
        if (vn < total)
            while (vn >= t.pref[s + 1])
                s++;
        rr = p.rings[any ? t.ring[s] : 1];
        float c2, s2;
        int bp = beam_pixel_f64<LOAD_SKY>(p.hb, p.fc64, c, c2psi, s2psi,
                                          ring_phi(p.hs, cur, r, pix), cur.cdp, cur.sdp, c2, s2);
        bp = act ? bp : -1;
        if constexpr (LOAD_SKY) {   /* float alpha: uncertain floors to the queue */
            const unsigned long long m = __ballot(bp == kNeedExact);
            if (m) {
                if (bp == kNeedExact) {
                    const int pos = *qn + lanes_below(m);
                    dq->qpix[pos] = pix;
                    dq->qring[pos] = r;
                    dq->qgrp[pos] = (int)(__lane_id() / SUB);
                }
                *qn += __popcll(m);
            }
        }
        if (q.live)
            use(q);
        if constexpr (LOAD_SKY) {
            if (*qn >= kWaveSize) {
                drain_sub<SUB, VT>(p, *dq, *qn - kWaveSize, kWaveSize, (int)__lane_id());
                *qn -= kWaveSize;
            }
        }
        q.live = (unsigned)bp < (unsigned)p.npix_beam;
        q.pix = pix;
        q.r = r;
        q.sp = cur.startpix;
        q.bp = bp;
        q.c2 = c2;
        q.s2 = s2;
        if constexpr (LOAD_SKY) {
            const float *row = p.beam + 8 * (long)(q.live ? bp : 0);
            q.v0 = *(const float4 *)row;
            q.v1 = *(const float2 *)(row + 4);
            if constexpr (VT)
                q.v2 = *(const float2 *)(row + 6);
            const float4 sv = p.sky4[q.live ? pix : 0];
            q.sI = sv.x;
            q.sQ = sv.y;
            q.sU = sv.z;
            q.sV = sv.w;
        } else if (q.live) {
            const float *row = p.beam + 8 * (long)bp;
            q.v0 = *(const float4 *)row;
            q.v1 = *(const float2 *)(row + 4);
        }
    }
    if (q.live)
        use(q);
}

/* Per-lane-group sample set-up of the several-samples-per-wave kernels:
 * beam-centre constants, cos/sin 2 psi, the disc and its ring count (0 for a
 * group without a sample). */
__device__ inline int sub_setup(const PairParams &p, long i, bool valid, Centre &c,
                                double &c2psi, double &s2psi, Disc &d)
{
    c = centre_make(p.ra[i], p.dec[i], p.pa[i]);
    pl_sincos(2.0 * c.psi, &s2psi, &c2psi);
    disc_setup(p.hs, kHalfPi - c.dec, c.ra, p.rho_max, p.cos_rmax, d);
    return valid && d.r_last >= d.r_first ? d.r_last - d.r_first + 1 : 0;
}

template <int SUB, bool VT = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_fwd_sub(PairParams p)
{
    constexpr int G = kWaveSize / SUB;
    __shared__ SubTable<SUB> tables[kWavesPerBlock * G];
    __shared__ SubDefer<SUB> defers[kWavesPerBlock];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWaveSize);
    const int lane = threadIdx.x % kWaveSize;
    const int grp = lane / SUB, sl = lane % SUB;
    SubTable<SUB> &t = tables[wave * G + grp];
    SubDefer<SUB> &dq = defers[wave];
    const long lb = xcd_remap(blockIdx.x, gridDim.x);
    const long w0 = p.s_begin + (lb * kWavesPerBlock + wave) * G;   /* the wave's first sample */
    if (w0 >= p.s_end)
        return;   /* wave-uniform */
    const long i = w0 + grp;
    const bool valid = i < p.s_end;
    Centre c;
    double c2psi, s2psi;
    Disc d;
    const int nrings = sub_setup(p, valid ? i : w0, valid, c, c2psi, s2psi, d);
    const int npass = (wave_max_uniform(nrings) + SUB - 1) / SUB;
    if (sl == 0) {   /* the group's constants and sums of deferred visits */
        dq.ra[grp] = c.ra;
        dq.psi[grp] = c.psi;
        dq.cdc[grp] = c.cdc;
        dq.sdc[grp] = c.sdc;
        dq.c2psi[grp] = c2psi;
        dq.s2psi[grp] = s2psi;
        dq.acc[grp][0] = 0.0;
        dq.acc[grp][1] = 0.0;
    }
    wave_sync();
    int qn = 0;   /* queued visits (wave-uniform) */
    double acc_a = 0.0, acc_b = 0.0;
    auto use = [&](const Pending &e) {
        float ta, tb;
        visit_tod<VT>(e, ta, tb);
        acc_a += ta;
        acc_b += tb;
    };
    for (int ps = 0; ps < npass; ps++) {
        const int total = build_sub<SUB>(p.hs, d, nrings, ps, sl, t);
        walk_sub<SUB, true, decltype(use), VT>(p, t, total, wave_max_uniform(total), sl, c, c2psi,
                                               s2psi, use, &dq, &qn);
        wave_sync();
    }
    if (qn > 0)
        drain_sub<SUB, VT>(p, dq, 0, qn, lane);
#pragma unroll
    for (int m = SUB / 2; m > 0; m >>= 1) {
        acc_a += __shfl_xor(acc_a, m, SUB);
        acc_b += __shfl_xor(acc_b, m, SUB);
    }
    if (valid && sl == 0) {
        acc_a += dq.acc[grp][0];
        acc_b += dq.acc[grp][1];
        if (p.tod_a)
            p.tod_a[i] = (float)acc_a;
        if (p.tod_b)
            p.tod_b[i] = (float)acc_b;
    }
}

/* ---- beam interpolation mode ----------------------------------------------
 *
 * The convolution with the beam interpolated at (rho, sigma) from the four
 * get_interpol neighbours that lie inside the beam map, normalised by the
 * sum of their weights: the interpolating form of convolver.cpp:124-147
 * (commented out there), in its precision (float values accumulated from
 * double products, divided by the double weight sum), with the values
 * starting from zero at every visit.  A visit with no neighbour in the beam
 * contributes nothing.  The mapping is the exact FP64 rho_sigma_chi of the
 * sky pixel; one wave per sample walks the disc as k_pair does, without the
 * pipelining of the nearest-pixel kernels.  Sky adjoint and beam adjoint use
 * FP64 atomics; the beam adjoint spreads each visit over the neighbours used
 * with weights w / sum(w). */
template <int OP>
__global__ __launch_bounds__(kBlock) void k_pair_interp(PairParams p)
{
    __shared__ RunTable tables[kWavesPerBlock];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWaveSize);
    const int lane = threadIdx.x % kWaveSize;
    const long lb = xcd_remap(blockIdx.x, gridDim.x);
    const long i = p.s_begin + lb * kWavesPerBlock + wave;
    if (i >= p.s_end)
        return;
    RunTable &t = tables[wave];
    Centre c;
    SampleF sf;
    SampleD sd;
    Disc d;
    sample_setup(p, i, p.pa[i], c, sf, sd, d);
    double ga = 0.0, gb = 0.0;
    if (OP != OP_FWD) {
        ga = p.g_a ? (double)p.g_a[i] : 0.0;
        gb = p.g_b ? (double)p.g_b[i] : 0.0;
    }
    double acc_a = 0.0, acc_b = 0.0;
    for (int r0 = d.r_first; r0 <= d.r_last; r0 += kWaveSize) {
        const int total = build_runs(p.hs, d, r0, lane, t);
        int s = 0;
        for (int v = lane; v < total; v += kWaveSize) {
            while (v >= t.pref[s + 1])
                s++;
            const int pix = t.pix0[s] + (v - t.pref[s]);
            const int r = t.ring[s];
            const RingRec rr = p.rings[r];
            const double ra_pix = ring_phi(p.hs, rr, r, pix);
            double rho, sigma, chi;
            visit_rsc(c, ra_pix, rr.cdp, rr.sdp, rho, sigma, chi);
            int nb[4];
            double w[4];
            get_interpol(p.hb, rho, sigma, nb, w);
            float va[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            double ws = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if ((unsigned)nb[k] < (unsigned)p.npix_beam) {
                    ws += w[k];
                    const float *row = p.beam + 8 * (long)nb[k];
#pragma unroll
                    for (int b = 0; b < 6; b++)
                        va[b] = (float)((double)va[b] + (double)row[b] * w[k]);
                }
            }
            if (!(ws > 0.0))
                continue;
#pragma unroll
            for (int b = 0; b < 6; b++)
                va[b] = (float)((double)va[b] / ws);
            double c2d, s2d;
            pl_sincos(2.0 * chi, &s2d, &c2d);
            const float c2 = (float)c2d, s2 = (float)s2d;
            const float A0 = va[0], A1 = va[1], A2 = va[2];
            const float B0 = va[3], B1 = va[4], B2 = va[5];
            if (OP == OP_FWD) {
                const float sI = p.I[pix], sQ = p.Q[pix], sU = p.U[pix];
                float ta = sI * (A0) + sQ * (A1 * c2 - A2 * s2) + sU * (A2 * c2 + A1 * s2);
                float tb = sI * (B0) + sQ * (-B1 * c2 + B2 * s2) + sU * (-B2 * c2 - B1 * s2);
                acc_a += ta;
                acc_b += tb;
            } else if (OP == OP_ADJ_MAP) {
                double cc = c2, ss = s2;
                unsafeAtomicAdd(&p.gI[pix], ga * A0 + gb * B0);
                unsafeAtomicAdd(&p.gQ[pix], ga * (A1 * cc - A2 * ss) + gb * (-B1 * cc + B2 * ss));
                unsafeAtomicAdd(&p.gU[pix], ga * (A2 * cc + A1 * ss) + gb * (-B2 * cc - B1 * ss));
            } else {
                double cc = c2, ss = s2;
                double qq = p.Q[pix], u = p.U[pix], sI = p.I[pix];
                const double v6[6] = {ga * sI, ga * (qq * cc + u * ss), ga * (-qq * ss + u * cc),
                                      gb * sI, gb * (-qq * cc - u * ss), gb * (qq * ss - u * cc)};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    if ((unsigned)nb[k] < (unsigned)p.npix_beam) {
                        const double f = w[k] / ws;
                        double *g = p.gbeam + 8 * (long)nb[k];
#pragma unroll
                        for (int b = 0; b < 6; b++)
                            unsafeAtomicAdd(&g[b], v6[b] * f);
                    }
                }
            }
        }
        wave_sync();
    }
    if (OP == OP_FWD) {
        acc_a = wave_sum(acc_a);
        acc_b = wave_sum(acc_b);
        if (lane == 0) {
            if (p.tod_a)
                p.tod_a[i] = (float)acc_a;
            if (p.tod_b)
                p.tod_b[i] = (float)acc_b;
        }
    }
}

/* Item order of the batched launches: samples in chunks of kBatchChunk, and
 * within a chunk all items of one pair (or pair group) before the next.  The
 * XCD-aware block order hands each XCD whole chunks, so every XCD sees every
 * pair (the pairs' discs differ widely in size) while consecutive waves
 * share one beam.  Returns false for the padding items of a partial last
 * chunk. */
constexpr long kBatchChunk = 1024;

__device__ inline bool batch_item(long w, long nsamp, int nunits, long &unit, long &sample)
{
    const long per = kBatchChunk * nunits;
    const long c = w / per;
    const long rem = w - c * per;
    const long cs = c * kBatchChunk;
    const long len = (nsamp - cs) < kBatchChunk ? (nsamp - cs) : kBatchChunk;
    if (len <= 0 || rem >= len * nunits)
        return false;
    unit = rem / len;
    sample = cs + (rem - unit * len);
    return true;
}

__host__ inline long batch_items(long nsamp, int nunits)
{
    return (nsamp + kBatchChunk - 1) / kBatchChunk * kBatchChunk * nunits;
}

/* PairParams of one pair of a batch: its beam, beam-map geometry, disc and
 * outputs (TOD rows 2 out, 2 out + 1 of [pairs][2][nsamples]). */
__device__ inline PairParams batch_pair_params(const BatchParams &b, const BatchPair &q)
{
    PairParams p;
    p.hs = b.hs;
    p.rings = b.rings;
    p.sky4 = b.sky4;
    p.I = b.I;
    p.Q = b.Q;
    p.U = b.U;
    p.hb = hpx_make(q.nside);
    p.npix_beam = q.npix;
    p.rho_max = q.rho_max;
    p.cos_rmax = q.cos_rmax;
    p.fc.omc_lo = b.fc.omc_lo;
    p.fc.omc_hi = b.fc.omc_hi;
    p.fc.marg_k = q.marg_k;
    p.fc.nside = q.nside_f;
    p.fc64.crho_big = b.fc64.crho_big;
    p.fc64.crho_small = b.fc64.crho_small;
    p.fc64.marg_k = 100.0 * 4.4e-16 * q.nside * 1.7320508075688772;
    p.beam = q.beam;
    p.beam_last = nullptr;
    /* a pointing per pair (ptg_stride > 0) is found by the pair's caller index */
    const long po = (long)(q.out < 0 ? 0 : q.out) * b.ptg_stride;
    p.ra = b.ra + po;
    p.dec = b.dec + po;
    p.pa = b.pa + po;
    p.s_begin = b.s_begin;
    p.s_end = b.s_end;
    const long row = 2 * (long)(q.out < 0 ? 0 : q.out) * b.nsamples;
    p.tod_a = b.tod ? b.tod + row : nullptr;
    p.tod_b = b.tod ? b.tod + row + b.nsamples : nullptr;
    p.g_a = b.g ? b.g + row : nullptr;
    p.g_b = b.g ? b.g + row + b.nsamples : nullptr;
    p.gI = b.gI;
    p.gQ = b.gQ;
    p.gU = b.gU;
    p.gbeam = nullptr;
    return p;
}

/* Batched form of k_pair: one wave per (pair, sample) item in the order of
 * batch_item; each item is exactly the single-pair computation with
 * pa = (float)(pa + pa_offset). */
template <int OP>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5, 8)))
void k_batch_pair(BatchParams b)
{
    __shared__ RunTable tables[kWavesPerBlock];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWaveSize);
    const int lane = threadIdx.x % kWaveSize;
    const long lb = xcd_remap(blockIdx.x, gridDim.x);
    const long w = lb * kWavesPerBlock + wave;
    long k, j;
    if (!batch_item(w, b.s_end - b.s_begin, b.npairs, k, j))
        return;
    const BatchPair &q = b.pairs[k];
    if (q.out < 0)
        return;   /* inert padding */
    const long i = b.s_begin + j;
    const PairParams p = batch_pair_params(b, q);
    pair_sample<OP>(p, i, p.pa[i] + q.pa_off, tables[wave], lane);
}

/* ---- batched forward: many detector pairs on one pointing ----------------
 *
 * Detector pairs that share the boresight pointing (each with its own beam,
 * disc radius and position-angle offset) are convolved in one launch.  A wave
 * takes one (sample, group of G pairs), in the item order of batch_item; the
 * pairs are sorted by disc radius,
 * so the group's union disc is the disc of its first pair.  Lanes walk the
 * visits of that disc as in k_pair; per visit the parts that do not depend on
 * the pair (sky values, the float geometry: 1 - cos rho, alpha, the chi
 * phase) are computed once and each pair runs only its tail: disc membership
 * (its own rho_max), beam pixel for its position angle, beam gather, sum.
 * Membership is decided in float away from the disc edge and by the exact
 * query_disc ring test (disc_ring at the pair's radius) within 4e-6 of it;
 * like uncertain beam pixels, such visits go to the per-wave queue and are
 * evaluated exactly, their sums kept in LDS.  The pair table is padded to a
 * multiple of G with inert pairs (npix 0, never members). */

/* Per-wave LDS of a group of G pairs; the queue holds < 64 entries between
 * drains and one visit adds at most G x 64. */
template <int G>
struct BatchTable {
    RunTable rt;                       /* run table (its own queue is not used) */
    int qpix[(G + 1) * kWaveSize];     /* queue: sky pixel, ring, pair in the group */
    int qring[(G + 1) * kWaveSize];
    int qpair[(G + 1) * kWaveSize];
    double acc[G][2];                  /* forward: exact-path sums of each pair */
    double psi[G];                     /* psi of each pair at this sample */
    float ga[G], gb[G];                /* adjoint: cotangents of each pair */
};

/* Exact mapping of a queued visit for one pair: membership by the query_disc
 * ring test at the pair's radius, then the exact beam pixel and chi; the
 * visit's beam row (and its ring's first pixel) when it is inside the beam. */
__device__ inline bool batch_exact_visit(const BatchParams &p, const BatchPair &bp_,
                                         const Centre &c, double psi, int pix, int r,
                                         ExactVisit &e, float4 &v0, float2 &v1, int &sp)
{
    Disc dk;
    disc_setup(p.hs, kHalfPi - c.dec, c.ra, bp_.rho_max, bp_.cos_rmax, dk);
    if (r < dk.r_first || r > dk.r_last)
        return false;
    int p0[2] = {0, 0}, ln[2] = {0, 0};
    const int nrun = disc_ring(p.hs, dk, r, p0, ln);
    bool in = false;
#pragma unroll
    for (int k = 0; k < 2; k++)
        in = in || (k < nrun && (unsigned)(pix - p0[k]) < (unsigned)ln[k]);
    if (!in)
        return false;
    const RingRec rr = p.rings[r];
    const double ra_pix = ring_phi(p.hs, rr, r, pix);
    e = beam_visit_exact(bp_.nside, c.ra, psi, c.cdc, c.sdc, ra_pix, rr.cdp, rr.sdp);
    if ((unsigned)e.bp >= (unsigned)bp_.npix)
        return false;
    const float *row = bp_.beam + 8 * (long)e.bp;
    v0 = *(const float4 *)row;
    v1 = *(const float2 *)(row + 4);
    sp = rr.startpix;
    return true;
}

/* Wave-uniform constants of the pairs d0 .. d0+G-1 at one sample. */
template <int G>
struct BatchConsts {
    float T0[G], C2P[G], S2P[G];   /* (pi - psi) 2/pi mod 4, cos/sin 2 psi */
    float GA[G], GB[G];            /* adjoint cotangents */
    float NSF[G], MK[G], CUT_IN[G], CUT_OUT[G];
    int NPIX[G];
    const float *BEAM[G];
};

template <int G, bool ADJ>
__device__ inline void batch_consts(const BatchParams &p, BatchTable<G> &bt, long i, int d0,
                                    int lane, BatchConsts<G> &K)
{
    float t0l = 0.f, c2l = 1.f, s2l = 0.f, gal = 0.f, gbl = 0.f;
    if (lane < G) {   /* lane k: pair d0 + k */
        const BatchPair &q = p.pairs[d0 + lane];
        const float pa_k = p.pa[i] + q.pa_off;   /* the pair's own pa */
        const double psi = -(double)pa_k;
        double s2, c2;
        pl_sincos(2.0 * psi, &s2, &c2);
        t0l = (float)fmodulo((kPi - psi) * kInvHalfPi, 4.0);
        c2l = (float)c2;
        s2l = (float)s2;
        if (ADJ && q.out >= 0) {
            const long row = 2 * (long)q.out * p.nsamples;
            gal = p.g[row + i];
            gbl = p.g[row + p.nsamples + i];
        }
        bt.psi[lane] = psi;
        bt.acc[lane][0] = 0.0;
        bt.acc[lane][1] = 0.0;
        bt.ga[lane] = gal;
        bt.gb[lane] = gbl;
    }
#pragma unroll
    for (int k = 0; k < G; k++) {
        K.T0[k] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, t0l), k));
        K.C2P[k] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, c2l), k));
        K.S2P[k] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, s2l), k));
        K.GA[k] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, gal), k));
        K.GB[k] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, gbl), k));
        const BatchPair &q = p.pairs[d0 + k];
        K.NSF[k] = q.nside_f;
        K.MK[k] = q.marg_k;
        K.CUT_IN[k] = q.omc_cut * (1.0f - 4e-6f);
        K.CUT_OUT[k] = q.omc_cut * (1.0f + 4e-6f);
        K.NPIX[k] = q.npix;
        K.BEAM[k] = q.beam;
    }
    wave_sync();
}

/* A visit of the group waiting for its loads: per pair the beam row and
 * chi, and the visit's sky values / pixel. */
template <int G>
struct BatchPend {
    float4 q0[G];
    float2 q1[G];
    float qc[G], qs[G];
    float sI, sQ, sU;
    int pix, r, sp;
    unsigned live;   /* bit k: pair k has this visit inside its beam */
};

/* The walk over the visits of the group's union disc (the disc of its first
 * pair), pipelined as walk_visits: per visit the shared float geometry, then
 * each pair's tail (membership, beam pixel, gather); consume(q) receives a
 * visit once its loads are due (after the next visit's geometry).  Queued
 * visits (uncertain membership or beam pixel) are drained 64 at a time by the
 * whole wave; exact(k, pix, r, sp, e, v0, v1) receives those inside pair k's
 * disc and beam. */
template <int G, bool LOAD_SKY, typename Consume, typename Exact>
__device__ inline void batch_walk(const BatchParams &p, BatchTable<G> &bt, const Disc &d, int lane,
                                  const Centre &c, const SampleF &sf, const BatchConsts<G> &K,
                                  int d0, Consume &consume, Exact &exact)
{
    RunTable &t = bt.rt;
    int qn = 0;
    auto drain = [&](int base, int cnt) {
        wave_sync();
        if (lane < cnt) {
            const int pix = bt.qpix[base + lane], r = bt.qring[base + lane];
            const int k = bt.qpair[base + lane];
            ExactVisit e;
            float4 v0;
            float2 v1;
            int sp;
            if (batch_exact_visit(p, p.pairs[d0 + k], c, bt.psi[k], pix, r, e, v0, v1, sp))
                exact(k, pix, r, sp, e, v0, v1);
        }
        __builtin_amdgcn_s_waitcnt(0);
        wave_sync();
    };
    for (int r0 = d.r_first; r0 <= d.r_last; r0 += kWaveSize) {
        const int total = build_runs(p.hs, d, r0, lane, t);
        if (total <= 0)
            continue;
        int s = 0;
        int v = lane;
        if (v < total)
            while (v >= t.pref[s + 1])
                s++;
        RingRec rr = p.rings[t.ring[s]];
        __builtin_amdgcn_s_waitcnt(0);
        BatchPend<G> q;
        q.live = 0;
        q.sI = q.sQ = q.sU = 0.f;
        q.pix = q.r = q.sp = 0;
        for (int v0 = 0; v0 < total; v0 += kWaveSize, v += kWaveSize) {
            const bool act = v < total;
            const int pix = t.pix0[s] + (act ? v - t.pref[s] : 0);
            const int r = t.ring[s];
            const VisitRing g = visit_ring_f32(c, ring_phi(p.hs, rr, r, pix), rr.cdp, rr.sdp);
            const int startpix = rr.startpix;
            const int vn = v + kWaveSize;
            if (vn < total)
                while (vn >= t.pref[s + 1])
                    s++;
            rr = p.rings[t.ring[s]];
            const VisitGeo vg = visit_geo_f32(p.fc, sf, g);

            if (q.live)
                consume(q);
            /* the only drain site of the loop: the previous visit is consumed
             * and the next loads are not issued yet */
            while (qn >= kWaveSize) {
                drain(qn - kWaveSize, kWaveSize);
                qn -= kWaveSize;
            }
            /* this visit, per pair (the queue holds < 64 entries here and
             * grows by at most G x 64 below) */
            q.live = 0;
#pragma unroll
            for (int k = 0; k < G; k++) {
                const bool in = act && vg.omc < K.CUT_IN[k];
                const bool edge = act && !in && !(vg.omc > K.CUT_OUT[k]);
                float c2 = 0.f, s2 = 0.f;
                int bp = -1;
                if (in)
                    bp = beam_pixel_f32_det(vg, K.T0[k], K.C2P[k], K.S2P[k], K.NSF[k], K.MK[k],
                                            c2, s2);
                const bool defer = edge || bp == kNeedExact;
                const unsigned long long m = __ballot(defer);
                if (m) {
                    if (defer) {
                        const int pos = qn + lanes_below(m);
                        bt.qpix[pos] = pix;
                        bt.qring[pos] = r;
                        bt.qpair[pos] = k;
                    }
                    qn += __popcll(m);
                }
                const bool live = (unsigned)bp < (unsigned)K.NPIX[k];
                q.live |= live ? (1u << k) : 0u;
                q.qc[k] = c2;
                q.qs[k] = s2;
                const float *row = K.BEAM[k] + 8 * (long)(live ? bp : 0);
                q.q0[k] = *(const float4 *)row;
                q.q1[k] = *(const float2 *)(row + 4);
            }
            q.pix = pix;
            q.r = r;
            q.sp = startpix;
            if (LOAD_SKY) {
                const float4 sv = p.sky4[act ? pix : 0];
                q.sI = sv.x;
                q.sQ = sv.y;
                q.sU = sv.z;
            }
        }
        if (q.live)
            consume(q);
        wave_sync();
    }
    while (qn > 0) {
        const int cnt = qn < kWaveSize ? qn : kWaveSize;
        drain(qn - cnt, cnt);
        qn -= cnt;
    }
}

template <int G>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
void k_batch_fwd(BatchParams p)
{
    __shared__ BatchTable<G> tabs[kWavesPerBlock];
    /* each lane's FP64 TOD sums of the group's pairs in LDS, one slot per
     * lane (the same additions in the same order as in registers).  In
     * registers they spilled, and the spill traffic went through the vector
     * memory path this gather-bound kernel is limited by: 49.1 -> 40.8 ms at
     * 32 pairs of FWHM 0.2-1 deg, beam NSIDE 2048, 2.5e5 samples, even with
     * the LDS now allowing three waves per SIMD instead of four. */
    __shared__ double accb_s[kWavesPerBlock][G][kWaveSize];
    __shared__ double acca_s[kWavesPerBlock][G][kWaveSize];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWaveSize);
    const int lane = threadIdx.x % kWaveSize;
    BatchTable<G> &bt = tabs[wave];
    double (*accb)[kWaveSize] = accb_s[wave];
    double (*acca)[kWaveSize] = acca_s[wave];

    const long lb = xcd_remap(blockIdx.x, gridDim.x);
    const long w = lb * kWavesPerBlock + wave;
    long grp, j;
    if (!batch_item(w, p.s_end - p.s_begin, p.ngroups, grp, j))
        return;
    const long i = p.s_begin + j;
    const int d0 = (int)grp * G;

    Centre c = centre_make(p.ra[i], p.dec[i], p.pa[i]);
    c.ra = uni(c.ra);
    c.dec = uni(c.dec);
    c.cdc = uni(c.cdc);
    c.sdc = uni(c.sdc);
    SampleF sf;
    sf.cdc = unif((float)c.cdc);
    sf.sdc = unif((float)c.sdc);
    BatchConsts<G> K;
    batch_consts<G, false>(p, bt, i, d0, lane, K);
    Disc d;
    disc_setup(p.hs, kHalfPi - c.dec, c.ra, p.pairs[d0].rho_max, p.pairs[d0].cos_rmax, d);

#pragma unroll
    for (int k = 0; k < G; k++) {
        acca[k][lane] = 0.0;
        accb[k][lane] = 0.0;
    }
    auto consume = [&](const BatchPend<G> &q) {
#pragma unroll
        for (int k = 0; k < G; k++) {
            if (q.live & (1u << k)) {
                const float A0 = q.q0[k].x, A1 = q.q0[k].y, A2 = q.q0[k].z, B0 = q.q0[k].w;
                const float B1 = q.q1[k].x, B2 = q.q1[k].y, cc = q.qc[k], ss = q.qs[k];
                const float ta = q.sI * A0 + q.sQ * (A1 * cc - A2 * ss) + q.sU * (A2 * cc + A1 * ss);
                acca[k][lane] = acca[k][lane] + (double)ta;
                const float tb = q.sI * B0 + q.sQ * (-B1 * cc + B2 * ss) + q.sU * (-B2 * cc - B1 * ss);
                accb[k][lane] = accb[k][lane] + (double)tb;
            }
        }
    };
    auto exact = [&](int k, int pix, int r, int sp, const ExactVisit &e, const float4 &v0,
                     const float2 &v1) {
        (void)r;
        (void)sp;
        const float4 sv = p.sky4[pix];
        const float sI = sv.x, sQ = sv.y, sU = sv.z;
        const float ta =
            sI * v0.x + sQ * (v0.y * e.c2 - v0.z * e.s2) + sU * (v0.z * e.c2 + v0.y * e.s2);
        const float tb =
            sI * v0.w + sQ * (-v1.x * e.c2 + v1.y * e.s2) + sU * (-v1.y * e.c2 - v1.x * e.s2);
        atomicAdd(&bt.acc[k][0], (double)ta);
        atomicAdd(&bt.acc[k][1], (double)tb);
    };
    batch_walk<G, true>(p, bt, d, lane, c, sf, K, d0, consume, exact);
    wave_sync();
#pragma unroll
    for (int k = 0; k < G; k++) {
        const double ta = wave_sum(acca[k][lane]) + bt.acc[k][0];
        const double tb = wave_sum(accb[k][lane]) + bt.acc[k][1];
        const int out = p.pairs[d0 + k].out;
        if (lane == 0 && out >= 0) {
            p.tod[(2 * (long)out) * p.nsamples + i] = (float)ta;
            p.tod[(2 * (long)out + 1) * p.nsamples + i] = (float)tb;
        }
    }
}

/* ---- adjoint with respect to the sky, LDS-privatised ---------------------
 *
 * A block (16 waves) owns a chunk of K consecutive samples.  Their discs
 * overlap almost entirely, so the block first computes, per ring, the pixel
 * window that covers every disc of the chunk (pass 1: ring spans unwrapped
 * around the chunk's first centre, atomic min/max in LDS), lays the windows
 * out contiguously in LDS (prefix sum), accumulates every visit's I,Q,U
 * gradient into LDS with ds_add_f32 (pass 2: wave per sample, as the
 * forward), and finally adds the non-zero window entries to the FP64 global
 * gradient (one global atomic per pixel and chunk instead of one per visit).
 * A chunk whose windows do not fit falls back to direct global atomics. */

struct AdjLayout {
    int chunk;       /* samples per chunk */
    int cap;         /* LDS pixel capacity of the window */
    int rw;          /* ring capacity of the window tables */
    long nchunks;
    /* batch mode (pairs != NULL): every pair of b.pairs adds into the same
     * windows; the windows use radius rho_w (the pairs' union support) */
    const BatchPair *pairs;
    int npairs;
    double rho_w, cos_w;
    BatchParams b;
    /* deterministic mode: contributions are rounded to multiples of
     * 1 / *fx_scale and summed as 64-bit integers (LDS windows and the global
     * accumulators, whose FP64 planes then hold int64 bits) */
    const double *fx_scale;
};

#ifndef PISCO_ADJ_WAVES
#define PISCO_ADJ_WAVES 16
#endif
constexpr int kAdjWaves = PISCO_ADJ_WAVES;   /* waves per block of the sky adjoint */
constexpr int kAdjBlock = kAdjWaves * kWaveSize;

/* rint(v * s) as a two's-complement 64-bit integer, for |v * s| < 2^51 and s
 * a power of two (v * s is then exact): adding 1.5 * 2^52 rounds it to the
 * nearest integer, ties to even as rint does (1.5 * 2^52 is even), and
 * leaves that integer in the low bits of the sum's mantissa, whose exponent
 * is fixed.  Two instructions instead of rint and a 64-bit conversion. */
__device__ inline unsigned long long fix64(float v, double s)
{
    const double t = fma((double)v, s, 0x1.8p52);
    return (unsigned long long)__double_as_longlong(t) - 0x4338000000000000ull;
}
constexpr int kBigI = 0x3fffffff;
/* A chunk whose window exceeds the LDS capacity is split into at most this
 * many slabs of consecutive window rings (see k_adjmap_lds). */
constexpr int kMaxSlabs = 63;
constexpr int kAdjCtl = 4 + kMaxSlabs + 1;   /* ints of the control block */

/* I,Q,U gradient of one visit, in float: the values are accumulated in float
 * LDS windows, so FP64 products would be rounded away there anyway. */
__device__ inline void adj_visit_values(float ga, float gb, float4 v0, float2 v1, float c2,
                                        float s2, float &vI, float &vQ, float &vU)
{
    float A0 = v0.x, A1 = v0.y, A2 = v0.z, B0 = v0.w, B1 = v1.x, B2 = v1.y;
    vI = ga * A0 + gb * B0;
    vQ = ga * (A1 * c2 - A2 * s2) + gb * (-B1 * c2 + B2 * s2);
    vU = ga * (A2 * c2 + A1 * s2) + gb * (-B2 * c2 - B1 * s2);
}

/* Per-wave LDS table of the sky adjoint: a run table, or the batch table of
 * a group of BG >= 2 pairs. */
template <int BG>
struct AdjWaveTable {
    using type = BatchTable<BG>;
};
template <>
struct AdjWaveTable<0> {
    using type = RunTable;
};
template <>
struct AdjWaveTable<1> {
    using type = RunTable;
};

/* BG: 0 one pair; 1 a batch, one wave per (pair, sample); 2, 4 a batch, one
 * wave per (group of BG pairs, sample) sharing each visit's geometry.
 * DET: deterministic mode: every contribution (with BG >= 2, each visit's
 * sum over the group's pairs, formed in pair order) is rounded once to a
 * multiple of 1 / scale and all sums are 64-bit integer sums, which do not
 * depend on the order of the additions; the result is the same bits for any
 * schedule and launch shape (for a given grouping of the pairs). */
template <int BG, bool DET = false, int SUB = kWaveSize>
__global__ __launch_bounds__(kAdjBlock) void k_adjmap_lds(PairParams p, AdjLayout L)
{
    static_assert(SUB == kWaveSize || (BG == 0 && sizeof(SubTable<SUB>) * (kWaveSize / SUB) <=
                                                      sizeof(RunTable)),
                  "several samples per wave: single pair, tables within the run table");
    constexpr bool BATCH = BG > 0;
    using WT = typename AdjWaveTable<BG>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    WT *wtab = (WT *)smem;
    int *win_lo = (int *)(wtab + kAdjWaves);
    int *win_hi = win_lo + L.rw;
    int *win_base = win_hi + L.rw;          /* rw + 1 */
    int *ctl = win_base + L.rw + 1;         /* 0 R0, 1 R1, 2 W, 3 slabs, 4.. slab starts */
    int *slab = ctl + 4;                    /* first window ring of each slab, then nring */
    float *acc = (float *)(ctl + kAdjCtl);
    /* deterministic mode: int64 windows, 8-byte aligned after ctl.  Formed
     * by indexing smem (16-byte aligned), not through an integer: a pointer
     * that passes through an integer loses its LDS address space, and its
     * atomics compile to flat instead of ds_add_u64 (which measured the
     * same here: the window's adds are off the critical path). */
    const size_t acc_off = (size_t)((unsigned char *)(ctl + kAdjCtl) - smem);
    unsigned long long *acc64 = (unsigned long long *)(smem + ((acc_off + 7) & ~(size_t)7));
    const double fxs = DET ? *L.fx_scale : 0.0;

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWaveSize);
    const int lane = tid % kWaveSize;
    RunTable &t = *(RunTable *)&wtab[wave];   /* the first member of either table */

    /* Window radius.  A visit adds a non-zero value only if its beam pixel
     * is at or before row `last`, i.e. lies in a beam ring <= R of that row.
     * A point in a ring-R pixel has colatitude below the centre colatitude of
     * ring R+1 (pixel vertices lie on the neighbouring rings), so every
     * contributing sky pixel is within theta(R+1) of the sample; the windows
     * use theta(R+2), one more ring of margin.  Visits outside a window still
     * take the global-atomic path below, so the windows only need to be
     * right for speed, not for the result. */
    double rho_w = p.rho_max, cos_w = p.cos_rmax;
    if (BATCH) {
        rho_w = L.rho_w;
        cos_w = L.cos_w;
    } else if (p.beam_last) {
        int last = __builtin_amdgcn_readfirstlane(*p.beam_last);
        int rq = (last < 0 ? 0 : pix2ring(p.hb, last)) + 2;
        if (rq < 4 * p.hb.nside) {
            double th = ring_theta(p.hb, rq);
            if (th < rho_w) {
                rho_w = th;
                cos_w = pl_cos(th);
            }
        }
    }

    long lb = xcd_remap(blockIdx.x, gridDim.x);
    for (long ch = lb; ch < L.nchunks; ch += gridDim.x) {
        long s0 = p.s_begin + ch * L.chunk;
        long s1 = s0 + L.chunk < p.s_end ? s0 + L.chunk : p.s_end;
        int ns = (int)(s1 - s0);

        /* pass 0: ring range of the chunk */
        if (tid == 0) {
            ctl[0] = kBigI;
            ctl[1] = -1;
        }
        __syncthreads();
        for (int j = tid; j < ns; j += kAdjBlock) {
            Centre c = centre_make(p.ra[s0 + j], p.dec[s0 + j], 0.f);
            Disc d;
            disc_setup(p.hs, kHalfPi - c.dec, c.ra, rho_w, cos_w, d);
            if (d.r_first <= d.r_last) {
                atomicMin(&ctl[0], d.r_first);
                atomicMax(&ctl[1], d.r_last);
            }
        }
        __syncthreads();
        const int R0 = __builtin_amdgcn_readfirstlane(ctl[0]);
        const int nring = __builtin_amdgcn_readfirstlane(ctl[1]) - R0 + 1;
        /* no window ring (a disc of radius rho_w can miss every ring centre
         * while the rho_max discs still have visits) or too many: global
         * atomics for this chunk */
        bool fallback = nring <= 0 || nring > L.rw;
        double phi_ref;
        {
            Centre c = centre_make(p.ra[s0], p.dec[s0], 0.f);
            Disc d;
            disc_setup(p.hs, kHalfPi - c.dec, c.ra, rho_w, cos_w, d);
            phi_ref = d.phi;
        }
        if (!fallback) {
            for (int k = tid; k < nring; k += kAdjBlock) {
                win_lo[k] = kBigI;
                win_hi[k] = -kBigI;
            }
            __syncthreads();
            /* pass 1: windows (SUB lanes per sample, 64/SUB samples per wave) */
            for (int j0 = wave * (kWaveSize / SUB); j0 < ns; j0 += kAdjWaves * (kWaveSize / SUB)) {
                const int j = j0 + lane / SUB;
                Centre c = centre_make(p.ra[s0 + (j < ns ? j : j0)], p.dec[s0 + (j < ns ? j : j0)],
                                       0.f);
                Disc d;
                disc_setup(p.hs, kHalfPi - c.dec, c.ra, rho_w, cos_w, d);
                const int r_end = j < ns ? d.r_last : d.r_first - 1;
                for (int r = d.r_first + lane % SUB; r <= r_end; r += SUB) {
                    int lo, hi, ipix1, nr;
                    int kind = disc_ring_span(p.hs, d, r, lo, hi, ipix1, nr);
                    if (kind == 0)
                        continue;
                    if (kind == 2) {
                        lo = -kBigI;
                        hi = kBigI;
                    } else {
                        double ref = nr * kInvTwoPi * phi_ref;
                        double mid = 0.5 * (lo + hi);
                        int sh = (int)rint((ref - mid) / nr);
                        lo += sh * nr;
                        hi += sh * nr;
                    }
                    atomicMin(&win_lo[r - R0], lo);
                    atomicMax(&win_hi[r - R0], hi);
                }
            }
            __syncthreads();
            /* finalise: window start (mod nr) and length per ring; prefix sum */
            if (wave == 0) {
                int carry = 0;
                for (int k0 = 0; k0 < nring; k0 += kWaveSize) {
                    int k = k0 + lane;
                    int len = 0;
                    if (k < nring) {
                        int lo = win_lo[k], hi = win_hi[k];
                        int sp, nr;
                        bool shf;
                        ring_info(p.hs, R0 + k, sp, nr, shf);
                        if (lo == kBigI) {
                            len = 0;
                            lo = 0;
                        } else if (lo <= -kBigI / 2 || hi - lo + 1 >= nr) {
                            len = nr;
                            lo = 0;
                        } else {
                            len = hi - lo + 1;
                            lo %= nr;
                            if (lo < 0)
                                lo += nr;
                        }
                        win_lo[k] = lo;
                        win_hi[k] = len;
                    }
                    int inc = wave_incl_scan(len, lane);
                    if (k < nring)
                        win_base[k] = carry + inc - len;
                    carry += __shfl(inc, kWaveSize - 1, kWaveSize);
                }
                if (lane == 0) {
                    win_base[nring] = carry;
                    ctl[2] = carry;
                }
            }
            __syncthreads();
            /* slabs: consecutive window rings whose windows add up to at
             * most cap pixels (one slab, the whole window, when it fits) */
            if (tid == 0) {
                const int W = ctl[2];
                int n = 1;
                slab[0] = 0;
                if (W > L.cap) {
                    int k0 = 0;
                    n = 0;
                    while (k0 < nring && n < kMaxSlabs) {
                        int k1 = k0 + 1;
                        while (k1 < nring && win_base[k1 + 1] - win_base[k0] <= L.cap)
                            k1++;
                        slab[n++] = k0;
                        k0 = k1;
                    }
                    /* beyond kMaxSlabs slabs the last one takes the remaining
                     * rings: it exceeds cap, and its visits use global atomics */
                }
                slab[n] = nring;
                ctl[3] = n;
            }
            __syncthreads();
        }
        /* block-uniform slab bounds, read from LDS into scalar registers */
        const int nslab = fallback ? 1 : __builtin_amdgcn_readfirstlane(ctl[3]);
        for (int si = 0; si < nslab; si++) {
        /* this slab: window rings [k0, k1) and their Ws window pixels; the
         * first slab also walks the disc rings before the window rings and
         * the last slab those after (their visits take global atomics) */
        const int k0 = fallback ? 0 : __builtin_amdgcn_readfirstlane(slab[si]);
        const int k1 = fallback ? 0 : __builtin_amdgcn_readfirstlane(slab[si + 1]);
        const int base0 = fallback ? 0 : __builtin_amdgcn_readfirstlane(win_base[k0]);
        const int Ws = fallback ? 0 : __builtin_amdgcn_readfirstlane(win_base[k1]) - base0;
        const bool sfall = fallback || Ws > L.cap;
        const int rlo = (fallback || si == 0) ? -kBigI : R0 + k0;
        const int rhi = (fallback || si == nslab - 1) ? kBigI : R0 + k1 - 1;
        auto clamp = [&](Disc &d) {
            d.r_first = d.r_first > rlo ? d.r_first : rlo;
            d.r_last = d.r_last < rhi ? d.r_last : rhi;
        };
        if (!sfall) {
            for (int k = tid; k < 3 * Ws; k += kAdjBlock) {
                if constexpr (DET)
                    acc64[k] = 0ull;
                else
                    acc[k] = 0.f;
            }
        }
        __syncthreads();

        /* pass 2: visits; adds I,Q,U of one visit to the LDS window (or, when no
         * window holds the pixel, to the global FP64 gradient) */
        const int W = sfall ? 0 : Ws;
        auto scatter = [&](int pix, int r, int sp, float vI, float vQ, float vU) {
            if (vI == 0.f && vQ == 0.f && vU == 0.f)
                return;   /* adds nothing */
            const int k = r - R0;
            int slot = -1;
            if (!sfall && k >= k0 && k < k1) {
                int nr = 4 * (r < p.hs.nside ? r
                              : (r > 3 * p.hs.nside ? 4 * p.hs.nside - r : p.hs.nside));
                int off = pix - sp - win_lo[k];
                if (off < 0)
                    off += nr;
                if (off < win_hi[k])   /* inside the ring's window */
                    slot = win_base[k] - base0 + off;
            }
            if constexpr (DET) {
                const unsigned long long fI = fix64(vI, fxs);
                const unsigned long long fQ = fix64(vQ, fxs);
                const unsigned long long fU = fix64(vU, fxs);
                if (slot >= 0) {
                    atomicAdd(&acc64[slot], fI);
                    atomicAdd(&acc64[W + slot], fQ);
                    atomicAdd(&acc64[2 * W + slot], fU);
                } else {   /* no window holds the pixel: global int64 atomics */
                    atomicAdd((unsigned long long *)&p.gI[pix], fI);
                    atomicAdd((unsigned long long *)&p.gQ[pix], fQ);
                    atomicAdd((unsigned long long *)&p.gU[pix], fU);
                }
            } else if (slot >= 0) {
                atomicAdd(&acc[slot], vI);
                atomicAdd(&acc[W + slot], vQ);
                atomicAdd(&acc[2 * W + slot], vU);
            } else {   /* no window holds the pixel: global FP64 atomics */
                unsafeAtomicAdd(&p.gI[pix], (double)vI);
                unsafeAtomicAdd(&p.gQ[pix], (double)vQ);
                unsafeAtomicAdd(&p.gU[pix], (double)vU);
            }
        };
        if constexpr (BG >= 2) {
            /* one wave per (group, sample), groups outermost */
            BatchTable<BG> &bt = wtab[wave];
            const int ngr = L.npairs / BG;
            for (int j = wave; j < ns * ngr; j += kAdjWaves) {
                const int jg = j / ns;
                const long i = s0 + (j - jg * ns);
                const int d0 = jg * BG;
                Centre c = centre_make(p.ra[i], p.dec[i], p.pa[i]);
                c.ra = uni(c.ra);
                c.dec = uni(c.dec);
                c.cdc = uni(c.cdc);
                c.sdc = uni(c.sdc);
                SampleF sf;
                sf.cdc = unif((float)c.cdc);
                sf.sdc = unif((float)c.sdc);
                BatchConsts<BG> K;
                batch_consts<BG, true>(L.b, bt, i, d0, lane, K);
                Disc d;
                disc_setup(p.hs, kHalfPi - c.dec, c.ra, L.pairs[d0].rho_max,
                           L.pairs[d0].cos_rmax, d);
                clamp(d);
                auto consume = [&](const BatchPend<BG> &q) {
                    float vI = 0.f, vQ = 0.f, vU = 0.f;
#pragma unroll
                    for (int k = 0; k < BG; k++) {
                        if (q.live & (1u << k)) {
                            float dI, dQ, dU;
                            adj_visit_values(K.GA[k], K.GB[k], q.q0[k], q.q1[k], q.qc[k],
                                             q.qs[k], dI, dQ, dU);
                            vI += dI;
                            vQ += dQ;
                            vU += dU;
                        }
                    }
                    scatter(q.pix, q.r, q.sp, vI, vQ, vU);
                };
                auto exact = [&](int k, int pix, int r, int sp, const ExactVisit &e,
                                 const float4 &v0, const float2 &v1) {
                    float vI, vQ, vU;
                    adj_visit_values(bt.ga[k], bt.gb[k], v0, v1, e.c2, e.s2, vI, vQ, vU);
                    scatter(pix, r, sp, vI, vQ, vU);
                };
                batch_walk<BG, false>(L.b, bt, d, lane, c, sf, K, d0, consume, exact);
            }
        } else if constexpr (SUB < kWaveSize) {
            /* single pair, small discs: 64/SUB samples per wave, SUB lanes
             * each, with the FP64 mapping (as k_fwd_sub) */
            constexpr int G = kWaveSize / SUB;
            const int grp = lane / SUB, sl = lane % SUB;
            SubTable<SUB> &tg = ((SubTable<SUB> *)&wtab[wave])[grp];
            for (int j0 = wave * G; j0 < ns; j0 += kAdjWaves * G) {
                const int j = j0 + grp;
                const bool valid = j < ns;
                const long i = s0 + (valid ? j : j0);
                Centre c;
                double c2psi, s2psi;
                Disc d;
                int nrings = sub_setup(p, i, valid, c, c2psi, s2psi, d);
                clamp(d);
                nrings = nrings > 0 && d.r_last >= d.r_first ? d.r_last - d.r_first + 1 : 0;
                const float ga = (valid && p.g_a) ? p.g_a[i] : 0.f;
                const float gb = (valid && p.g_b) ? p.g_b[i] : 0.f;
                auto use = [&](const Pending &q) {
                    float vI, vQ, vU;
                    adj_visit_values(ga, gb, q.v0, q.v1, q.c2, q.s2, vI, vQ, vU);
                    scatter(q.pix, q.r, q.sp, vI, vQ, vU);
                };
                const int npass = (wave_max_uniform(nrings) + SUB - 1) / SUB;
                for (int ps = 0; ps < npass; ps++) {
                    const int total = build_sub<SUB>(p.hs, d, nrings, ps, sl, tg);
                    walk_sub<SUB, false>(p, tg, total, wave_max_uniform(total), sl, c, c2psi,
                                         s2psi, use);
                    wave_sync();
                }
            }
        } else {
        /* one wave per (pair, sample), pairs outermost */
        const int np = BATCH ? L.npairs : 1;
        for (int j = wave; j < ns * np; j += kAdjWaves) {
            const int jp = j / ns;
            const long i = s0 + (j - jp * ns);
            float pa_i;
            PairParams pp;
            if constexpr (BATCH) {
                const BatchPair &q = L.pairs[jp];
                if (q.out < 0)
                    continue;   /* inert padding (uniform per wave) */
                pp = batch_pair_params(L.b, q);
                pa_i = pp.pa[i] + q.pa_off;
            } else {
                pp = p;
                pa_i = p.pa[i];
            }
            Centre c;
            SampleF sf;
            SampleD sd;
            Disc d;
            sample_setup(pp, i, pa_i, c, sf, sd, d);
            clamp(d);
            const float ga = unif(pp.g_a ? pp.g_a[i] : 0.f);
            const float gb = unif(pp.g_b ? pp.g_b[i] : 0.f);
            auto use = [&](const Pending &q) {
                float vI, vQ, vU;
                adj_visit_values(ga, gb, q.v0, q.v1, q.c2, q.s2, vI, vQ, vU);
                scatter(q.pix, q.r, q.sp, vI, vQ, vU);
            };
            int qn = 0;
            for (int r0 = d.r_first; r0 <= d.r_last; r0 += kWaveSize) {
                int total = build_runs(pp.hs, d, r0, lane, t);
                walk_visits<false, false>(pp, t, total, lane, c, sf, sd, qn, use);
                wave_sync();
            }
            if (qn > 0)
                drain_exact<false>(pp, t, 0, qn, lane, c, sd.c2psi, sd.s2psi, use);
        }
        }
        __syncthreads();

        /* flush the slab's window */
        if (!sfall) {
            for (int slot = tid; slot < W; slot += kAdjBlock) {
                float vI = 0.f, vQ = 0.f, vU = 0.f;
                unsigned long long fI = 0, fQ = 0, fU = 0;
                if constexpr (DET) {
                    fI = acc64[slot];
                    fQ = acc64[W + slot];
                    fU = acc64[2 * W + slot];
                    if ((fI | fQ | fU) == 0ull)
                        continue;
                } else {
                    vI = acc[slot];
                    vQ = acc[W + slot];
                    vU = acc[2 * W + slot];
                    if (vI == 0.f && vQ == 0.f && vU == 0.f)
                        continue;
                }
                int lo = k0, hi = k1;   /* win_base[lo] <= base0 + slot < win_base[hi] */
                while (hi - lo > 1) {
                    int mid = (lo + hi) >> 1;
                    if (win_base[mid] - base0 <= slot)
                        lo = mid;
                    else
                        hi = mid;
                }
                int sp, nr;
                bool shf;
                ring_info(p.hs, R0 + lo, sp, nr, shf);
                int ip = win_lo[lo] + (slot - (win_base[lo] - base0));
                if (ip >= nr)
                    ip -= nr;
                int pix = sp + ip;
                if constexpr (DET) {
                    if (fI)
                        atomicAdd((unsigned long long *)&p.gI[pix], fI);
                    if (fQ)
                        atomicAdd((unsigned long long *)&p.gQ[pix], fQ);
                    if (fU)
                        atomicAdd((unsigned long long *)&p.gU[pix], fU);
                } else {
                    if (vI != 0.f)
                        unsafeAtomicAdd(&p.gI[pix], (double)vI);
                    if (vQ != 0.f)
                        unsafeAtomicAdd(&p.gQ[pix], (double)vQ);
                    if (vU != 0.f)
                        unsafeAtomicAdd(&p.gU[pix], (double)vU);
                }
            }
        }
        __syncthreads();
        }   /* slabs */
    }
}

/* ---- beam adjoint in LDS bands --------------------------------------------
 *
 * The beam adjoint adds six values per visit to the visit's beam row.  With
 * global FP64 atomics that is bound by the atomic rate (every sample's disc
 * touches the whole beam support, so no short span of samples repeats a
 * row).  Here the beam rows are split into bands of `rows` consecutive rows
 * that fit the LDS; a work item is (band, chunk of samples).  A point in a
 * pixel of beam ring R lies between the centres of rings R-1 and R+1, so the
 * visits whose beam pixel falls in rows [lo, hi) of rings Rf..Rl have
 * rho in [theta(Rf-1), theta(Rl+1)]: the item walks only that annulus of each
 * sample's disc (annulus_ring), keeps the visits whose exact beam pixel is in
 * the band, sums their values in float in LDS and adds the band to the FP64
 * gradient once per item. */
constexpr int kBandRuns = 4 * kWaveSize;   /* <= 4 runs per ring, 64 rings per pass */

struct BandRunTable {
    int pref[kBandRuns + 1];
    int pix0[kBandRuns];
    int ring[kBandRuns];
    int spare[7];   /* room for eight 8-lane tables (SubBandTable<8>) in the same slot */
};

struct BeamBands {
    int rows;      /* rows per band */
    int nband;
    int chunk;     /* samples per work item */
    long nchunks;
};

__device__ inline int build_band_runs(const Hpx &hs, const Disc &dout, const Disc &din,
                                      bool inner, int r0, int lane, BandRunTable &t)
{
    const int r = r0 + lane;
    int p0[4] = {0, 0, 0, 0}, ln[4] = {0, 0, 0, 0};
    int ns = 0;
    if (r <= dout.r_last)
        ns = annulus_ring<true>(hs, dout, din, inner, r, p0, ln);
    int tot = 0;
#pragma unroll
    for (int k = 0; k < 4; k++)
        tot += k < ns ? ln[k] : 0;
    const int cnt_in = wave_incl_scan(ns, lane);
    const int tot_in = wave_incl_scan(tot, lane);
    const int cbase = cnt_in - ns;
    int off = tot_in - tot;
#pragma unroll
    for (int k = 0; k < 4; k++) {   /* constant indices keep p0, ln in registers */
        if (k < ns) {
            t.pix0[cbase + k] = p0[k];
            t.ring[cbase + k] = r;
            t.pref[cbase + k] = off;
            off += ln[k];
        }
    }
    const int nrun = __shfl(cnt_in, kWaveSize - 1, kWaveSize);
    const int total = __shfl(tot_in, kWaveSize - 1, kWaveSize);
    if (lane == 0)
        t.pref[nrun] = total;
    wave_sync();
    return total;
}

/* The walk of one annulus pass: FP64 certified mapping, software pipelined
 * as walk_visits; every load of the loop issued unconditionally.  use(q) is
 * called for each visit whose beam pixel lies in rows [lo, lo + nrow). */
template <typename Use>
__device__ inline void walk_band(const PairParams &p, const BandRunTable &t, int total, int lane,
                                 const Centre &c, double c2psi, double s2psi, int lo, int nrow,
                                 Use &use)
{
    if (total <= 0)
        return;
    int s = 0;
    int v = lane;
    if (v < total)
        while (v >= t.pref[s + 1])
            s++;
    RingRec rr = p.rings[t.ring[s]];
    __builtin_amdgcn_s_waitcnt(0);   /* as in walk_visits */
    Pending q;
    q.live = false;
    for (int v0 = 0; v0 < total; v0 += kWaveSize, v += kWaveSize) {
        const bool act = v < total;
        const int pix = t.pix0[s] + (act ? v - t.pref[s] : 0);
        const int r = t.ring[s];
        const RingRec cur = rr;
        const int vn = v + kWaveSize;
        if (vn < total)
            while (vn >= t.pref[s + 1])
                s++;
        rr = p.rings[t.ring[s]];
        float c2, s2;
        const int bp = beam_pixel_f64(p.hb, p.fc64, c, c2psi, s2psi,
                                      ring_phi(p.hs, cur, r, pix), cur.cdp, cur.sdp, c2, s2);
        if (q.live)
            use(q);
        q.live = act && (unsigned)(bp - lo) < (unsigned)nrow;
        q.pix = pix;
        q.bp = bp;
        q.c2 = c2;
        q.s2 = s2;
        const float4 sv = p.sky4[q.live ? pix : 0];
        q.sI = sv.x;
        q.sQ = sv.y;
        q.sU = sv.z;
    }
    if (q.live)
        use(q);
}

/* Several samples per wave (64/SUB, SUB lanes each) for the band walks:
 * per (band, sample) the annulus holds few visits and its set-up is a long
 * dependent chain, so lane groups keep several chains in flight. */
template <int SUB>
struct SubBandTable {
    int pref[4 * SUB + 1];
    int pix0[4 * SUB];
    int ring[4 * SUB];
};

template <int SUB>
__device__ inline int build_band_sub(const Hpx &hs, const Disc &dout, const Disc &din, bool inner,
                                     int nrings, int ps, int sl, SubBandTable<SUB> &t)
{
    const int r = dout.r_first + ps * SUB + sl;
    int p0[4] = {0, 0, 0, 0}, ln[4] = {0, 0, 0, 0};
    int ns = 0;
    if (ps * SUB + sl < nrings)
        ns = annulus_ring<true>(hs, dout, din, inner, r, p0, ln);
    int tot = 0;
#pragma unroll
    for (int k = 0; k < 4; k++)
        tot += k < ns ? ln[k] : 0;
    const int cnt_in = seg_incl_scan<SUB>(ns, sl);
    const int tot_in = seg_incl_scan<SUB>(tot, sl);
    const int cbase = cnt_in - ns;
    int off = tot_in - tot;
#pragma unroll
    for (int k = 0; k < 4; k++) {   /* constant indices keep p0, ln in registers */
        if (k < ns) {
            t.pix0[cbase + k] = p0[k];
            t.ring[cbase + k] = r;
            t.pref[cbase + k] = off;
            off += ln[k];
        }
    }
    const int nrun = __shfl(cnt_in, SUB - 1, SUB);
    const int total = __shfl(tot_in, SUB - 1, SUB);
    if (sl == 0)
        t.pref[nrun] = total;
    wave_sync();
    return total;
}

/* walk_band for a lane group: its `total` visits SUB at a time, for wtot
 * (the wave's largest total) iterations' worth; a lane without a visit
 * computes the first visit of its run (or pixel 0 of ring 1) and discards
 * it. */
template <int SUB, typename Use>
__device__ inline void walk_band_sub(const PairParams &p, const SubBandTable<SUB> &t, int total,
                                     int wtot, int sl, const Centre &c, double c2psi,
                                     double s2psi, int lo, int nrow, Use &use)
{
    if (wtot <= 0)
        return;
    const bool any = total > 0;
    int s = 0;
    int v = sl;
    if (v < total)
        while (v >= t.pref[s + 1])
            s++;
    RingRec rr = p.rings[any ? t.ring[s] : 1];
    __builtin_amdgcn_s_waitcnt(0);   /* as in walk_visits */
    Pending q;
    q.live = false;
    for (int v0 = 0; v0 < wtot; v0 += SUB, v += SUB) {
        const bool act = v < total;
        const int r = any ? t.ring[s] : 1;
        const int pix = any ? t.pix0[s] + (act ? v - t.pref[s] : 0) : 0;
        const RingRec cur = rr;
        const int vn = v + SUB;
        if (vn < total)
            while (vn >= t.pref[s + 1])
                s++;
        rr = p.rings[any ? t.ring[s] : 1];
        float c2, s2;
        const int bp = beam_pixel_f64(p.hb, p.fc64, c, c2psi, s2psi,
                                      ring_phi(p.hs, cur, r, pix), cur.cdp, cur.sdp, c2, s2);
        if (q.live)
            use(q);
        q.live = act && (unsigned)(bp - lo) < (unsigned)nrow;
        q.pix = pix;
        q.bp = bp;
        q.c2 = c2;
        q.s2 = s2;
        const float4 sv = p.sky4[q.live ? pix : 0];
        q.sI = sv.x;
        q.sQ = sv.y;
        q.sU = sv.z;
    }
    if (q.live)
        use(q);
}

template <int SUB>
__global__ __launch_bounds__(kAdjBlock) void k_adjbeam_band_sub(PairParams p, BeamBands B)
{
    constexpr int G = kWaveSize / SUB;
    static_assert(sizeof(SubBandTable<SUB>) * G <= sizeof(BandRunTable),
                  "lane-group tables within the wave's run table");
    extern __shared__ __align__(16) unsigned char smem[];
    BandRunTable *tabs = (BandRunTable *)smem;
    float *acc = (float *)(tabs + kAdjWaves);
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWaveSize);
    const int lane = tid % kWaveSize;
    const int grp = lane / SUB, sl = lane % SUB;
    SubBandTable<SUB> &t = ((SubBandTable<SUB> *)&tabs[wave])[grp];
    const long nitems = B.nchunks * B.nband;
    for (long it = xcd_remap(blockIdx.x, gridDim.x); it < nitems; it += gridDim.x) {
        const long ch = it / B.nband;
        const int k = (int)(it - ch * B.nband);
        const int lo = k * B.rows;
        const int nrow = min(p.npix_beam - lo, B.rows);
        bool inner;
        double r_in, c_in, r_out, c_out;
        band_annulus_radii(p.hb, p.rho_max, p.cos_rmax, lo, nrow, inner, r_in, c_in, r_out, c_out);
        for (int i = tid; i < 6 * nrow; i += kAdjBlock)
            acc[i] = 0.f;
        __syncthreads();
        const long s0 = p.s_begin + ch * B.chunk;
        const long s1 = s0 + B.chunk < p.s_end ? s0 + B.chunk : p.s_end;
        for (long j0 = s0 + (long)wave * G; j0 < s1; j0 += (long)kAdjWaves * G) {
            const long i = j0 + grp;
            const bool valid = i < s1;
            const long iv = valid ? i : j0;
            const float ga = valid && p.g_a ? p.g_a[iv] : 0.f;
            const float gb = valid && p.g_b ? p.g_b[iv] : 0.f;
            const Centre c = centre_make(p.ra[iv], p.dec[iv], p.pa[iv]);
            double s2psi, c2psi;
            pl_sincos(2.0 * c.psi, &s2psi, &c2psi);
            Disc dout, din;
            disc_setup(p.hs, kHalfPi - c.dec, c.ra, r_out, c_out, dout);
            if (inner)
                disc_setup(p.hs, kHalfPi - c.dec, c.ra, r_in, c_in, din);
            else
                din = dout;   /* not read */
            const bool work = valid && (ga != 0.f || gb != 0.f);
            const int nrings = work && dout.r_last >= dout.r_first ? dout.r_last - dout.r_first + 1
                                                                    : 0;
            auto use = [&](const Pending &q) {
                const float X = q.sQ * q.c2 + q.sU * q.s2;
                const float Y = -q.sQ * q.s2 + q.sU * q.c2;
                float *a = acc + 6 * (q.bp - lo);
                atomicAdd(&a[0], ga * q.sI);
                atomicAdd(&a[1], ga * X);
                atomicAdd(&a[2], ga * Y);
                atomicAdd(&a[3], gb * q.sI);
                atomicAdd(&a[4], -(gb * X));
                atomicAdd(&a[5], -(gb * Y));
            };
            const int npass = (wave_max_uniform(nrings) + SUB - 1) / SUB;
            for (int ps = 0; ps < npass; ps++) {
                const int total = build_band_sub<SUB>(p.hs, dout, din, inner, nrings, ps, sl, t);
                walk_band_sub<SUB>(p, t, total, wave_max_uniform(total), sl, c, c2psi, s2psi, lo,
                                   nrow, use);
                wave_sync();
            }
        }
        __syncthreads();
        for (int rr = tid; rr < nrow; rr += kAdjBlock) {
            const float *a = acc + 6 * rr;
            double *g = p.gbeam + 8 * (long)(lo + rr);
#pragma unroll
            for (int m = 0; m < 6; m++)
                if (a[m] != 0.f)
                    unsafeAtomicAdd(&g[m], (double)a[m]);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kAdjBlock) void k_adjbeam_band(PairParams p, BeamBands B)
{
    extern __shared__ __align__(16) unsigned char smem[];
    BandRunTable *tabs = (BandRunTable *)smem;
    float *acc = (float *)(tabs + kAdjWaves);
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWaveSize);
    const int lane = tid % kWaveSize;
    BandRunTable &t = tabs[wave];
    const long nitems = B.nchunks * B.nband;
    for (long it = xcd_remap(blockIdx.x, gridDim.x); it < nitems; it += gridDim.x) {
        /* consecutive items: the bands of one chunk (its samples stay in L2) */
        const long ch = it / B.nband;
        const int k = (int)(it - ch * B.nband);
        const int lo = k * B.rows;
        const int nrow = min(p.npix_beam - lo, B.rows);
        /* the annulus of rows [lo, lo + nrow) */
        bool inner;
        double r_in, c_in, r_out, c_out;
        band_annulus_radii(p.hb, p.rho_max, p.cos_rmax, lo, nrow, inner, r_in, c_in, r_out,
                           c_out);
        for (int i = tid; i < 6 * nrow; i += kAdjBlock)
            acc[i] = 0.f;
        __syncthreads();
        const long s0 = p.s_begin + ch * B.chunk;
        const long s1 = s0 + B.chunk < p.s_end ? s0 + B.chunk : p.s_end;
        for (long i = s0 + wave; i < s1; i += kAdjWaves) {
            const float ga = unif(p.g_a ? p.g_a[i] : 0.f);
            const float gb = unif(p.g_b ? p.g_b[i] : 0.f);
            if (ga == 0.f && gb == 0.f)
                continue;   /* adds nothing (wave-uniform) */
            Centre c = centre_make(p.ra[i], p.dec[i], p.pa[i]);
            c.ra = uni(c.ra);
            c.dec = uni(c.dec);
            c.psi = uni(c.psi);
            c.cdc = uni(c.cdc);
            c.sdc = uni(c.sdc);
            double s2psi, c2psi;
            pl_sincos(2.0 * c.psi, &s2psi, &c2psi);
            c2psi = uni(c2psi);
            s2psi = uni(s2psi);
            Disc dout, din;
            disc_setup(p.hs, kHalfPi - c.dec, c.ra, r_out, c_out, dout);
            if (inner)
                disc_setup(p.hs, kHalfPi - c.dec, c.ra, r_in, c_in, din);
            else
                din = dout;   /* not read */
            auto use = [&](const Pending &q) {
                const float X = q.sQ * q.c2 + q.sU * q.s2;
                const float Y = -q.sQ * q.s2 + q.sU * q.c2;
                float *a = acc + 6 * (q.bp - lo);
                atomicAdd(&a[0], ga * q.sI);
                atomicAdd(&a[1], ga * X);
                atomicAdd(&a[2], ga * Y);
                atomicAdd(&a[3], gb * q.sI);
                atomicAdd(&a[4], -(gb * X));
                atomicAdd(&a[5], -(gb * Y));
            };
            for (int r0 = dout.r_first; r0 <= dout.r_last; r0 += kWaveSize) {
                const int total = build_band_runs(p.hs, dout, din, inner, r0, lane, t);
                walk_band(p, t, total, lane, c, c2psi, s2psi, lo, nrow, use);
                wave_sync();
            }
        }
        __syncthreads();
        /* add the band to the FP64 gradient */
        for (int rr = tid; rr < nrow; rr += kAdjBlock) {
            const float *a = acc + 6 * rr;
            double *g = p.gbeam + 8 * (long)(lo + rr);
#pragma unroll
            for (int m = 0; m < 6; m++)
                if (a[m] != 0.f)
                    unsafeAtomicAdd(&g[m], (double)a[m]);
        }
        __syncthreads();
    }
}

/* ---- beam adjoint through visit records -----------------------------------
 *
 * The banded kernels above map each sample's annulus once per band of beam
 * rows and set up every (band, sample) pair.  Here each visit is mapped
 * twice, once per pass, and the gradient is formed from 16-byte visit
 * records, per chunk of samples:
 *  1. k_beamrec<false> (one wave per sample, the sky adjoint's walk and
 *     certified mapping) counts the sample's visits inside the beam per band
 *     of `rows` consecutive beam rows;
 *  2. k_band_scan + k_scan_excl: per band, the offsets of the samples' records
 *     within the band; the bands' starts;
 *  3. k_beamrec<true> maps the visits again and writes each one's record
 *     (row within the band | sample << kRecRowBits, I, X = Q c2 + U s2,
 *     Y = -Q s2 + U c2) into its band, within the range of the samples of
 *     its block;
 *  4. k_beamrec_sum: a block per (band, part: the records of a share of the
 *     chunk's samples) sums its records, times the cotangents of each
 *     record's sample, into the band's
 *     rows in LDS as 64-bit fixed-point integers, then adds them to its own
 *     FP64 partial rows (a (band, part) slice belongs to one block in every
 *     chunk: no global atomics).
 * k_beamrec_reduce adds the parts into the FP64 gradient at the end.
 * Integers because float atomic adds to random LDS addresses ran about ten
 * times slower than 64-bit integer ones here (6.8 against 1.0 ms per chunk
 * of 41,000 samples at the 5-degree, beam NSIDE 4096 shape); integer sums
 * also make the result independent of the order of the additions. */
struct BeamRec {
    unsigned key;   /* row within the band | sample within the chunk << kRecRowBits */
    float sI, X, Y;
};
constexpr int kRecRowBits = 13;
constexpr int kRecUnroll = 4;
constexpr int kRecWriteWaves = 8;   /* samples per block of the write pass (shared cursors) */
constexpr int kScanBlock = 1024;

/* In-place exclusive scan of x[0 .. n) (one block); x[n] receives the total. */
__global__ __launch_bounds__(kScanBlock) void k_scan_excl(long *x, long n)
{
    __shared__ long part[kScanBlock];
    const int t = threadIdx.x;
    const long seg = (n + kScanBlock - 1) / kScanBlock;
    const long a = std::min<long>(n, t * seg), b = std::min<long>(n, a + seg);
    long s = 0;
    for (long k = a; k < b; k++)
        s += x[k];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {
        const long v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long run = part[t] - s;
    for (long k = a; k < b; k++) {
        const long v = x[k];
        x[k] = run;
        run += v;
    }
    if (t == kScanBlock - 1)
        x[n] = part[t];
}

__device__ inline int rec_band(const RecParams &R, int bp)
{
    int b = (int)((float)bp * R.inv_rows);
    if (b * R.rows > bp)
        b--;
    else if ((b + 1) * R.rows <= bp)
        b++;
    return b;
}

/* Steps 1 and 3: one wave per sample of the chunk [p.s_begin, p.s_end),
 * the sky adjoint's walk and certified single-precision mapping (every
 * visit's beam pixel is the exact one; cos/sin 2chi within 9.2e-7), with the
 * deferred visits mapped by the FP64-alpha form.  The count pass loads no
 * beam rows; the write pass loads each visit's sky values, pipelined. */
template <bool WRITE>
__global__ __launch_bounds__(WRITE ? kRecWriteWaves * kWaveSize : kBlock)
__attribute__((amdgpu_waves_per_eu(4, 8)))
void k_beamrec(PairParams p, RecParams R)
{
    constexpr int NW = WRITE ? kRecWriteWaves : kWavesPerBlock;   /* waves (samples) per block */
    __shared__ RunTable tables[NW];
    /* counts: one row per wave (sample); write cursors: one row per block */
    __shared__ unsigned curs[WRITE ? 1 : NW][kRecMaxBands];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWaveSize);
    const int lane = threadIdx.x % kWaveSize;
    RunTable &t = tables[wave];
    const long j0 = xcd_remap(blockIdx.x, gridDim.x) * NW;
    const long j = j0 + wave;
    const long i = p.s_begin + j;
    unsigned *cur;
    if constexpr (WRITE) {
        /* The block's samples j0 .. j0+NW-1 own adjacent ranges in every
         * band (the offsets are an exclusive scan over samples), so the
         * block fills each band's range through one shared cursor: a
         * 128-byte line of records is completed by the block's waves
         * together, while it is still in L2, instead of by one wave over
         * many iterations (partly written lines left L2 and doubled the
         * pass's HBM writes). */
        cur = curs[0];
        for (int b = threadIdx.x; b < R.nband; b += NW * kWaveSize)
            cur[b] = (unsigned)(R.bbase[b] + R.bcnt[b * R.chunk + j0]);
        __syncthreads();
        if (i >= p.s_end)
            return;
    } else {
        if (i >= p.s_end)
            return;
        cur = curs[wave];
        for (int b = lane; b < R.nband; b += kWaveSize)
            cur[b] = 0u;
        wave_sync();
    }
    const float ga = unif(p.g_a ? p.g_a[i] : 0.f);
    const float gb = unif(p.g_b ? p.g_b[i] : 0.f);
    if (ga != 0.f || gb != 0.f) {   /* else it adds nothing (wave-uniform) */
        Centre c;
        SampleF sf;
        SampleD sd;
        Disc d;
        sample_setup(p, i, p.pa[i], c, sf, sd, d);
        int qn = 0;
        if constexpr (WRITE) {
            const unsigned smp_key = (unsigned)j << kRecRowBits;
            auto write = [&](const Pending &q) {
                const int b = rec_band(R, q.bp);
                const unsigned slot = atomicAdd(&cur[b], 1u);
                if (slot >= (unsigned)R.nrec_cap)
                    return;   /* never: both passes count the same visits */
                BeamRec r;
                r.key = (unsigned)(q.bp - b * R.rows) | smp_key;
                r.sI = q.sI;
                r.X = q.sQ * q.c2 + q.sU * q.s2;
                r.Y = -q.sQ * q.s2 + q.sU * q.c2;
                R.brec[slot] = r;
            };
            for (int r0 = d.r_first; r0 <= d.r_last; r0 += kWaveSize) {
                const int total = build_runs(p.hs, d, r0, lane, t);
                walk_visits<true, false, decltype(write), false, false>(p, t, total, lane, c, sf,
                                                                        sd, qn, write);
                wave_sync();
            }
            if (qn > 0)
                drain_exact<true>(p, t, 0, qn, lane, c, sd.c2psi, sd.s2psi, write);
        } else {
            auto count = [&](const Pending &q) { atomicAdd(&cur[rec_band(R, q.bp)], 1u); };
            for (int r0 = d.r_first; r0 <= d.r_last; r0 += kWaveSize) {
                const int total = build_runs(p.hs, d, r0, lane, t);
                walk_visits<false, false, decltype(count), false, false>(p, t, total, lane, c, sf,
                                                                         sd, qn, count);
                wave_sync();
            }
            if (qn > 0)
                drain_exact<false>(p, t, 0, qn, lane, c, sd.c2psi, sd.s2psi, count);
        }
    }
    if constexpr (!WRITE) {
        wave_sync();
        for (int b = lane; b < R.nband; b += kWaveSize)
            R.bcnt[b * R.chunk + j] = cur[b];
    }
}

/* Step 2: per band (one block each), the exclusive offsets of the samples'
 * records within the band (in place) and the band's total in bbase[b].
 * Each thread scans kScanItems consecutive samples per round. */
constexpr int kScanItems = 16;
__global__ __launch_bounds__(kScanBlock) void k_band_scan(RecParams R, long nsamp)
{
    __shared__ long part[kScanBlock];
    const int b = blockIdx.x, t = threadIdx.x;
    long *x = R.bcnt + (long)b * R.chunk;
    long carry = 0;
    for (long r0 = 0; r0 < nsamp; r0 += (long)kScanBlock * kScanItems) {
        long v[kScanItems];
        long s = 0;
#pragma unroll
        for (int q = 0; q < kScanItems; q++) {
            const long k = r0 + (long)t * kScanItems + q;
            v[q] = k < nsamp ? x[k] : 0;
            s += v[q];
        }
        part[t] = s;
        __syncthreads();
        for (int o = 1; o < kScanBlock; o <<= 1) {
            const long w = t >= o ? part[t - o] : 0;
            __syncthreads();
            part[t] += w;
            __syncthreads();
        }
        long run = carry + part[t] - s;
#pragma unroll
        for (int q = 0; q < kScanItems; q++) {
            const long k = r0 + (long)t * kScanItems + q;
            if (k < nsamp)
                x[k] = run;
            run += v[q];
        }
        carry += part[kScanBlock - 1];
        __syncthreads();
    }
    if (t == 0)
        R.bbase[b] = carry;
}

/* Step 4: block (band, part) of nsplit.  Each value v = g * (I, X or Y) is
 * rounded once to a multiple of 2^-k and summed as a 64-bit integer, with k
 * the largest integer such that the block's records, each |v| <= gmax *
 * smax * (1 + 1e-3), cannot sum past 2^62 in magnitude. */
__global__ __launch_bounds__(kAdjBlock) void k_beamrec_sum(PairParams p, RecParams R)
{
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned long long *acc = (unsigned long long *)smem;
    const int tid = threadIdx.x;
    const int b = blockIdx.x / R.nsplit, part = blockIdx.x % R.nsplit;
    const int lo = b * R.rows;
    const int nrow = min(p.npix_beam - lo, R.rows);
    const long r0 = R.bbase[b], cnt = R.bbase[b + 1] - r0;
    /* the part's samples: a share of the chunk's, cut at multiples of the
     * write pass's samples per block, so that the records of a part are the
     * same set however the block's waves interleaved them */
    const long ns = p.s_end - p.s_begin;
    const long sa = (ns * part / R.nsplit) / kRecWriteWaves * kRecWriteWaves;
    const long sb = part + 1 == R.nsplit
                        ? ns
                        : (ns * (part + 1) / R.nsplit) / kRecWriteWaves * kRecWriteWaves;
    const long k0 = sa < ns ? R.bcnt[b * R.chunk + sa] : cnt;
    const long j1 = sb < ns ? R.bcnt[b * R.chunk + sb] : cnt;
    if (k0 >= j1)
        return;   /* no records (uniform per block) */
    const double vmax = (double)R.vmax[0] * (double)R.vmax[1] * (1.0 + 1e-3);
    if (!(vmax > 0.0))
        return;   /* every value is zero */
    /* every sum below 2^62 and (for fix64) every single value below 2^50 */
    const int e = (int)std::ceil(std::log2((double)(j1 - k0) * vmax));
    const double scale = std::ldexp(1.0, std::min(62 - e, 49 - ilogb(vmax)));
    for (int k = tid; k < 6 * nrow; k += kAdjBlock)
        acc[k] = 0ull;
    __syncthreads();
    /* an absent cotangent reads the other one and adds zeros (with neither
     * there are no records) */
    const bool has_a = p.g_a != nullptr, has_b = p.g_b != nullptr;
    const float *gA = has_a ? p.g_a : p.g_b, *gB = has_b ? p.g_b : p.g_a;
    if (gA == nullptr)
        return;
    /* records of the next batch are loaded while the current one is summed */
    BeamRec nx[kRecUnroll];
#pragma unroll
    for (int u = 0; u < kRecUnroll; u++) {
        const long k = k0 + u * kAdjBlock + tid;
        nx[u] = R.brec[r0 + (k < j1 ? k : k0)];
    }
    for (long kk = k0; kk < j1; kk += (long)kRecUnroll * kAdjBlock) {
        BeamRec r[kRecUnroll];
        float ga[kRecUnroll], gb[kRecUnroll];
#pragma unroll
        for (int u = 0; u < kRecUnroll; u++) {
            const long k = kk + u * kAdjBlock + tid;
            r[u] = nx[u];
            if (k >= j1)
                r[u].sI = r[u].X = r[u].Y = 0.f;   /* adds zeros */
        }
        /* the cotangents first, then the prefetch: loads complete in order,
         * so waiting for the cotangents does not wait for the prefetch */
#pragma unroll
        for (int u = 0; u < kRecUnroll; u++) {
            const long i = p.s_begin + (long)(r[u].key >> kRecRowBits);
            const float a = gA[i], bb = gB[i];   /* unconditional: no branch between loads */
            ga[u] = has_a ? a : 0.f;
            gb[u] = has_b ? bb : 0.f;
        }
#pragma unroll
        for (int u = 0; u < kRecUnroll; u++) {
            const long kn = kk + u * kAdjBlock + tid + (long)kRecUnroll * kAdjBlock;
            nx[u] = R.brec[r0 + (kn < j1 ? kn : k0)];
        }
#pragma unroll
        for (int u = 0; u < kRecUnroll; u++) {
            unsigned long long *a = acc + 6 * (int)(r[u].key & ((1u << kRecRowBits) - 1));
            const float v[6] = {ga[u] * r[u].sI, ga[u] * r[u].X,    ga[u] * r[u].Y,
                                gb[u] * r[u].sI, -(gb[u] * r[u].X), -(gb[u] * r[u].Y)};
#pragma unroll
            for (int m = 0; m < 6; m++)
                if (v[m] != 0.f)
                    atomicAdd(&a[m], fix64(v[m], scale));
        }
    }
    __syncthreads();
    const double inv = 1.0 / scale;
    double *dst = R.part + ((long)part * p.npix_beam + lo) * 6;
    for (int k = tid; k < 6 * nrow; k += kAdjBlock)
        dst[k] += (double)(long long)acc[k] * inv;
}

/* The parts' sums into the FP64 gradient (rows of eight, columns 0..5), in
 * part order. */
__global__ void k_beamrec_reduce(const double *part, int nsplit, long npix, double *gbeam)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; k < 6 * npix; k += (long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < nsplit; q++)
            s += part[(long)q * 6 * npix + k];
        const long row = k / 6;
        gbeam[8 * row + (k - 6 * row)] += s;
    }
}

/* vmax[0] = max |g_a|, |g_b| over the samples; vmax[1] = max over sky
 * pixels of max(|I|, |Q| + |U|) (bounds |I|, |X| and |Y| of every record).
 * vmax must hold zeros on entry; non-negative floats order as their bits. */
__global__ void k_rec_vmax(PairParams p, float *vmax)
{
    float g = 0.f, sk = 0.f;
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = p.s_end - p.s_begin, stride = (long)gridDim.x * blockDim.x;
    for (long j = k; j < n; j += stride) {
        if (p.g_a)
            g = fmaxf(g, fabsf(p.g_a[p.s_begin + j]));
        if (p.g_b)
            g = fmaxf(g, fabsf(p.g_b[p.s_begin + j]));
    }
    for (long j = k; j < p.hs.npix; j += stride) {
        const float4 v = p.sky4[j];
        sk = fmaxf(sk, fmaxf(fabsf(v.x), fabsf(v.y) + fabsf(v.z)));
    }
#pragma unroll
    for (int m = kWaveSize / 2; m > 0; m >>= 1) {
        g = fmaxf(g, __shfl_xor(g, m, kWaveSize));
        sk = fmaxf(sk, __shfl_xor(sk, m, kWaveSize));
    }
    if (threadIdx.x % kWaveSize == 0) {
        atomicMax((unsigned *)&vmax[0], __float_as_uint(g));
        atomicMax((unsigned *)&vmax[1], __float_as_uint(sk));
    }
}

__global__ void k_fill_f64_n(double *x, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = 0.0;
}

__global__ void k_fill_i32(int *x, int v) { *x = v; }

/* Last row with a non-zero value in columns 0..5 (wave max, one atomic per
 * wave). */
__global__ void k_beam_support(const float4 *beam8, long npix, int *last)
{
    int best = -1;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < npix; i += (long)gridDim.x * blockDim.x) {
        float4 a = beam8[2 * i];
        float4 b = beam8[2 * i + 1];
        if (a.x != 0.f || a.y != 0.f || a.z != 0.f || a.w != 0.f || b.x != 0.f || b.y != 0.f)
            best = (int)i;
    }
#pragma unroll
    for (int m = kWaveSize / 2; m > 0; m >>= 1)
        best = max(best, __shfl_xor(best, m, kWaveSize));
    if (threadIdx.x % kWaveSize == 0 && best >= 0)
        atomicMax(last, best);
}

/* Ring records of every ring, indexed by ring number 1 .. 4N-1 (entry 0
 * unused). */
__global__ void k_ring_table(Hpx h, RingRec *out)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (r >= 4 * h.nside)
        return;
    out[r] = ring_record(h, r);
}

__global__ void k_fill_f64(double *x, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = 0.0;
}

__global__ void k_add_offset(const float *x, float off, float *y, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = x[i] + off;
}

__global__ void k_pack_sky(const float *I, const float *Q, const float *U, const float *V, long n,
                           float4 *out)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = make_float4(I[i], Q[i], U[i], V ? V[i] : 0.f);
}

/* Deterministic sky adjoint: out[0] = max over beam rows of the sum of |A0
 * A1 A2 B0 B1 B2|, out[1] = max |g| (as float bits; atomicMax on the bits of
 * non-negative floats orders them as the values). */
__global__ void k_absmax_rows(const float *beam8, long npix, unsigned *out)
{
    float best = 0.f;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < npix; i += (long)gridDim.x * blockDim.x) {
        const float *r = beam8 + 8 * i;
        float v = fabsf(r[0]) + fabsf(r[1]) + fabsf(r[2]) + fabsf(r[3]) + fabsf(r[4]) +
                  fabsf(r[5]);
        best = fmaxf(best, v);
    }
    for (int m = kWaveSize / 2; m > 0; m >>= 1)
        best = fmaxf(best, __shfl_xor(best, m, kWaveSize));
    if (threadIdx.x % kWaveSize == 0)
        atomicMax(out, __float_as_uint(best));
}

__global__ void k_absmax(const float *x, long n, unsigned *out)
{
    float best = 0.f;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x)
        best = fmaxf(best, fabsf(x[i]));
    for (int m = kWaveSize / 2; m > 0; m >>= 1)
        best = fmaxf(best, __shfl_xor(best, m, kWaveSize));
    if (threadIdx.x % kWaveSize == 0)
        atomicMax(out, __float_as_uint(best));
}

/* scale = 2^k, the largest power of two with count * bound * scale <= 2^62,
 * bound = max row sum * max |g| (the largest |contribution| of a visit) and
 * count the most contributions one pixel can receive (one per sample and
 * pair); every pixel's sum then fits a signed 64-bit integer. */
__global__ void k_fix_scale(const unsigned *m, long count, double *scale)
{
    const double one = (double)__uint_as_float(m[0]) * (double)__uint_as_float(m[1]);
    const double bound = one * (double)(count > 0 ? count : 1);
    double s = 1.0;
    if (bound > 0.0 && bound < INFINITY) {
        /* every sum below 2^62, and (for fix64) every single value below
         * 2^50: one value is at most m[0] m[1] */
        int k = min(62 - ilogb(bound) - 1, 49 - ilogb(one));
        k = k > 1000 ? 1000 : (k < -1000 ? -1000 : k);
        s = ldexp(1.0, k);
    }
    *scale = s;
}

__global__ void k_fix_to_f32(const unsigned long long *x, const double *scale, float *y, long n)
{
    const double inv = 1.0 / *scale;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = (float)((double)(long long)x[i] * inv);
}

__global__ void k_fill_u32(unsigned *x, int n, unsigned v)
{
    if ((int)threadIdx.x < n)
        x[threadIdx.x] = v;
}

__global__ void k_f64_to_f32(const double *x, float *y, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = (float)x[i];
}

/* eight float maps -> rows A0 A1 A2 B0 B1 B2 A3 B3 */
__global__ void k_pack_beam(long npix, const float *a0, const float *a1, const float *a2,
                            const float *a3, const float *b0, const float *b1,
                            const float *b2, const float *b3, float4 *out)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < npix; i += (long)gridDim.x * blockDim.x) {
        out[2 * i] = make_float4(a0[i], a1[i], a2[i], b0[i]);
        out[2 * i + 1] = make_float4(b1[i], b2[i], a3 ? a3[i] : 0.f, b3 ? b3[i] : 0.f);
    }
}

/* ---- test kernels: the HEALPix and matching primitives on their own ---- */

__global__ void k_pix2ang(Hpx h, long n, const int *pix, double *th, double *ph)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x)
        pix2ang(h, pix[i], th[i], ph[i]);
}

__global__ void k_ang2pix(Hpx h, long n, const double *th, const double *ph, int *pix)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x)
        pix[i] = ang2pix(h, th[i], ph[i]);
}

__global__ void k_interpol(Hpx h, long n, const double *th, const double *ph, int *pix,
                           double *wgt)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x)
        get_interpol(h, th[i], ph[i], pix + 4 * i, wgt + 4 * i);
}

/* Disc of each pointing, one thread per pointing: count (pixels == NULL) or
 * fill pixels at offsets[i]. */
__global__ void k_disc(Hpx h, long n, const float *ra, const float *dec, double radius,
                       double cosr, const long *offsets, int *counts, int *pixels)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    Centre c = centre_make(ra[i], dec[i], 0.f);
    Disc d;
    disc_setup(h, kHalfPi - c.dec, c.ra, radius, cosr, d);
    long k = pixels ? offsets[i] : 0;
    int cnt = 0;
    for (int r = d.r_first; r <= d.r_last; r++) {
        int p0[2] = {0, 0}, ln[2] = {0, 0};
        int ns = disc_ring(h, d, r, p0, ln);
#pragma unroll
        for (int j = 0; j < 2; j++) {
            if (j < ns) {
                if (pixels)
                    for (int q = 0; q < ln[j]; q++)
                        pixels[k + cnt + q] = p0[j] + q;
                cnt += ln[j];
            }
        }
    }
    if (counts)
        counts[i] = cnt;
}

/* Sky pixel -> (beam pixel, chi) for pixel lists laid out as offsets/pixels
 * (the quantity kernel 1 of the reference computes, cudaConv2.cu:59-128).
 * mode 0: the exact FP64 mapping, chi[k] = chi.
 * mode 1: what the adjoint kernels use: the float form, and the exact
 *         mapping where the float form cannot certify the pixel; chi[k]
 *         receives float cos(2chi), sin(2chi) packed as two floats.
 * mode 2: the float form alone; bpix[k] = kNeedExact (-2) where it declines.
 * mode 3: the certified double-precision form of the small-disc kernels and
 *         of the deferred-visit drain (output as mode 1).
 * mode 4: what the one-wave-per-sample forward uses: the double-precision
 *         form with the float alpha, and mode 3 where it declines.
 * mode 5: the double-precision form with the float alpha alone; bpix[k] =
 *         kNeedExact where it declines. */
__global__ void k_match(Hpx hs, const RingRec *rings, Hpx hb, FastConsts fc, FastConsts64 fc64,
                        int mode, long n,
                        const float *ra, const float *dec, const float *pa,
                        const long *offsets, const int *pixels, int *bpix, double *chi)
{
    long i = (long)blockIdx.x;
    if (i >= n)
        return;
    Centre c = centre_make(ra[i], dec[i], pa[i]);
    double s2psi, c2psi;
    pl_sincos(2.0 * c.psi, &s2psi, &c2psi);
    const SampleF sf = sample_f32(c, c2psi, s2psi);
    for (long k = offsets[i] + threadIdx.x; k < offsets[i + 1]; k += blockDim.x) {
        int pix = pixels[k];
        int r = pix2ring(hs, pix);
        RingRec rr = rings[r];
        if (mode > 0) {
            float c2 = 0.f, s2 = 0.f;
            double ra_pix = ring_phi(hs, rr, r, pix);
            int bp;
            if (mode >= 4) {
                bp = beam_pixel_f64<true>(hb, fc64, c, c2psi, s2psi, ra_pix, rr.cdp, rr.sdp, c2,
                                          s2);
                if (bp == kNeedExact && mode == 4)
                    bp = beam_pixel_f64(hb, fc64, c, c2psi, s2psi, ra_pix, rr.cdp, rr.sdp, c2, s2);
            } else if (mode == 3) {
                bp = beam_pixel_f64(hb, fc64, c, c2psi, s2psi, ra_pix, rr.cdp, rr.sdp, c2, s2);
            } else {
                bp = beam_pixel_f32(fc, c, sf, ra_pix, rr.cdp, rr.sdp, c2, s2);
            }
            if (bp == kNeedExact && mode == 1) {
                ExactVisit e =
                    beam_visit_exact(hb.nside, c.ra, c.psi, c.cdc, c.sdc, ra_pix, rr.cdp, rr.sdp);
                bp = e.bp;
                c2 = e.c2;
                s2 = e.s2;
            }
            bpix[k] = bp;
            float2 *o = (float2 *)&chi[k];
            *o = make_float2(c2, s2);
        } else {
            int sp, nr;
            bool sh;
            ring_info(hs, r, sp, nr, sh);
            double ra_pix = pixel_phi(hs, r, pix, sp);
            double x;
            bpix[k] = beam_pixel(hb, c, ra_pix, rr.cdp, rr.sdp, x);
            chi[k] = x;
        }
    }
}

/* Element-wise FP64 primitives, for checking host/device bit equality. */
__global__ void k_libm(int op, long n, const double *x, const double *y, double *o)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x) {
        double a = x[i], b = y ? y[i] : 0.0, r = 0.0;
        switch (op) {
        case 0: r = pl_sin(a); break;
        case 1: r = pl_cos(a); break;
        case 2: r = pl_acos(a); break;
        case 3: r = pl_atan2(a, b); break;
        case 4: r = sqrt(a); break;
        case 5: r = a / b; break;
        case 6: r = fmodulo(a, b); break;
        case 7: { double s, c; pl_sincos(a, &s, &c); r = s + 2.0 * c; } break;
        case 8: r = __longlong_as_double((long long)fix64((float)a, b)); break;
        default: r = 0.0;
        }
        o[i] = r;
    }
}

/* ---- launchers ---------------------------------------------------------- */

static inline unsigned nblocks(long n, int per)
{
    long b = (n + per - 1) / per;
    if (b < 1)
        b = 1;
    return (unsigned)b;
}

static inline unsigned gs_blocks(long n)
{
    long b = (n + 255) / 256;
    if (b > 8192)
        b = 8192;
    if (b < 1)
        b = 1;
    return (unsigned)b;
}

hipError_t launch_pair(int op, const PairParams &p, hipStream_t st)
{
    long n = p.s_end - p.s_begin;
    if (n <= 0)
        return hipSuccess;
    unsigned g = nblocks(n, kWavesPerBlock);
    switch (op) {
    case OP_FWD: hipLaunchKernelGGL(k_pair<OP_FWD>, dim3(g), dim3(kBlock), 0, st, p); break;
    case OP_FWD_V: hipLaunchKernelGGL(k_pair<OP_FWD_V>, dim3(g), dim3(kBlock), 0, st, p); break;
    case OP_ADJ_MAP: hipLaunchKernelGGL(k_pair<OP_ADJ_MAP>, dim3(g), dim3(kBlock), 0, st, p); break;
    case OP_ADJ_BEAM: hipLaunchKernelGGL(k_pair<OP_ADJ_BEAM>, dim3(g), dim3(kBlock), 0, st, p); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool VT>
static hipError_t launch_fwd_sub_vt(int sub, const PairParams &p, long n, hipStream_t st)
{
    switch (sub) {
    case 8:
        hipLaunchKernelGGL((k_fwd_sub<8, VT>), dim3(nblocks(n, kWavesPerBlock * 8)), dim3(kBlock),
                           0, st, p);
        break;
    case 16:
        hipLaunchKernelGGL((k_fwd_sub<16, VT>), dim3(nblocks(n, kWavesPerBlock * 4)),
                           dim3(kBlock), 0, st, p);
        break;
    case 32:
        hipLaunchKernelGGL((k_fwd_sub<32, VT>), dim3(nblocks(n, kWavesPerBlock * 2)),
                           dim3(kBlock), 0, st, p);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fwd_sub(int sub, const PairParams &p, hipStream_t st, bool vterm)
{
    long n = p.s_end - p.s_begin;
    if (n <= 0)
        return hipSuccess;
    return vterm ? launch_fwd_sub_vt<true>(sub, p, n, st) : launch_fwd_sub_vt<false>(sub, p, n, st);
}

hipError_t launch_pair_interp(int op, const PairParams &p, hipStream_t st)
{
    long n = p.s_end - p.s_begin;
    if (n <= 0)
        return hipSuccess;
    unsigned g = nblocks(n, kWavesPerBlock);
    switch (op) {
    case OP_FWD: hipLaunchKernelGGL(k_pair_interp<OP_FWD>, dim3(g), dim3(kBlock), 0, st, p); break;
    case OP_ADJ_MAP:
        hipLaunchKernelGGL(k_pair_interp<OP_ADJ_MAP>, dim3(g), dim3(kBlock), 0, st, p);
        break;
    case OP_ADJ_BEAM:
        hipLaunchKernelGGL(k_pair_interp<OP_ADJ_BEAM>, dim3(g), dim3(kBlock), 0, st, p);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_forward(int group, const BatchParams &p, hipStream_t st)
{
    const long n = p.s_end - p.s_begin;
    if (n <= 0)
        return hipSuccess;
    const unsigned g = nblocks(batch_items(n, p.ngroups), kWavesPerBlock);
    switch (group) {
    case 1:
        hipLaunchKernelGGL(k_batch_pair<OP_FWD>,
                           dim3(nblocks(batch_items(n, p.npairs), kWavesPerBlock)),
                           dim3(kBlock), 0, st, p);
        break;
    case 2: hipLaunchKernelGGL(k_batch_fwd<2>, dim3(g), dim3(kBlock), 0, st, p); break;
    case 4: hipLaunchKernelGGL(k_batch_fwd<4>, dim3(g), dim3(kBlock), 0, st, p); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

size_t adjmap_lds_bytes(int cap, int rw, int group, bool det)
{
    const size_t wt = group == 2 ? sizeof(BatchTable<2>)
                      : group == 4 ? sizeof(BatchTable<4>)
                                   : sizeof(RunTable);
    return wt * kAdjWaves + sizeof(int) * (3 * (size_t)rw + 1 + kAdjCtl) +
           (det ? 8 + 8 * 3 * (size_t)cap : sizeof(float) * 3 * (size_t)cap);
}

/* Raises the dynamic-LDS limit of kernel fn on the current device to at
 * least `bytes`.  `configured[dev]` holds the largest limit set so far for
 * that kernel and device, so a later launch that needs more LDS than an
 * earlier one raises the limit again (it is never lowered). */
constexpr int kDevs = 64;
static hipError_t ensure_dyn_lds(const void *fn, size_t bytes, std::atomic<size_t> *configured)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    if (dev >= 0 && dev < kDevs && bytes <= configured[dev].load())
        return hipSuccess;
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess)
        return e;
    if (dev >= 0 && dev < kDevs) {
        size_t cur = configured[dev].load();
        while (bytes > cur && !configured[dev].compare_exchange_weak(cur, bytes)) {
        }
    }
    return hipSuccess;
}

template <int BG, bool DET, int SUB = kWaveSize>
static hipError_t launch_adjmap_bg(const PairParams &p, const AdjLayout &L, size_t bytes,
                                   long g, hipStream_t st)
{
    static std::atomic<size_t> configured[kDevs];
    hipError_t e = ensure_dyn_lds((const void *)k_adjmap_lds<BG, DET, SUB>, bytes, configured);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL((k_adjmap_lds<BG, DET, SUB>), dim3((unsigned)g), dim3(kAdjBlock), bytes, st,
                       p, L);
    return hipGetLastError();
}

hipError_t launch_adjmap_lds(const PairParams &p, int chunk, int cap, int rw, int nblocks,
                             hipStream_t st, const BatchParams *batch, double rho_w, int group,
                             const double *fx_scale, int lanes)
{
    long n = p.s_end - p.s_begin;
    if (n <= 0)
        return hipSuccess;
    AdjLayout L;
    memset(&L, 0, sizeof L);
    L.chunk = chunk;
    L.cap = cap;
    L.rw = rw;
    L.nchunks = (n + chunk - 1) / chunk;
    if (batch) {
        L.pairs = batch->pairs;
        L.npairs = batch->npairs;
        L.rho_w = rho_w;
        L.cos_w = pl_cos(rho_w);
        L.b = *batch;
    }
    const int bg = batch ? group : 0;
    const bool det = fx_scale != nullptr;
    L.fx_scale = fx_scale;
    const size_t bytes = adjmap_lds_bytes(cap, rw, bg, det);
    /* nblocks 0: one block per window.  Blocks start in index order, so the
     * windows in flight on an XCD stay consecutive in the scan; persistent
     * blocks (one per CU) drift apart over the launch, and measured 10%
     * slower on the 5-degree, NSIDE 512 sky adjoint. */
    const long g = nblocks > 0 ? std::min<long>(L.nchunks, nblocks) : L.nchunks;
    if (bg == 0 && lanes != kWaveSize) {   /* several samples per wave */
        switch (lanes) {
        case 8:
            return det ? launch_adjmap_bg<0, true, 8>(p, L, bytes, g, st)
                       : launch_adjmap_bg<0, false, 8>(p, L, bytes, g, st);
        case 16:
            return det ? launch_adjmap_bg<0, true, 16>(p, L, bytes, g, st)
                       : launch_adjmap_bg<0, false, 16>(p, L, bytes, g, st);
        default: return hipErrorInvalidValue;
        }
    }
    if (det) {
        switch (bg) {
        case 0: return launch_adjmap_bg<0, true>(p, L, bytes, g, st);
        case 1: return launch_adjmap_bg<1, true>(p, L, bytes, g, st);
        case 2: return launch_adjmap_bg<2, true>(p, L, bytes, g, st);
        case 4: return launch_adjmap_bg<4, true>(p, L, bytes, g, st);
        default: return hipErrorInvalidValue;
        }
    }
    switch (bg) {
    case 0: return launch_adjmap_bg<0, false>(p, L, bytes, g, st);
    case 1: return launch_adjmap_bg<1, false>(p, L, bytes, g, st);
    case 2: return launch_adjmap_bg<2, false>(p, L, bytes, g, st);
    case 4: return launch_adjmap_bg<4, false>(p, L, bytes, g, st);
    default: return hipErrorInvalidValue;
    }
}

int adjbeam_band_rows()
{
    return (int)((163840 - sizeof(BandRunTable) * kAdjWaves) / (6 * sizeof(float)));
}

hipError_t launch_adjbeam_bands(const PairParams &p, int chunk, int rows, int nblocks, int ncu,
                                int lanes, hipStream_t st)
{
    const long n = p.s_end - p.s_begin;
    if (n <= 0 || p.npix_beam <= 0)
        return hipSuccess;
    if (chunk < 1)
        return hipErrorInvalidValue;
    BeamBands B;
    B.rows = adjbeam_band_rows();
    if (rows > 0 && rows < B.rows)
        B.rows = rows;
    B.rows = std::min<long>(B.rows, p.npix_beam);   /* no LDS for rows past the beam */
    B.nband = (p.npix_beam + B.rows - 1) / B.rows;
    /* smaller items when there are too few to give every block two */
    const long fill = n * B.nband / (2L * std::max(ncu, 1));
    B.chunk = (int)std::max<long>(std::min<long>(chunk, std::max<long>(fill, 64)), 1);
    B.nchunks = (n + B.chunk - 1) / B.chunk;
    const size_t bytes = sizeof(BandRunTable) * kAdjWaves + 6 * sizeof(float) * (size_t)B.rows;
    if (lanes != 64 && lanes != 8)
        return hipErrorInvalidValue;
    const void *fn = lanes == 8 ? (const void *)k_adjbeam_band_sub<8> : (const void *)k_adjbeam_band;
    static std::atomic<size_t> configured[2][kDevs];
    hipError_t e = ensure_dyn_lds(fn, bytes, configured[lanes == 8 ? 1 : 0]);
    if (e != hipSuccess)
        return e;
    const long items = B.nchunks * B.nband;
    const long g = nblocks > 0 ? std::min<long>(items, nblocks) : items;   /* 0: one per item */
    if (lanes == 8)
        hipLaunchKernelGGL(k_adjbeam_band_sub<8>, dim3((unsigned)g), dim3(kAdjBlock), bytes, st, p,
                           B);
    else
        hipLaunchKernelGGL(k_adjbeam_band, dim3((unsigned)g), dim3(kAdjBlock), bytes, st, p, B);
    return hipGetLastError();
}

bool adjbeam_rec_plan(const PairParams &p, int rows, size_t pool_bytes, int ncu, RecPlan &pl)
{
    const long n = p.s_end - p.s_begin;
    /* six 64-bit sums per row in the band sums' LDS */
    const int max_rows = std::min<int>(1 << kRecRowBits, 163840 / (6 * 8));
    pl.rows = rows > 0 && rows < max_rows ? rows : max_rows;
    pl.rows = (int)std::max<long>(1, std::min<long>(pl.rows, p.npix_beam));
    pl.nband = (int)((p.npix_beam + pl.rows - 1) / pl.rows);
    if (pl.nband > kRecMaxBands || n <= 0)
        return false;
    /* Visits per sample, at most: the sky pixels whose centres lie within
     * rho_max are disjoint, each of area 4 pi / npix, and lie inside the cap
     * of radius rho_max + 2 s, s = sqrt(4 pi / npix) (more than any HEALPix
     * pixel's centre-to-boundary distance). */
    const double om = 4.0 * M_PI / (double)p.hs.npix;
    const double rc = std::min(M_PI, p.rho_max + 2.0 * std::sqrt(om));
    pl.bound = (long)(2.0 * M_PI * (1.0 - std::cos(rc)) / om) + 1;
    pl.bound = std::min<long>(pl.bound, p.hs.npix);
    /* per sample: its records and its offsets per band */
    const double per_sample = (double)sizeof(BeamRec) * pl.bound + 8.0 * pl.nband;
    double chunk = (double)pool_bytes / per_sample;
    chunk = std::max(chunk, (double)kBlock);
    chunk = std::min(chunk, (double)((1L << (32 - kRecRowBits)) - 1));
    chunk = std::min(chunk, (double)(0x7fffffffL / pl.bound));   /* 32-bit record positions */
    pl.chunk = std::max<long>(1, std::min<long>((long)chunk, n));
    /* parts per band: about eight blocks per CU (one fits a CU at a time),
     * so that the last round of blocks is a small part of the launch */
    pl.nsplit = (int)std::max<long>(1, std::min<long>(64, (8L * std::max(ncu, 1) + pl.nband - 1) /
                                                               pl.nband));
    pl.nrec = (size_t)pl.chunk * (size_t)pl.bound;
    pl.nidx = (size_t)pl.chunk * pl.nband + pl.nband + 1;
    pl.npart = (size_t)pl.nsplit * 6 * (size_t)p.npix_beam;
    return true;
}

hipError_t launch_adjbeam_records(const PairParams &p0, const RecPlan &pl, void *recs, long *idx,
                                  float *vmax, double *part, hipStream_t st)
{
    if (p0.s_end <= p0.s_begin || p0.npix_beam <= 0)
        return hipSuccess;
    RecParams R;
    R.rows = pl.rows;
    R.nband = pl.nband;
    R.inv_rows = 1.0f / (float)pl.rows;
    R.brec = (BeamRec *)recs;
    R.nrec_cap = (long)pl.nrec;
    R.bcnt = idx;
    R.bbase = idx + (size_t)pl.chunk * pl.nband;
    R.chunk = pl.chunk;
    R.vmax = vmax;
    R.part = part;
    R.nsplit = pl.nsplit;
    const size_t bytes = 6 * sizeof(unsigned long long) * (size_t)pl.rows;
    static std::atomic<size_t> configured[kDevs];
    hipError_t e = ensure_dyn_lds((const void *)k_beamrec_sum, bytes, configured);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_fill_f64_n, dim3(gs_blocks((long)pl.npart)), dim3(256), 0, st, part,
                       (long)pl.npart);
    hipLaunchKernelGGL(k_fill_f64_n, dim3(1), dim3(64), 0, st, (double *)vmax, 1L);
    hipLaunchKernelGGL(k_rec_vmax, dim3(1024), dim3(256), 0, st, p0, vmax);
    for (long s0 = p0.s_begin; s0 < p0.s_end; s0 += pl.chunk) {
        PairParams p = p0;
        p.s_begin = s0;
        p.s_end = std::min(s0 + pl.chunk, p0.s_end);
        const long ns = p.s_end - p.s_begin;
        const dim3 g1((unsigned)((ns + kWavesPerBlock - 1) / kWavesPerBlock));
        hipLaunchKernelGGL(k_beamrec<false>, g1, dim3(kBlock), 0, st, p, R);
        hipLaunchKernelGGL(k_band_scan, dim3((unsigned)pl.nband), dim3(kScanBlock), 0, st, R, ns);
        hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(kScanBlock), 0, st, R.bbase,
                           (long)pl.nband);
        const dim3 g3((unsigned)((ns + kRecWriteWaves - 1) / kRecWriteWaves));
        hipLaunchKernelGGL(k_beamrec<true>, g3, dim3(kRecWriteWaves * kWaveSize), 0, st, p, R);
        hipLaunchKernelGGL(k_beamrec_sum, dim3((unsigned)(pl.nband * pl.nsplit)), dim3(kAdjBlock),
                           bytes, st, p, R);
    }
    hipLaunchKernelGGL(k_beamrec_reduce, dim3(gs_blocks(6 * (long)p0.npix_beam)), dim3(256), 0, st,
                       part, pl.nsplit, (long)p0.npix_beam, p0.gbeam);
    return hipGetLastError();
}

int adjmap_describe(char *buf, size_t n, int cap, int rw)
{
    hipFuncAttributes a;
    hipError_t e = hipFuncGetAttributes(&a, (const void *)k_adjmap_lds<0, false>);
    if (e != hipSuccess)
        return snprintf(buf, n, "hipFuncGetAttributes: %s", hipGetErrorString(e));
    return snprintf(buf, n,
                    "block %d threads, kernel max %d, regs %d, static LDS %zu, dynamic LDS "
                    "%zu (max %d), private %zu",
                    kAdjBlock, a.maxThreadsPerBlock, a.numRegs, a.sharedSizeBytes,
                    adjmap_lds_bytes(cap, rw, 0, false), a.maxDynamicSharedSizeBytes,
                    a.localSizeBytes);
}

hipError_t launch_beam_support(const float *beam8, long npix, int *last, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, st, last, -1);
    if (npix > 0)
        hipLaunchKernelGGL(k_beam_support, dim3(gs_blocks(npix)), dim3(256), 0, st,
                           (const float4 *)beam8, npix, last);
    return hipGetLastError();
}

hipError_t launch_ring_table(const Hpx &h, RingRec *out, hipStream_t st)
{
    long n = 4L * h.nside - 1;
    hipLaunchKernelGGL(k_ring_table, dim3(nblocks(n, 256)), dim3(256), 0, st, h, out);
    return hipGetLastError();
}

hipError_t launch_zero_f64(double *x, long n, hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3(gs_blocks(n)), dim3(256), 0, st, x, n);
    return hipGetLastError();
}

hipError_t launch_add_offset(const float *x, float off, float *y, long n, hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_add_offset, dim3(gs_blocks(n)), dim3(256), 0, st, x, off, y, n);
    return hipGetLastError();
}

hipError_t launch_pack_sky(const float *I, const float *Q, const float *U, long npix,
                           float4 *sky4, hipStream_t st, const float *V)
{
    if (npix <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_pack_sky, dim3(gs_blocks(npix)), dim3(256), 0, st, I, Q, U, V, npix,
                       sky4);
    return hipGetLastError();
}

hipError_t launch_fix_scale(const float *const *beams, const long *npix, int nbeams,
                            const float *g_a, const float *g_b, long nsamples, long count,
                            unsigned *work, double *scale, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, work, 2, 0u);
    for (int k = 0; k < nbeams; k++)
        if (npix[k] > 0)
            hipLaunchKernelGGL(k_absmax_rows, dim3(gs_blocks(npix[k])), dim3(256), 0, st,
                               beams[k], npix[k], work);
    if (g_a && nsamples > 0)
        hipLaunchKernelGGL(k_absmax, dim3(gs_blocks(nsamples)), dim3(256), 0, st, g_a, nsamples,
                           work + 1);
    if (g_b && nsamples > 0)
        hipLaunchKernelGGL(k_absmax, dim3(gs_blocks(nsamples)), dim3(256), 0, st, g_b, nsamples,
                           work + 1);
    hipLaunchKernelGGL(k_fix_scale, dim3(1), dim3(1), 0, st, (const unsigned *)work, count,
                       scale);
    return hipGetLastError();
}

hipError_t launch_fix_to_f32(const double *x, const double *scale, float *y, long n,
                             hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_fix_to_f32, dim3(gs_blocks(n)), dim3(256), 0, st,
                       (const unsigned long long *)x, scale, y, n);
    return hipGetLastError();
}

hipError_t launch_f64_to_f32(const double *x, float *y, long n, hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(gs_blocks(n)), dim3(256), 0, st, x, y, n);
    return hipGetLastError();
}

hipError_t launch_pack_beam(long npix, const float *const in[8], float *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_pack_beam, dim3(gs_blocks(npix)), dim3(256), 0, st, npix, in[0],
                       in[1], in[2], in[3], in[4], in[5], in[6], in[7], (float4 *)out);
    return hipGetLastError();
}

hipError_t launch_pix2ang(const Hpx &h, long n, const int *pix, double *th, double *ph,
                          hipStream_t st)
{
    hipLaunchKernelGGL(k_pix2ang, dim3(gs_blocks(n)), dim3(256), 0, st, h, n, pix, th, ph);
    return hipGetLastError();
}

hipError_t launch_ang2pix(const Hpx &h, long n, const double *th, const double *ph,
                          int *pix, hipStream_t st)
{
    hipLaunchKernelGGL(k_ang2pix, dim3(gs_blocks(n)), dim3(256), 0, st, h, n, th, ph, pix);
    return hipGetLastError();
}

hipError_t launch_interpol(const Hpx &h, long n, const double *th, const double *ph,
                           int *pix, double *wgt, hipStream_t st)
{
    hipLaunchKernelGGL(k_interpol, dim3(gs_blocks(n)), dim3(256), 0, st, h, n, th, ph, pix,
                       wgt);
    return hipGetLastError();
}

hipError_t launch_disc(const Hpx &h, long n, const float *ra, const float *dec,
                       double radius, double cosr, const long *offsets, int *counts,
                       int *pixels, hipStream_t st)
{
    hipLaunchKernelGGL(k_disc, dim3(nblocks(n, 64)), dim3(64), 0, st, h, n, ra, dec, radius,
                       cosr, offsets, counts, pixels);
    return hipGetLastError();
}

hipError_t launch_match(const Hpx &hs, const RingRec *rings, const Hpx &hb, int mode, long n,
                        const float *ra, const float *dec, const float *pa,
                        const long *offsets, const int *pixels, int *bpix, double *chi,
                        hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_match, dim3((unsigned)n), dim3(256), 0, st, hs, rings, hb,
                       fast_consts(hb), fast_consts64(hb), mode, n, ra, dec, pa, offsets, pixels,
                       bpix, chi);
    return hipGetLastError();
}

/* Constants of beam_pixel_f32 for a beam map.  The rho > 1e-6 band edges are
 * 1 - cos(x) = x^2/2 (the x^4/24 term is below float resolution). */
FastConsts fast_consts(const Hpx &hb)
{
    FastConsts f;
    f.omc_lo = (float)(0.5 * 0.5e-6 * 0.5e-6);
    f.omc_hi = (float)(0.5 * 1.5e-6 * 1.5e-6);
    f.marg_k = (float)(1.2e-13 * hb.nside);
    f.nside = (float)hb.nside;
    return f;
}

/* Constants of beam_pixel_f64: thresholds of the rho > 1e-6 test on crho
 * (rho >= 1.5e-6 is surely above 1e-6, rho <= 0.5e-6 surely below) and the
 * floor margin factor. */
FastConsts64 fast_consts64(const Hpx &hb)
{
    FastConsts64 f;
    f.crho_big = pl_cos(1.5e-6);
    f.crho_small = pl_cos(0.5e-6);
    f.marg_k = 100.0 * 4.4e-16 * hb.nside * 1.7320508075688772;
    return f;
}

hipError_t launch_libm(int op, long n, const double *x, const double *y, double *o,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_libm, dim3(gs_blocks(n)), dim3(256), 0, st, op, n, x, y, o);
    return hipGetLastError();
}

} /* namespace pisco */
