/*
 * noalloc.hip -- pisco_dev_forward_noalloc: the single-pair forward of
 * pisco_dev_forward with its workspace in a fixed device array instead of
 * memory the library allocates, for callers that cannot allocate device
 * memory during the call (code under graph capture, serving loops).
 *
 * The forward's workspace is the sky ring table (one RingRec per ring) and
 * the sky packed as one float4 (I, Q, U, 0) per pixel.  Both live in
 * `scratch`, sized for sky NSIDE up to kNoAllocMaxNside: 4 N RingRec records
 * and 12 N^2 float4 pixels, 48.1 MiB at N = 512.  Every call rewrites both
 * parts completely before the forward reads them, so nothing needs clearing;
 * the forward writes each TOD sample once.  Every call on a device shares
 * that device's copy of the array: calls that overlap in time on one device
 * (on different streams, or from different host threads) race.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>

#include "pisco.h"
#include "kernels.hpp"

using namespace pisco;

namespace {

constexpr int kNoAllocMaxNside = 512;
constexpr size_t kRingBytes = sizeof(RingRec) * 4 * (size_t)kNoAllocMaxNside;
constexpr size_t kSkyOffset = (kRingBytes + 255) / 256 * 256;
constexpr size_t kSkyBytes = sizeof(float4) * 12 * (size_t)kNoAllocMaxNside * kNoAllocMaxNside;
constexpr size_t kScratchBytes = kSkyOffset + kSkyBytes;
constexpr long kChunk = 1L << 22;   /* samples per launch, as the torch binding's contexts */
constexpr int kMaxDevices = 64;

} // namespace

/* The workspace (each device holds its own copy for the life of the
 * process).  Host code never uses this name as an address: it uses the
 * per-device pointer from hipGetSymbolAddress below. */
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

/* Per device: the device address of `scratch`, looked up once with that
 * device current (NULL until then). */
std::atomic<unsigned char *> g_scratch[kMaxDevices];

int errorf(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int errorf(int code, const char *fmt, ...)
{
    char buf[320];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return set_error(code, buf);
}

int scratch_address(unsigned char **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return errorf(PISCO_ERR_DEVICE, "hipGetDevice failed: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return errorf(PISCO_ERR_DEVICE, "device %d beyond the %d the library tracks", dev,
                      kMaxDevices);
    unsigned char *p = g_scratch[dev].load();
    if (!p) {
        e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return errorf(PISCO_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                          hipGetErrorString(e));
        g_scratch[dev].store(p);
    }
    *out = p;
    return PISCO_OK;
}

} // namespace

extern "C" {

int pisco_dev_forward_noalloc(void *stream, int nside_sky, const float *I, const float *Q,
                              const float *U, int nside_beam, long npix_beam,
                              const float *beam8, long nsamples, const float *ra,
                              const float *dec, const float *pa, float *tod_a, float *tod_b)
{
    if (!I || !Q || !U || !beam8 || (nsamples > 0 && (!ra || !dec || !pa)))
        return errorf(PISCO_ERR_INVALID, "NULL argument");
    if (nside_sky > kNoAllocMaxNside)
        return errorf(PISCO_ERR_INVALID,
                      "pisco_dev_forward_noalloc holds skies up to NSIDE %d in its fixed "
                      "%zu-byte workspace (NSIDE %d given); pisco_dev_forward takes any NSIDE",
                      kNoAllocMaxNside, kScratchBytes, nside_sky);
    PairParams p;
    int rc = pair_params(nside_sky, I, Q, U, nside_beam, npix_beam, beam8, nsamples, ra, dec,
                         pa, p);
    if (rc != PISCO_OK)
        return rc;
    unsigned char *base = nullptr;
    rc = scratch_address(&base);
    if (rc != PISCO_OK)
        return rc;
    hipStream_t st = (hipStream_t)stream;
    RingRec *rings = (RingRec *)base;
    float4 *sky4 = (float4 *)(base + kSkyOffset);
    hipError_t e = launch_ring_table(p.hs, rings, st);
    if (e == hipSuccess)
        e = launch_pack_sky(I, Q, U, p.hs.npix, sky4, st);
    if (e != hipSuccess)
        return errorf(PISCO_ERR_DEVICE, "workspace set-up launch failed: %s",
                      hipGetErrorString(e));
    p.rings = rings;
    p.sky4 = sky4;
    p.tod_a = tod_a;
    p.tod_b = tod_b;
    const int lanes = forward_lanes_auto(p);
    for (long s = 0; s < nsamples; s += kChunk) {
        p.s_begin = s;
        p.s_end = std::min(nsamples, s + kChunk);
        e = lanes < 64 ? launch_fwd_sub(lanes, p, st) : launch_pair(OP_FWD, p, st);
        if (e != hipSuccess)
            return errorf(PISCO_ERR_DEVICE, "forward launch failed: %s", hipGetErrorString(e));
    }
    return PISCO_OK;
}

} /* extern "C" */
