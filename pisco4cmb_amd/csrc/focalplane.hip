/*
 * focalplane.hip -- detector pointing from array-centre pointing on the
 * device: BPoint::focal_plane_to_sky (src/Bpoint/bpoint.cpp:553-674).
 *
 * One thread per array-centre sample; a block covers 256 samples and
 * kFpDets detectors, whose constants (sin/cos of the offsets, offset radius
 * and angle) are computed once into LDS.  The sample's own sines and cosines
 * are shared by the block's detectors.  Every expression keeps the
 * reference's operation order and uses the shared FP64 library, so the
 * results equal the CPU restatement bit for bit.  Outputs are
 * detector-major [ndets][nsamples]; the float outputs are the doubles
 * rounded, ready to be a Scan's ra, dec, pa.
 */
#include <hip/hip_runtime.h>

#include "device_hpx.h"
#include "kernels.hpp"
#include "pisco_libm.h"

namespace pisco {

namespace {

constexpr int kFpDets = 16;
constexpr int kFpThreads = 256;

struct FpDet {
    double cdx, sdx, cdy, sdy, cr, sr, alpha, pol;
};

} // namespace

__global__ __launch_bounds__(kFpThreads) void k_focal_plane(FocalPlaneParams q)
{
    __shared__ FpDet dg[kFpDets];
    const int d0 = blockIdx.y * kFpDets;
    const int nd = min(kFpDets, q.ndets - d0);
    if ((int)threadIdx.x < nd) {
        int det = d0 + threadIdx.x;
        FpDet g;
        double dx = q.dx[det];
        pl_sincos(dx, &g.sdx, &g.cdx);
        pl_sincos(q.dy[det], &g.sdy, &g.cdy);
        g.pol = q.pol[det];
        g.cr = g.cdx * g.cdy;
        g.sr = sqrt((1.0 - g.cr) * (1.0 + g.cr));
        g.alpha = pl_atan2(g.sdy, g.sdx * g.cdy);
        dg[threadIdx.x] = g;
    }
    __syncthreads();
    const long s = (long)blockIdx.x * kFpThreads + threadIdx.x;
    if (s >= q.n)
        return;
    const double xa = q.x_ac[s], rot = q.rot_ac[s];
    double sy, cy, seta, ceta;
    pl_sincos(q.y_ac[s], &sy, &cy);
    pl_sincos(rot, &seta, &ceta);
    const double cl = cy * ceta;
    for (int k = 0; k < nd; k++) {
        const FpDet g = dg[k];
        const long idx = (long)(d0 + k) * q.n + s;
        double gamma = kHalfPi - g.alpha - rot;
        double sg, cg;
        pl_sincos(gamma, &sg, &cg);
        double y = pl_asin(g.cr * sy + g.sr * cy * cg);
        double sl_seps = cy * seta * g.cdx - sy * g.sdx;
        double sl_ceps = sy * g.cdx + cy * seta * g.sdx;
        double pa_det = pl_atan2(sl_seps, cl * g.cdy - sl_ceps * g.sdy);
        double r = pa_det + g.pol;
        double x;
        if (q.flag == 2) {
            x = -pa_det;
        } else {
            double dxd = pl_atan2(g.sr * sg, g.cr * cy - g.sr * sy * cg);
            x = (q.flag == 0) ? xa + dxd : xa - dxd;
        }
        if (q.x)
            q.x[idx] = x;
        if (q.y)
            q.y[idx] = y;
        if (q.rot)
            q.rot[idx] = r;
        if (q.x32)
            q.x32[idx] = (float)x;
        if (q.y32)
            q.y32[idx] = (float)y;
        if (q.rot32)
            q.rot32[idx] = (float)r;
    }
}

hipError_t launch_focal_plane(const FocalPlaneParams &q, hipStream_t st)
{
    if (q.ndets <= 0 || q.n <= 0)
        return hipSuccess;
    dim3 grid((unsigned)((q.n + kFpThreads - 1) / kFpThreads),
              (unsigned)((q.ndets + kFpDets - 1) / kFpDets));
    hipLaunchKernelGGL(k_focal_plane, grid, dim3(kFpThreads), 0, st, q);
    return hipGetLastError();
}

} /* namespace pisco */
