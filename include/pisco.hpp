/*
 * pisco.hpp -- C++ drop-in classes over the C-ABI of libpisco (pisco.h).
 *
 * The classes keep the reference's names, constructors and methods so that
 * a program written against PISCO's headers builds against this one:
 *   Sky          src/Sky/sky.hpp:12-38            (non-owning view)
 *   Scan         src/Scan/scan.hpp:9-34           (non-owning view)
 *   PolBeam      src/Polbeam/polbeam.hpp:23-93    (owns its maps)
 *   GPUConvolver src/GpuConvolver/gpuconvolver.hpp:17-82
 *   Convolver    src/Convolver/convolver.hpp:14-36 (computed on the GPU)
 *   SkyMap       src/SkyMap/skymap.hpp:13-55      (normal equations on the GPU)
 *   BPoint       src/Bpoint/BPoint.h:30-103, focal_plane_to_sky only
 *   CUDACONV::RunConfig src/GpuConvolver/cuda/cudaConv2.h:13-26
 * Errors are exceptions: std::invalid_argument where the reference threw it
 * (invalid configuration, gpuconvolver.cu:74-78; bad polFlag,
 * convolver.cpp:66-70) and std::runtime_error where it printed and exited.
 *
 * The classes live in namespace pisco and are also made visible at global
 * scope under the reference's names; define PISCO_NO_GLOBAL_NAMES before
 * including this header to keep them namespaced only.
 *
 * Link with -lpisco (pisco4cmb_amd/libpisco.so).
 */
#ifndef PISCO_HPP
#define PISCO_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "pisco.h"

namespace pisco {

inline void check(int rc)
{
    if (rc == PISCO_OK)
        return;
    std::string msg = pisco_last_error();
    if (rc == PISCO_ERR_INVALID)
        throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

using RunConfig = pisco_run_config;

class Sky {
  public:
    Sky(int nside_, const float *I, const float *Q, const float *U, const float *V)
        : sI(I), sQ(Q), sU(U), sV(V), nside(nside_), nPixels(12 * nside_ * nside_)
    {
    }
    int get_nside() const { return nside; }
    int size() const { return nPixels; }
    const float *sI;
    const float *sQ;
    const float *sU;
    const float *sV;

  private:
    int nside;
    int nPixels;
};

class Scan {
  public:
    Scan(long nsamples, float *ra_, float *dec_, float *pa_)
        : ra(ra_), dec(dec_), pa(pa_), nsamp(nsamples)
    {
    }
    long get_nsamples() const { return nsamp; }
    float *get_ra_ptr() { return ra; }
    float *get_dec_ptr() { return dec; }
    float *get_pa_ptr() { return pa; }

  private:
    float *ra;
    float *dec;
    float *pa;
    long nsamp;
};

class PolBeam {
  public:
    PolBeam(int nside_, long nPixels_, double epsilon_, char enabledDets_)
        : nside(nside_), nPixels(nPixels_), enabledDets(enabledDets_), epsilon(epsilon_)
    {
        if (enabledDets != 'a' && enabledDets != 'b' && enabledDets != 'p')
            throw std::invalid_argument("enabledDets must be 'a', 'b' or 'p'");
        rhoMax = pisco_beam_rho_max(nside, nPixels);
        if (rhoMax < 0)
            throw std::invalid_argument("invalid beam geometry");
        store.assign(16, std::vector<float>((size_t)nPixels, 0.0f));
        float **tilde[8] = {&Ia, &Qa, &Ua, &Va, &Ib, &Qb, &Ub, &Vb};
        for (int k = 0; k < 8; k++)
            *tilde[k] = store[k].data();
        for (int k = 0; k < 4; k++) {
            aBeams[k] = store[8 + k].data();
            bBeams[k] = store[12 + k].data();
        }
    }
    PolBeam(const PolBeam &) = delete;
    PolBeam &operator=(const PolBeam &) = delete;

    int get_nside() const { return nside; }
    unsigned long size() const { return (unsigned long)nPixels; }
    double get_rho_max() const { return rhoMax; }

    /* polbeam.cpp:117-184 */
    void beam_from_fields(char polFlag, double *magEco, double *phsEco, double *magEcx,
                          double *phsEcx)
    {
        if (polFlag != 'a' && polFlag != 'b')
            throw std::invalid_argument("polFlag must be 'a' or 'b'");
        float *I = polFlag == 'a' ? Ia : Ib, *Q = polFlag == 'a' ? Qa : Qb;
        float *U = polFlag == 'a' ? Ua : Ub, *V = polFlag == 'a' ? Va : Vb;
        check(pisco_beam_from_fields(nPixels, magEco, phsEco, magEcx, phsEcx, I, Q, U, V));
    }

    /* polbeam.cpp:186-288 */
    void build(int nsideSky)
    {
        const float *tilde[8] = {Ia, Qa, Ua, Va, Ib, Qb, Ub, Vb};
        float *out[8] = {aBeams[0], aBeams[1], aBeams[2], aBeams[3],
                         bBeams[0], bBeams[1], bBeams[2], bBeams[3]};
        check(pisco_polbeam_build(nPixels, nside, nsideSky, epsilon, enabledDets, tilde, out));
    }

    /* polbeam.cpp:290-372 (FWHM and phi0 in degrees) */
    void make_unpol_gaussian_elliptical_beams(double fwhmx, double fwhmy, double phi0)
    {
        std::vector<double> f[4];
        for (auto &v : f)
            v.assign((size_t)nPixels, 0.0);
        check(pisco_beam_gaussian_fields(nside, nPixels, fwhmx, fwhmy, phi0, f[0].data(),
                                         f[1].data(), f[2].data(), f[3].data()));
        if (enabledDets == 'a' || enabledDets == 'p')
            beam_from_fields('a', f[0].data(), f[1].data(), f[2].data(), f[3].data());
        if (enabledDets == 'b' || enabledDets == 'p')
            beam_from_fields('b', f[0].data(), f[1].data(), f[2].data(), f[3].data());
    }

    float *aBeams[4];
    float *bBeams[4];
    int nside;
    long nPixels;
    char enabledDets;
    float *Ia, *Qa, *Ua, *Va;
    float *Ib, *Qb, *Ub, *Vb;

  private:
    double rhoMax;
    double epsilon;
    std::vector<std::vector<float>> store;
};

class GPUConvolver {
  public:
    explicit GPUConvolver(RunConfig cfg_) : cfg(cfg_) {}
    ~GPUConvolver() { pisco_destroy(ctx); }
    GPUConvolver(const GPUConvolver &) = delete;
    GPUConvolver &operator=(const GPUConvolver &) = delete;

    /* gpuconvolver.cu:66-87: an invalid configuration throws
     * std::invalid_argument */
    void configure_execution()
    {
        if (ctx)
            return;
        int rc = pisco_create(&cfg, &ctx);
        if (rc == PISCO_ERR_INVALID || rc == PISCO_ERR_MEMORY)
            throw std::invalid_argument(std::string("GPUConvolver has an invalid "
                                                    "configuration: ") +
                                        pisco_last_error());
        check(rc);
    }
    void update_sky(Sky *sky)
    {
        check(pisco_update_sky(need(), sky->get_nside(), sky->sI, sky->sQ, sky->sU, sky->sV));
        npixSky = sky->size();
    }
    void update_beam(PolBeam *beam)
    {
        check(pisco_update_beam(need(), beam->get_nside(), (long)beam->size(), beam->aBeams,
                                beam->bBeams));
        npixBeam = (long)beam->size();
    }
    void update_scan(Scan *scan)
    {
        check(pisco_update_scan(need(), scan->get_nsamples(), scan->get_ra_ptr(),
                                scan->get_dec_ptr(), scan->get_pa_ptr()));
    }
    /* gpuconvolver.cu:382-505: both detectors */
    void exec_convolution(float *data_a, float *data_b)
    {
        check(pisco_exec_convolution(need(), 'p', data_a, data_b));
    }
    /* with Convolver's polFlag ('a', 'b' or 'p') */
    void exec_convolution(char polFlag, float *data_a, float *data_b)
    {
        check(pisco_exec_convolution(need(), polFlag, data_a, data_b));
    }
    /* adjoints (no counterpart in the reference) */
    void exec_adjoint_map(const float *g_a, const float *g_b, float *gI, float *gQ, float *gU)
    {
        check(pisco_exec_adjoint_map(need(), g_a, g_b, gI, gQ, gU));
    }
    void exec_adjoint_beam(const float *g_a, const float *g_b, float *const ga[4],
                           float *const gb[4])
    {
        check(pisco_exec_adjoint_beam(need(), g_a, g_b, ga, gb));
    }
    void set_tuning(const char *key, long value) { check(pisco_set_tuning(need(), key, value)); }

  private:
    pisco_ctx *need()
    {
        if (!ctx)
            throw std::runtime_error("configure_execution() must be called first");
        return ctx;
    }
    RunConfig cfg;
    pisco_ctx *ctx = nullptr;
    int npixSky = 0;
    long npixBeam = 0;
};

/* Convolver(Scan*, Sky*, PolBeam*) with exec_convolution(polFlag, a, b);
 * the reference runs it on host cores, this one on the GPU. */
class Convolver {
  public:
    Convolver(Scan *scan_, Sky *sky_, PolBeam *beam_, RunConfig cfg = default_config())
        : scan(scan_), sky(sky_), beam(beam_), gpu(cfg)
    {
        gpu.configure_execution();
    }
    void exec_convolution(char polFlag, float *data_a, float *data_b)
    {
        if (polFlag != 'a' && polFlag != 'b' && polFlag != 'p')
            throw std::invalid_argument("valid polFlags are 'a', 'b' and 'p'");
        gpu.update_sky(sky);
        gpu.update_beam(beam);
        gpu.update_scan(scan);
        gpu.exec_convolution(polFlag, polFlag == 'b' ? nullptr : data_a,
                             polFlag == 'a' ? nullptr : data_b);
    }
    static RunConfig default_config()
    {
        RunConfig c{};
        c.nStreams = 1;
        c.maxMemUsage = (size_t)1 << 40;
        c.deviceId = 0;
        c.ptgPerConv = 1 << 20;
        c.pixelsPerDisc = 0;
        return c;
    }

  private:
    Scan *scan;
    Sky *sky;
    PolBeam *beam;
    GPUConvolver gpu;
};

/* SkyMap (src/SkyMap/skymap.hpp:13-55): binned I, Q, U map-maker whose
 * normal equations live on a GPU; the accumulators start at zero. */
class SkyMap {
  public:
    explicit SkyMap(int nside_, int device = 0)
        : nside(nside_), nPixels(12L * nside_ * nside_)
    {
        check(pisco_skymap_create(device, nside, &map));
        stokesI.assign((size_t)nPixels, 0.0f);
        stokesQ.assign((size_t)nPixels, 0.0f);
        stokesU.assign((size_t)nPixels, 0.0f);
        hitsMap.assign((size_t)nPixels, 0);
    }
    ~SkyMap() { pisco_skymap_destroy(map); }
    SkyMap(const SkyMap &) = delete;
    SkyMap &operator=(const SkyMap &) = delete;

    /* skymap.cpp:68-132; arrays detector-major, ndetectors * nsamples */
    void accumulate_data(int nsamples, float ra[], float dec[], float pa[], int ndetectors,
                         float detPolAngles[], float data[])
    {
        check(pisco_skymap_accumulate(map, nsamples, ra, dec, pa, ndetectors, detPolAngles,
                                      data));
    }
    /* skymap.cpp:134-231 */
    void solve_map(void)
    {
        check(pisco_skymap_solve(map, stokesI.data(), stokesQ.data(), stokesU.data(),
                                 hitsMap.data(), nullptr));
    }
    int get_nside(void) const { return nside; }
    int size(void) const { return (int)nPixels; }
    const float *get_stokes_I(void) { return stokesI.data(); }
    const float *get_stokes_Q(void) { return stokesQ.data(); }
    const float *get_stokes_U(void) { return stokesU.data(); }
    const int *get_hitmap(void) { return hitsMap.data(); }

  private:
    pisco_skymap *map = nullptr;
    int nside;
    long nPixels;
    std::vector<float> stokesI, stokesQ, stokesU;
    std::vector<int> hitsMap;
};

/* BPoint::focal_plane_to_sky (src/Bpoint/bpoint.cpp:553-674) on the GPU,
 * FP64; the rest of BPoint (SOFA astrometry) is not part of this library. */
class BPoint {
  public:
    explicit BPoint(int deviceId = 0)
    {
        RunConfig c = Convolver::default_config();
        c.deviceId = deviceId;
        check(pisco_create(&c, &ctx));
    }
    ~BPoint() { pisco_destroy(ctx); }
    BPoint(const BPoint &) = delete;
    BPoint &operator=(const BPoint &) = delete;
    void focal_plane_to_sky(int ndets, double det_dx[], double det_dy[], double det_pol_angle[],
                            int nsamples, double x_ac[], double y_ac[], double rot_ac[],
                            int flag, double x_det[], double y_det[], double rot_det[])
    {
        check(pisco_focal_plane_to_sky(ctx, ndets, det_dx, det_dy, det_pol_angle, nsamples,
                                       x_ac, y_ac, rot_ac, flag, x_det, y_det, rot_det));
    }

  private:
    pisco_ctx *ctx = nullptr;
};

} /* namespace pisco */

#ifndef PISCO_NO_GLOBAL_NAMES
namespace CUDACONV {
using RunConfig = pisco::RunConfig;
}
using pisco::Convolver;
using pisco::GPUConvolver;
using pisco::PolBeam;
using pisco::BPoint;
using pisco::Scan;
using pisco::Sky;
using pisco::SkyMap;
#endif

#endif /* PISCO_HPP */
