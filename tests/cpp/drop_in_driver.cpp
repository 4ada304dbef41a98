/*
 * drop_in_driver.cpp -- a C++ program written against the reference's class
 * surface (Sky, Scan, PolBeam, GPUConvolver, Convolver; the order of calls of
 * GPUConvolver: ctor, configure_execution, update_*, exec_convolution,
 * gpuconvolver.hpp:21-34), built against include/pisco.hpp and libpisco.
 *
 * It synthesises a small seeded case, runs both convolvers and writes every
 * input and output to one binary file that the GPU test compares with the
 * CPU oracle:
 *   int32 nside_sky, nside_beam; int64 npix_beam, nsamples;
 *   float32 ra, dec, pa [nsamples]; I, Q, U [12 nside_sky^2];
 *   aBeams[0..3], bBeams[0..3] [npix_beam];
 *   tod_a, tod_b (GPUConvolver) [nsamples]; tod_a (Convolver, 'a') [nsamples];
 *   int32 flags: 1 if Convolver rejected polFlag 'x' with
 *   std::invalid_argument, 1 if GPUConvolver rejected ptgPerConv = 0 the same way;
 *   with "crosspolar": float64 magEco, phsEco, magEcx, phsEcx of detector a,
 *   then of detector b [npix_beam each].
 *
 * usage: drop_in_driver OUTPUT [crosspolar]
 *   crosspolar: the beams come from PolBeam::beam_from_fields with a
 *   cross-polar field of its own phase (polbeam.cpp:117-184), so every beam
 *   map is non-zero; otherwise make_unpol_gaussian_elliptical_beams.
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "pisco.hpp"

int main(int argc, char **argv)
{
    if (argc != 2 && argc != 3) {
        std::fprintf(stderr, "usage: %s OUTPUT [crosspolar]\n", argv[0]);
        return 2;
    }
    const bool xpol = argc == 3 && std::string(argv[2]) == "crosspolar";
    const int nside_sky = 64, nside_beam = 512;
    const double deg = M_PI / 180.0;
    const long npix_beam = pisco_ang2pix(nside_beam, 4.0 * deg, 0.0);
    const long npix_sky = 12L * nside_sky * nside_sky;
    const long n = 700;

    std::mt19937_64 rng(2024);
    std::normal_distribution<float> gauss(0.0f, 1.0f);
    std::vector<float> I(npix_sky), Q(npix_sky), U(npix_sky), V(npix_sky, 0.0f);
    for (long p = 0; p < npix_sky; p++) {
        I[p] = 10.0f * gauss(rng);
        Q[p] = gauss(rng);
        U[p] = gauss(rng);
    }
    std::vector<float> ra(n), dec(n), pa(n);
    for (long i = 0; i < n; i++) {   /* a raster over 20 x 20 deg, pa cycling */
        ra[i] = (float)((40.0 + 20.0 * (double)(i % 35) / 35.0) * deg);
        dec[i] = (float)((-30.0 + 20.0 * (double)(i / 35) / 20.0) * deg);
        pa[i] = (float)(45.0 * (double)(i % 4) * deg);
    }

    Sky sky(nside_sky, I.data(), Q.data(), U.data(), V.data());
    Scan scan(n, ra.data(), dec.data(), pa.data());
    PolBeam beam(nside_beam, npix_beam, 0.01, 'p');
    std::vector<double> fields[8];   /* magEco, phsEco, magEcx, phsEcx of a, then b */
    if (xpol) {
        /* co-polar: sqrt of a Gaussian of 1.5 deg FWHM with a quadratic phase;
         * cross-polar: 0.3 (a) or 0.2 (b) of it, modulated in sigma, with its
         * own phase */
        const double sig = 1.5 * deg / 2.35482;
        for (auto &v : fields)
            v.assign((size_t)npix_beam, 0.0);
        for (long p = 0; p < npix_beam; p++) {
            double th = 0.0, ph = 0.0;
            pisco_pix2ang(nside_beam, p, &th, &ph);
            const double u = th * th / (sig * sig);
            const double co = std::exp(-0.25 * u);
            for (int d = 0; d < 2; d++) {
                fields[4 * d + 0][p] = co;
                fields[4 * d + 1][p] = 0.2 * u;
                fields[4 * d + 2][p] = (d ? 0.2 : 0.3) * co * (0.5 + 0.5 * std::cos(ph - d));
                fields[4 * d + 3][p] = (d ? -0.8 : 0.5) + 0.4 * std::sin(ph);
            }
        }
        beam.beam_from_fields('a', fields[0].data(), fields[1].data(), fields[2].data(),
                              fields[3].data());
        beam.beam_from_fields('b', fields[4].data(), fields[5].data(), fields[6].data(),
                              fields[7].data());
    } else {
        beam.make_unpol_gaussian_elliptical_beams(1.5, 1.4, 0.0);
    }
    beam.build(nside_sky);

    CUDACONV::RunConfig cfg{};
    cfg.nStreams = 1;
    cfg.maxMemUsage = (size_t)1 << 40;
    cfg.deviceId = 0;
    cfg.ptgPerConv = 256;   /* several launches and a tail */
    cfg.pixelsPerDisc = 0;
    GPUConvolver gpu(cfg);
    gpu.configure_execution();
    gpu.update_sky(&sky);
    gpu.update_beam(&beam);
    gpu.update_scan(&scan);
    std::vector<float> ta(n, 0.0f), tb(n, 0.0f), ta2(n, 0.0f), unused(n, 0.0f);
    gpu.exec_convolution(ta.data(), tb.data());

    Convolver cpu_api(&scan, &sky, &beam);
    cpu_api.exec_convolution('a', ta2.data(), unused.data());

    int32_t flags[2] = {0, 0};
    try {
        cpu_api.exec_convolution('x', ta2.data(), unused.data());
    } catch (const std::invalid_argument &) {
        flags[0] = 1;
    }
    CUDACONV::RunConfig bad = cfg;
    bad.ptgPerConv = 0;
    GPUConvolver rejected(bad);
    try {
        rejected.configure_execution();
    } catch (const std::invalid_argument &) {
        flags[1] = 1;
    }

    FILE *f = std::fopen(argv[1], "wb");
    if (!f)
        return 1;
    int32_t hdr[2] = {nside_sky, nside_beam};
    int64_t sz[2] = {npix_beam, n};
    std::fwrite(hdr, sizeof hdr, 1, f);
    std::fwrite(sz, sizeof sz, 1, f);
    for (auto *v : {&ra, &dec, &pa, &I, &Q, &U})
        std::fwrite(v->data(), sizeof(float), v->size(), f);
    for (int k = 0; k < 4; k++)
        std::fwrite(beam.aBeams[k], sizeof(float), (size_t)npix_beam, f);
    for (int k = 0; k < 4; k++)
        std::fwrite(beam.bBeams[k], sizeof(float), (size_t)npix_beam, f);
    for (auto *v : {&ta, &tb, &ta2})
        std::fwrite(v->data(), sizeof(float), v->size(), f);
    std::fwrite(flags, sizeof flags, 1, f);
    if (xpol)
        for (auto &v : fields)
            std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
    return 0;
}
