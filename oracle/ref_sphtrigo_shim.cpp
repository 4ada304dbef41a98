// C-callable wrapper around the reference's own rho_sigma_chi_pix
// (src/Sphtrigo/sphtrigo.cpp:4-55), compiled from the reference sources by
// oracle/Makefile into oracle/_ref/.  Test infrastructure only.
#include "Sphtrigo/sphtrigo.hpp"

extern "C" void ref_rho_sigma_chi(long n, const double *ra_bc,
                                  const double *dec_bc, const double *psi_bc,
                                  const double *ra_pix, const double *dec_pix,
                                  double *rho, double *sigma, double *chi)
{
    for (long i = 0; i < n; i++)
        SphericalTransformations::rho_sigma_chi_pix(
            &rho[i], &sigma[i], &chi[i], ra_bc[i], dec_bc[i], psi_bc[i],
            ra_pix[i], dec_pix[i]);
}
