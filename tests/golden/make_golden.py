"""Writes tests/golden/sphtrigo_ref.npz: input/output vectors of the
reference's own SphericalTransformations::rho_sigma_chi_pix
(src/Sphtrigo/sphtrigo.cpp:4-55), computed by that source compiled
unmodified into oracle/_ref/libsphtrigo_ref.so (oracle/Makefile builds it
from the reference tree).  The inputs cover the cases the function branches
on: generic offsets, rho below and above 1e-6, sigma wrapping below 0 and
above 2 pi, chi below -pi, the poles and ra across 0 / 2 pi.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind  # noqa: E402


def inputs(seed=20240101):
    rng = np.random.default_rng(seed)
    n = 4000
    ra_bc = rng.uniform(-7, 7, n)
    dec_bc = rng.uniform(-np.pi / 2, np.pi / 2, n)
    psi = rng.uniform(-7, 7, n)
    # offsets: small (typical beam), tiny (rho near 1e-6), large
    scale = np.concatenate([np.full(n // 2, 0.05), np.full(n // 4, 2e-6),
                            np.full(n - n // 2 - n // 4, 1.0)])
    ra_pix = ra_bc + rng.normal(0, 1, n) * scale / np.maximum(np.cos(dec_bc), 0.05)
    dec_pix = np.clip(dec_bc + rng.normal(0, 1, n) * scale, -np.pi / 2, np.pi / 2)
    # exact coincidences and the poles
    ra_pix[:20] = ra_bc[:20]
    dec_pix[:20] = dec_bc[:20]
    dec_bc[20:30] = np.pi / 2
    dec_bc[30:40] = -np.pi / 2
    ra_pix[40:60] = ra_bc[40:60] + 2 * np.pi
    return ra_bc, dec_bc, psi, ra_pix, dec_pix


def main():
    ref = oracle_bind.reference_sphtrigo()
    if ref is None:
        sys.exit("oracle/_ref/libsphtrigo_ref.so is not built (needs the reference tree)")
    x = inputs()
    rho, sigma, chi = ref(*x)
    np.savez_compressed(os.path.join(HERE, "sphtrigo_ref.npz"), ra_bc=x[0], dec_bc=x[1],
                        psi_bc=x[2], ra_pix=x[3], dec_pix=x[4], rho=rho, sigma=sigma, chi=chi)
    print("wrote sphtrigo_ref.npz (%d vectors)" % x[0].size)


if __name__ == "__main__":
    main()
