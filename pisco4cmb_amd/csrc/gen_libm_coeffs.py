"""Generate the polynomial coefficients and split constants of pisco_libm.h.

The FP64 elementary functions used on the pixel-index path (sin, cos, acos,
atan2) are evaluated with the same sequence of IEEE-754 operations on the host
and on gfx950, so that pixel indices computed on either side agree bit for bit.
This script derives every constant those functions use:

* Chebyshev-node fits (mpmath.chebyfit, near-minimax) of the reduced kernels
    sin(x) = x + x*z*S(z),          z = x^2 in [0, (pi/4)^2]
    cos(x) = 1 - z/2 + z^2*C(z),    z = x^2 in [0, (pi/4)^2]
    asin(x) = x + x*z*A(z),         z = x^2 in [0, 1/4]
    atan(u) = u + u*z*T(z),         z = u^2 in [0, tan(pi/12)^2]
* multi-word splits of pi/2, pi/6, pi/3, pi, 5pi/6, 2pi/3 and sqrt(3).

and the float32 kernels of the kernels' single-precision visit geometry
(pixel indices from it are certified by error margins, never trusted as is):
    sin(x) = x + x*z*SF(z),         z = x^2 in [0, (pi/2)^2]
    cos(x) = 1 - z/2 + z^2*CF(z),   z = x^2 in [0, (pi/2)^2]
    atan(t) = t + t*z*TF(z),        z = t^2 in [0, 1]

Run: python gen_libm_coeffs.py  (prints the C block pasted into pisco_libm.h)
"""
import struct
from mpmath import mp, mpf, sin, cos, asin, atan, sqrt, tan, pi, chebyfit

mp.dps = 60


def hexd(v):
    return float(v).hex()


def split(v, parts):
    out = []
    rem = mpf(v)
    for _ in range(parts):
        h = float(rem)
        out.append(h)
        rem -= mpf(h)
    return out


def fit(g, zmax, deg):
    coeffs, err = chebyfit(g, [0, zmax], deg + 1, error=True)
    return [float(c) for c in reversed(coeffs)], err  # ascending powers


def g_sin(z):
    if z == 0:
        return mpf(-1) / 6
    x = sqrt(z)
    return (sin(x) - x) / (x * z)


def g_cos(z):
    if z == 0:
        return mpf(1) / 24
    x = sqrt(z)
    return (cos(x) - 1 + z / 2) / (z * z)


def g_asin(z):
    if z == 0:
        return mpf(1) / 6
    x = sqrt(z)
    return (asin(x) - x) / (x * z)


def g_atan(z):
    if z == 0:
        return mpf(-1) / 3
    x = sqrt(z)
    return (atan(x) - x) / (x * z)


def main():
    lines = []
    for name, g, zmax, deg in (("S", g_sin, (pi / 4) ** 2 * mpf("1.0001"), 6),
                               ("C", g_cos, (pi / 4) ** 2 * mpf("1.0001"), 5),
                               ("A", g_asin, mpf("0.25"), 13),
                               ("T", g_atan, tan(pi / 12) ** 2 * mpf("1.0001"), 9)):
        c, err = fit(g, zmax, deg)
        lines.append("/* %s(z): degree %d, max abs fit error %.3g on [0, %.6g] */"
                     % (name, deg, float(err), float(zmax)))
        for i, v in enumerate(c):
            lines.append("#define PL_%s%d %s" % (name, i, hexd(v)))
    p1, p2, p3 = split(pi / 2, 3)
    lines += ["#define PL_PIO2_1 %s" % hexd(p1), "#define PL_PIO2_2 %s" % hexd(p2),
              "#define PL_PIO2_3 %s" % hexd(p3),
              "#define PL_TWO_OVER_PI %s" % hexd(2 / pi)]
    for nm, v in (("PI", pi), ("PIO2", pi / 2), ("PIO6", pi / 6), ("PIO3", pi / 3),
                  ("PI5O6", 5 * pi / 6), ("PI2O3", 2 * pi / 3), ("SQRT3", sqrt(3))):
        h, l = split(v, 2)
        lines.append("#define PL_%s_HI %s" % (nm, hexd(h)))
        lines.append("#define PL_%s_LO %s" % (nm, hexd(l)))
    lines.append("#define PL_TAN_PI_12 %s" % hexd(tan(pi / 12)))
    for name, g, zmax, deg in (("SF", g_sin, (pi / 2) ** 2 * mpf("1.0005"), 3),
                               ("CF", g_cos, (pi / 2) ** 2 * mpf("1.0005"), 3),
                               ("TF", g_atan, mpf("1.0001"), 8)):
        c, err = fit(g, zmax, deg)
        lines.append("/* %s(z): float32, degree %d, fit error %.3g on [0, %.6g] */"
                     % (name, deg, float(err), float(zmax)))
        for i, v in enumerate(c):
            f32 = struct.unpack("f", struct.pack("f", v))[0]
            lines.append("#define PL_%s%d %sf" % (name, i, float(f32).hex()))
    print("\n".join(lines))


if __name__ == "__main__":
    main()
