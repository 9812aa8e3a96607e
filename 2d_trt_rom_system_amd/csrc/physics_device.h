// Device restatements of the reference temperature functions (MainSrc/TEMP_FUNCTIONS.f08).
#pragma once
#include "trt_internal.h"

namespace trt {

// sigint, TEMP_FUNCTIONS.f08:114-141.  The band 1.95 < z <= 2.05 is filled by the
// reference's unvendored Interpolator; linear interpolation between the branch
// end points is used here.
__device__ inline double d_sigint(double z)
{
    const double pi4_15 = PI * PI * PI * PI / 15.0;
    if (z <= 1.95) return z * z * z * (1.0 / 3.0 - z / 8.0 + z * z / 62.4);
    if (z > 2.05) return pi4_15 - exp(-z) * (z * z * z + 3.0 * z * z + 6.0 * z + 7.28);
    const double x1 = 1.95, x2 = 2.05;
    const double y1 = x1 * x1 * x1 * (1.0 / 3.0 - x1 / 8.0 + x1 * x1 / 62.4);
    const double y2 = pi4_15 - exp(-x2) * (x2 * x2 * x2 + 3.0 * x2 * x2 + 6.0 * x2 + 7.28);
    return y1 + (y2 - y1) * (z - x1) / (x2 - x1);
}

// Bg_planck_calc, TEMP_FUNCTIONS.f08:90-108
__device__ inline double d_bg_planck(double T, double nl, double nh, double cu)
{
    if (T <= 1e-15) return 0.0;
    double sig = d_sigint(nh / T) - d_sigint(nl / T);
    sig = sig / (H_PLANCK * H_PLANCK * H_PLANCK * C_CONST * C_CONST * ERG * ERG * ERG * ERG * cu);
    return 2.0 * T * T * T * T * sig;
}

// kapB_calc, TEMP_FUNCTIONS.f08:11-32
__device__ inline double d_kapB(double T, double nl, double nh)
{
    const double sig0 = 27.0e9;
    const double z1 = nh / T, z2 = nl / T;
    if (z2 > 10.0) {
        const double sig = -exp(-(z1 - z2)) * (z1 * z1 * z1 + 3.0 * z1 * z1 + 6.0 * z1 + 7.28)
                           + (z2 * z2 * z2 + 3.0 * z2 * z2 + 6.0 * z2 + 7.28);
        return sig0 * (1.0 - exp(-(nh - nl) / T)) / (T * T * T * sig);
    }
    const double sig = d_sigint(z1) - d_sigint(z2);
    return sig0 * (exp(-nl / T) - exp(-nh / T)) / (T * T * T * sig);
}

// group-midpoint opacity sigma_g = 27 keV^3/cm (1 - exp(-nu_g/T)) / nu_g^3
__device__ inline double d_kap_mid(double T, double nl, double nh)
{
    const double nu = 0.5 * (nl + nh);
    return 27.0e9 * (1.0 - exp(-nu / T)) / (nu * nu * nu);
}

}  // namespace trt
