// Closed-form hot-region atmospheres and the beaming options, shared by the
// Stokes and general integrators' kernels:
//  - blackbody (xpsi/surface_radiation_field/hot_BB.pyx:54-98, atm 1; Q = 0);
//  - burst blackbody (hot_BB_burst.pyx:48-118, atm 3: I = B (0.421 + 0.868 mu),
//    Q = I * 0.1171 (mu - 1) / (1 + 3.582 mu));
//  - beaming (hot_wrapper.pyx:115-199, options 1-3; VEC[2..6] = a, b, c, d,
//    n_mu), applied to I.
#pragma once
#include "xh_common.hpp"

namespace xh {

__device__ __forceinline__ double planck(double E, double logT)
{
    const double kT = K_B_OVER_KEV * pow(10.0, logT);
    return E * E * E / (exp(E / kT) - 1.0);
}

__device__ __forceinline__ double hot_plain(int atm, double E, double mu, const double* VEC)
{
    const double I = planck(E, VEC[0]);
    return atm == 3 ? I * (0.421 + 0.868 * mu) : I;
}

// hot_wrapper.pyx:115-199
__device__ inline double hot_I(int atm, int beam, double E, double mu, const double* VEC)
{
    const double I_hot = hot_plain(atm, E, mu, VEC);
    if (beam == 0) return I_hot;
    const double a = VEC[2], b = VEC[3], c = VEC[4], d = VEC[5], nimu = VEC[6];
    const double fa = a * pow(E, c), fb = b * pow(E, d);
    double I_beam = 0.0;
    if (beam == 1) I_beam = (1.0 + fa * mu + fb * mu * mu) * I_hot;
    if (beam == 2) {
        const double anorm = 0.5 / (0.5 + (1.0 / 3.0) * fa + (1.0 / 4.0) * fb);
        I_beam = anorm * (1.0 + fa * mu + fb * mu * mu) * I_hot;
    }
    if (beam == 3) {
        // trapezoid over mu in (0, 1] of the blackbody with and without the
        // beaming factor; their ratio normalises the beamed intensity
        double mu_k = 0.0, nom = 0.0, den = 0.0;
        const long n = (long)nimu;
        const double I_k = planck(E, VEC[0]);
#pragma unroll 1
        for (long k = 0; k < n; k++) {
            mu_k = mu_k + (1.0 / nimu);
            const double dmu = (k == 0 || k == n - 1) ? (0.5 / nimu) : (1.0 / nimu);
            den = den + mu_k * (1.0 + fa * mu_k + fb * mu_k * mu_k) * I_k * dmu;
            nom = nom + mu_k * I_k * dmu;
        }
        I_beam = eq12(den, 0.0) ? 0.0 : (nom / den) * (1.0 + fa * mu + fb * mu * mu) * I_hot;
    }
    return I_beam < 0.0 ? 0.0 : I_beam;
}

// hot_wrapper.pyx:201-233 (no beaming)
__device__ __forceinline__ double hot_Q(int atm, double E, double mu, const double* VEC)
{
    if (atm != 3) return 0.0;
    return 0.1171 * (mu - 1.0) / (1.0 + 3.582 * mu) * hot_plain(3, E, mu, VEC);
}

}  // namespace xh
