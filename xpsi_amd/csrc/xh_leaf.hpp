// Geometry of one leaf of a parallel for one image order: the ray-table
// lookups (Steffen), the emission angle against the surface normal and the
// visibility test of the integrators' leaf loops
// (xpsi/cellmesh/integrator_for_azimuthal_invariance.pyx:343-414,
// xpsi/cellmesh/integrator.pyx:316-376).
#pragma once
#include "xh_common.hpp"

namespace xh {

__device__ __forceinline__ double image_deflection(int order, double psi)
{
    double sgn = (order % 2 == 0) ? 1.0 : -1.0;
    if (order % 2 != 0) return (double)(order + 1) * PI + sgn * psi;
    return (double)order * PI + sgn * psi;
}

struct LeafGeo {
    int calc;
    double phase_lag, mu, sin_alpha, deriv, sin_psi, psi;
};

// geometry of leaf k for image order I (integrator :343-414)
__device__ inline LeafGeo leaf_geometry(int I, double leaf, double inclination, double cos_i, double sin_i,
                                        double cos_th, double sin_th, double th_over_pi, double cos_gamma,
                                        double sin_gamma, double maxD, const double* D, const double* A,
                                        const double* Lg, int N_R, const double* xalt, const double* yalt,
                                        int nalt, bool need_all)
{
    LeafGeo g{};
    double cos_psi = cos_i * cos_th + sin_i * sin_th * cos(leaf);
    double psi = image_deflection(I, acos(cos_psi));
    double sin_psi = sin(psi);
    if (!eq12(psi, 0.0) && eq12(sin_psi, 0.0)) {
        double ii = (cos_i >= 0.0) ? inclination + inclination * 1.0e-6 : inclination - inclination * 1.0e-6;
        cos_psi = cos(ii) * cos_th + sin(ii) * sin_th * cos(leaf);
        psi = image_deflection(I, acos(cos_psi));
        sin_psi = sin(psi);
    }
    g.calc = 0;
    g.psi = psi;
    g.sin_psi = sin_psi;
    if (!(psi <= maxD)) return g;
    bool alt = (psi <= HALF_PI && cos_psi >= xalt[0]);
    double ca = alt ? steffen_eval(xalt, yalt, nalt, cos_psi, nullptr) : steffen_eval(D, A, N_R, psi, nullptr);
    double sa = sqrt(1.0 - ca * ca);
    double mu = ca * cos_gamma;
    if (!eq12(psi, 0.0)) {
        double cos_delta = (cos_i - cos_th * cos_psi) / (sin_th * sin_psi);
        if (th_over_pi < 0.5) mu = mu + sa * sin_gamma * cos_delta;
        else mu = mu - sa * sin_gamma * cos_delta;
    }
    g.calc = mu > 0.0 ? 1 : 0;
    if (!g.calc || !need_all) return g;
    g.mu = mu;
    g.sin_alpha = sa;
    double deriv;
    if (alt) {
        steffen_eval(xalt, yalt, nalt, cos_psi, &deriv);
    } else {
        steffen_eval(D, A, N_R, psi, &deriv);
        deriv = exp(log(fabs(deriv)) - log(fabs(sin_psi)));
    }
    g.deriv = deriv;
    g.phase_lag = steffen_eval(D, Lg, N_R, psi, nullptr);
    return g;
}

}  // namespace xh
