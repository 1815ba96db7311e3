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

// calc: the leaf is seen (mu > 0, not behind a disk).  psi, sin_psi and
// cos_psi are always set; mu, sin_alpha and cos_alpha when psi lies within the
// deflection table; deriv and phase_lag when the leaf is seen (with need_all).
struct LeafGeo {
    int calc;
    double phase_lag, mu, sin_alpha, cos_alpha, deriv, sin_psi, cos_psi, psi;
};

// Disk blocking (xpsi/cellmesh/common_functions.pyx disk_block, after
// Ibragimov & Poutanen 2009, eqs. B9 and C2): the radius at which the ray
// crosses the disk plane; a southern ray is seen only when that radius lies
// inside the disk's inner edge R_in.  Returns 1 when the ray is not blocked.
__device__ inline int disk_unblocked(double R_in, double cos_i, double cos_psi, double cos_th, double u,
                                     double radius, double sin_alpha, double th_over_pi)
{
    const double cpd = (cos_i * cos_psi - cos_th) /
                       sqrt(cos_i * cos_i + cos_th * cos_th - 2 * cos_i * cos_th * cos_psi);
    const double spd = sqrt(1 - cpd * cpd);
    const double r_s = u * radius;
    const double b = radius * sin_alpha / sqrt(1 - u);
    const double r_d = sqrt((r_s * r_s * (1 - cpd) * (1 - cpd)) / (4 * (1 + cpd) * (1 + cpd)) + ((b * b) / (spd * spd))) -
                       (r_s * (1 - cpd)) / (2 * (1 + cpd));
    return (th_over_pi < 0.5 || (th_over_pi > 0.5 && r_d < R_in)) ? 1 : 0;
}

// Polarisation angle of the emitted light at the observer
// (xpsi/cellmesh/common_functions.pyx compute_pol_ang; the oblate Schwarzschild
// approximation of Loktev et al. 2020): the flat-space angle chi_0 plus the
// surface-tilt term chi_1 and the aberration term chi'.
__device__ inline double pol_angle(double leaf, double sin_psi, double cos_psi, double sin_alpha, double cos_alpha,
                                   double sin_th, double cos_th, double sin_i, double cos_i, double sin_gamma,
                                   double cos_gamma, double Grav_z, double mu, double eta, double beta,
                                   double Lorentz, double cos_xi)
{
    const double chi_0 = atan2(-sin_th * sin(leaf), sin_i * cos_th - sin_th * cos_i * cos(leaf));
    const double sa_sp = !eq12(sin_psi, 0.0) ? sin_alpha / sin_psi : Grav_z;
    const double chi_1 = atan2(sin_gamma * sin_i * sin(leaf) * sa_sp, cos_gamma - cos_alpha * mu);
    const double sin_l = sin_th * cos_gamma - sin_gamma * cos_th;
    const double cos_l = cos_th * cos_gamma + sin_th * sin_gamma;
    const double cos_eps = sa_sp * (cos_i * sin_l - sin_i * cos_l * cos(leaf) + cos_psi * sin_gamma) -
                           cos_alpha * sin_gamma;
    const double chi_p = atan2(cos_eps * eta * mu * beta / Lorentz, 1.0 - mu * mu / (1.0 + beta * cos_xi));
    return chi_0 + chi_1 + chi_p;
}

// geometry of leaf k for image order I (integrator :343-414); a disk with
// inner edge R_in < 1e6 m blocks rays (:391-398; u = r_s / r and radius of the
// parallel), R_in >= 1e6 means no disk
__device__ inline LeafGeo leaf_geometry(int I, double leaf, double inclination, double cos_i, double sin_i,
                                        double cos_th, double sin_th, double th_over_pi, double cos_gamma,
                                        double sin_gamma, double maxD, const double* D, const double* A,
                                        const double* Lg, int N_R, const double* xalt, const double* yalt,
                                        int nalt, bool need_all, double R_in = 1.0e6, double u = 0.0,
                                        double radius = 0.0)
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
    g.cos_psi = cos_psi;
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
    if (g.calc && R_in < 1.0e6) g.calc = disk_unblocked(R_in, cos_i, cos_psi, cos_th, u, radius, sa, th_over_pi);
    g.mu = mu;
    g.sin_alpha = sa;
    g.cos_alpha = ca;
    g.cos_psi = cos_psi;
    if (!g.calc || !need_all) return g;
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
