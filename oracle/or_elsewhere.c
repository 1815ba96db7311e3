/*
 * Oracle: time-invariant spectrum of the star's closed mesh ("elsewhere").
 * TEST INFRASTRUCTURE ONLY.
 * Restates xpsi/cellmesh/integrator_for_time_invariance.pyx:59-338 with one
 * thread and the blackbody atmosphere (elsewhere_wrapper.pyx:23-81 dispatches
 * to hot_BB.pyx:54-98).  Note the geometric factor omits the 1/(1 + beta
 * cos xi) present in the hot-region integrator, as the reference does.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include "xpsi_oracle.h"

static const double C_SI = 2.99792458e8;
static const double KEV = 1.60217662e-16;
static const double K_B = 1.38064852e-23;
static const double BB_NORM = 1.0e-7 * 5.040366110812353e22;
static int eq12(double x, double y) { return fabs(x - y) < 1.0e-12; }

/* rays.pyx:53-57 */
static double image_deflection(int order, double psi)
{
    if (order % 2 != 0) return (double)(order + 1) * M_PI + pow(-1.0, (double)order) * psi;
    return (double)order * M_PI + pow(-1.0, (double)order) * psi;
}

/*
 * colat[S] ring colatitudes, phis[S] cell azimuths (every ring alike),
 * logT: the uniform elsewhere temperature; flux[N_E] out.
 * Returns 0, or 1 when a deflection falls outside the ray table.
 */
int or_integrate_time_invariant(double omega, double r_s, double inclination,
                                size_t S, double cellArea, const double *radial,
                                const double *r_s_over_r, const double *colat,
                                const double *phis, double logT, size_t N_R,
                                const double *deflection, const double *cos_alpha,
                                const double *maxDeflection, const double *cos_gammaArray,
                                size_t N_E, const double *energies,
                                int image_order_limit, double *flux)
{
    (void)r_s;
    const double hpi = 0.5 * M_PI;
    double sin_i = sin(inclination), cos_i = cos(inclination);
    int image_order = image_order_limit > 0 ? image_order_limit : 0;
    double kT = K_B / KEV * pow(10.0, logT);
    double *cdefl = (double *)malloc(N_R * sizeof(double));
    double *calt = (double *)malloc(N_R * sizeof(double));
    double *sa = (double *)malloc(4 * N_R * sizeof(double));
    double *st = (double *)malloc(4 * N_R * sizeof(double));
    memset(flux, 0, N_E * sizeof(double));
    int terminate = 0;
    for (size_t i = 0; i < S && !terminate; i++) {
        const double *D = deflection + i * N_R, *A = cos_alpha + i * N_R;
        size_t jn = 0;
        while (D[jn] <= hpi) jn++;
        for (size_t k = 0; k < N_R; k++) {
            cdefl[k] = cos(D[N_R - k - 1]);
            calt[k] = A[N_R - k - 1];
        }
        const double *xalt = cdefl + (N_R - jn - 1), *yalt = calt + (N_R - jn - 1);
        size_t nalt = jn + 1;
        or_steffen_init(xalt, yalt, nalt, st, st + N_R, st + 2 * N_R, st + 3 * N_R);
        or_steffen_init(D, A, N_R, sa, sa + N_R, sa + 2 * N_R, sa + 3 * N_R);
        double radius = radial[i];
        double Grav_z = sqrt(1.0 - r_s_over_r[i]);
        double cos_gamma = cos_gammaArray[i];
        double sin_gamma = sqrt(1.0 - cos_gamma * cos_gamma);
        double cos_th = cos(colat[i]), sin_th = sin(colat[i]);
        double thp = colat[i] / M_PI;
        double beta = radius * omega * sin_th / (C_SI * Grav_z);
        double Lorentz = sqrt(1.0 - beta * beta);
        for (size_t j = 0; j < S && !terminate; j++) {
            double c_psi0 = cos_i * cos_th + sin_i * sin_th * cos(phis[j]);
            double psi0 = acos(c_psi0);
            int IO = image_order == 0 ? (int)ceil(maxDeflection[i] / M_PI) : image_order;
            for (int I = 0; I < IO; I++) {
                double cos_psi = c_psi0;
                double psi = image_deflection(I, psi0);
                double sin_psi = sin(psi);
                if (!eq12(psi, 0.0) && eq12(sin_psi, 0.0)) {
                    double ii = cos_i >= 0.0 ? inclination + inclination * 1.0e-6
                                             : inclination - inclination * 1.0e-6;
                    cos_psi = cos(ii) * cos_th + sin(ii) * sin_th * cos(phis[j]);
                    psi = image_deflection(I, acos(cos_psi));
                    sin_psi = sin(psi);
                }
                if (psi > maxDeflection[i]) break;
                if (psi < D[0] || psi > D[N_R - 1]) {
                    printf("psi: %.16e\n", psi);
                    terminate = 1;
                    break;
                }
                int alt = (psi <= hpi && cos_psi >= xalt[0]);
                double ca = alt ? or_steffen_eval(xalt, nalt, st, st + N_R, st + 2 * N_R, st + 3 * N_R, cos_psi)
                                : or_steffen_eval(D, N_R, sa, sa + N_R, sa + 2 * N_R, sa + 3 * N_R, psi);
                double sin_alpha = sqrt(1.0 - ca * ca);
                double mu = ca * cos_gamma;
                if (!eq12(psi, 0.0)) {
                    double cos_delta = (cos_i - cos_th * cos_psi) / (sin_th * sin_psi);
                    if (thp < 0.5) mu = mu + sin_alpha * sin_gamma * cos_delta;
                    else mu = mu - sin_alpha * sin_gamma * cos_delta;
                }
                if (!(mu > 0.0)) continue;
                double eta;
                if (!eq12(sin_psi, 0.0)) {
                    double cos_xi = sin_alpha * sin_i * sin(phis[j]) / sin_psi;
                    eta = Lorentz / (1.0 + beta * cos_xi);
                } else eta = Lorentz;
                double deriv;
                if (alt) deriv = or_steffen_deriv(xalt, nalt, st, st + N_R, st + 2 * N_R, cos_psi);
                else {
                    deriv = or_steffen_deriv(D, N_R, sa, sa + N_R, sa + 2 * N_R, psi);
                    deriv = exp(log(fabs(deriv)) - log(fabs(sin_psi)));
                }
                double Z = eta * Grav_z;
                double GEOM = mu * fabs(deriv) * Grav_z * eta * eta * eta;
                for (size_t e = 0; e < N_E; e++) {
                    double Ep = energies[e] / Z;
                    flux[e] += Ep * Ep * Ep / (exp(Ep / kT) - 1.0) * GEOM;
                }
            }
        }
    }
    for (size_t e = 0; e < N_E; e++) flux[e] *= cellArea * BB_NORM / (energies[e] * KEV);
    free(cdefl); free(calt); free(sa); free(st);
    return terminate;
}
