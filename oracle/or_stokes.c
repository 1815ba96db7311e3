/*
 * Oracle: the Stokes (I, Q, U) pulse integrator exploiting azimuthal
 * invariance, with the hot-region options the plain integrator's restatement
 * leaves out.  TEST INFRASTRUCTURE ONLY.
 *
 * Restates xpsi/cellmesh/integratorIQU_for_azimuthal_invariance.pyx:68-746
 * (one thread).  With Q and U not requested it is the intensity of
 * xpsi/cellmesh/integrator_for_azimuthal_invariance.pyx:70-665 under the same
 * options, which both integrators share:
 *  - disk blocking, R_in < 1e6 (common_functions.pyx disk_block,
 *    integrator_for_azimuthal_invariance.pyx:392-398);
 *  - atmospheres: blackbody (hot_BB.pyx:54-98, atm_ext 1, Q = 0) and the
 *    burst blackbody (hot_BB_burst.pyx:48-118, atm_ext 3: I = B (0.421 +
 *    0.868 mu), Q = I * 0.1171 (mu - 1) / (1 + 3.582 mu));
 *  - beaming (hot_wrapper.pyx:115-199, beam_opt 1, 2 and 3 for the
 *    blackbody; VEC[2..6] = a, b, c, d, n_mu), applied to I only;
 *  - the polarisation angle (common_functions.pyx compute_pol_ang, the
 *    oblate Schwarzschild approximation of Loktev et al. 2020);
 *  - the blackbody elsewhere correction, per cell (correction [S][S][2]).
 * Image deflection: rays.pyx:53-57.
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

static double image_defl(int order, double psi)
{
    if (order % 2 != 0) return (double)(order + 1) * M_PI + pow(-1.0, (double)order) * psi;
    return (double)order * M_PI + pow(-1.0, (double)order) * psi;
}

/* hot_BB.pyx:54-82 */
static double planck(double E, double logT)
{
    double temp = K_B / KEV * pow(10.0, logT);
    return E * E * E / (exp(E / temp) - 1.0);
}

/* the un-beamed intensity of atmosphere 1 or 3 (hot_wrapper.pyx:135-141) */
static double hot_plain(int atm, double E, double mu, const double *VEC)
{
    double I = planck(E, VEC[0]);
    return atm == 3 ? I * (0.421 + 0.868 * mu) : I;
}

/* hot_wrapper.pyx:115-199: eval_hot_I with the beaming options */
double or_hot_I(int atm, int beam_opt, double E, double mu, const double *VEC)
{
    double I_hot = hot_plain(atm, E, mu, VEC);
    if (beam_opt == 0) return I_hot;
    double abb = VEC[2], bbb = VEC[3], cbb = VEC[4], dbb = VEC[5], nimu = VEC[6];
    double I_beam = 0.0;
    if (beam_opt == 1)
        I_beam = (1.0 + abb * pow(E, cbb) * mu + bbb * pow(E, dbb) * mu * mu) * I_hot;
    if (beam_opt == 2) {
        double anorm = 0.5 / (0.5 + (1.0 / 3.0) * abb * pow(E, cbb) + (1.0 / 4.0) * bbb * pow(E, dbb));
        I_beam = anorm * (1.0 + abb * pow(E, cbb) * mu + bbb * pow(E, dbb) * mu * mu) * I_hot;
    }
    if (beam_opt == 3) {
        /* trapezoid over mu in (0, 1]; the blackbody at each node */
        double mu_imu = 0.0, I_nom = 0.0, I_denom = 0.0;
        size_t n = (size_t)nimu;
        for (size_t imu = 0; imu < n; imu++) {
            mu_imu = mu_imu + (1.0 / nimu);
            double dmu = (imu == 0 || imu == n - 1) ? (0.5 / nimu) : (1.0 / nimu);
            double I_imu = planck(E, VEC[0]);
            double f = 1.0 + abb * pow(E, cbb) * mu_imu + bbb * pow(E, dbb) * mu_imu * mu_imu;
            I_denom = I_denom + mu_imu * f * I_imu * dmu;
            I_nom = I_nom + mu_imu * I_imu * dmu;
        }
        if (eq12(I_denom, 0.0)) I_beam = 0.0;
        else I_beam = (I_nom / I_denom) * (1.0 + abb * pow(E, cbb) * mu + bbb * pow(E, dbb) * mu * mu) * I_hot;
    }
    if (I_beam < 0.0) return 0.0;
    return I_beam;
}

/* hot_wrapper.pyx:201-233: Stokes Q of the atmosphere (no beaming) */
double or_hot_Q(int atm, double E, double mu, const double *VEC)
{
    if (atm != 3) return 0.0;
    double I = hot_plain(3, E, mu, VEC);
    double PD = 0.1171 * (mu - 1.0) / (1.0 + 3.582 * mu);
    return PD * I;
}

/* common_functions.pyx compute_pol_ang */
double or_pol_angle(double leaf, double sin_psi, double cos_psi, double sin_alpha, double cos_alpha,
                    double sin_theta, double cos_theta, double sin_i, double cos_i, double sin_gamma,
                    double cos_gamma, double Grav_z, double mu, double eta, double beta, double Lorentz,
                    double cos_xi)
{
    double chi_0 = atan2(-sin_theta * sin(leaf), sin_i * cos_theta - sin_theta * cos_i * cos(leaf));
    double sa_sp = !eq12(sin_psi, 0.0) ? sin_alpha / sin_psi : Grav_z;
    double chi_1 = atan2(sin_gamma * sin_i * sin(leaf) * sa_sp, cos_gamma - cos_alpha * mu);
    double sin_l = sin_theta * cos_gamma - sin_gamma * cos_theta;
    double cos_l = cos_theta * cos_gamma + sin_theta * sin_gamma;
    double cos_eps = sa_sp * (cos_i * sin_l - sin_i * cos_l * cos(leaf) + cos_psi * sin_gamma) - cos_alpha * sin_gamma;
    double chi_p = atan2(cos_eps * eta * mu * beta / Lorentz, 1.0 - mu * mu / (1.0 + beta * cos_xi));
    return chi_0 + chi_1 + chi_p;
}

/* common_functions.pyx disk_block (Ibragimov & Poutanen 2009): 1 = not blocked */
int or_disk_block(double R_in, double cos_i, double cos_psi, double cos_theta, double r_s_over_r, double radius,
                  double sin_alpha, double theta_over_pi)
{
    double cpd = (cos_i * cos_psi - cos_theta) /
                 sqrt(cos_i * cos_i + cos_theta * cos_theta - 2 * cos_i * cos_theta * cos_psi);
    double spd = sqrt(1 - cpd * cpd);
    double r_s = r_s_over_r * radius;
    double b = radius * sin_alpha / sqrt(1 - r_s_over_r);
    double r_d = sqrt((r_s * r_s * (1 - cpd) * (1 - cpd)) / (4 * (1 + cpd) * (1 + cpd)) + ((b * b) / (spd * spd))) -
                 (r_s * (1 - cpd)) / (2 * (1 + cpd));
    if (theta_over_pi < 0.5 || (theta_over_pi > 0.5 && r_d < R_in)) return 1;
    return 0;
}

int or_integrate_stokes(double omega, double inclination, size_t S_rows, size_t S_cols, const double *cellArea,
                        const double *radial, const double *r_s_over_r, const double *theta, const double *phi,
                        const double *cellParams, size_t n_params, const int *radiates, size_t N_R,
                        const double *deflection, const double *cos_alpha, const double *lag,
                        const double *maxDeflection, const double *cos_gammaArray, size_t N_E,
                        const double *energies, size_t N_L, const double *leaves, size_t N_P,
                        const double *phases, int atm_ext, int beam_opt, int image_order_limit, double R_in,
                        const double *correction, double *flux_I, double *flux_Q, double *flux_U)
{
    const double hpi = 0.5 * M_PI, two_pi = 2.0 * M_PI;
    const double sin_i = sin(inclination), cos_i = cos(inclination);
    const size_t leaf_lim = (N_L % 2 == 0) ? N_L / 2 : (N_L + 1) / 2;
    const int image_order = image_order_limit > 0 ? image_order_limit : 0;
    const int pol = flux_Q != NULL && flux_U != NULL;
    /* beam_opt 3 evaluates the user hook for the burst atmosphere (hot_wrapper.pyx:181-182) */
    if ((atm_ext != 1 && atm_ext != 3) || beam_opt < 0 || beam_opt > 3 || (beam_opt == 3 && atm_ext != 1) ||
        (beam_opt != 0 && n_params < 7) || n_params < 2)
        return 2;
    memset(flux_I, 0, N_E * N_P * sizeof(double));
    if (pol) {
        memset(flux_Q, 0, N_E * N_P * sizeof(double));
        memset(flux_U, 0, N_E * N_P * sizeof(double));
    }
    double *PHASE = (double *)malloc(N_L * sizeof(double));
    double *PI_ = (double *)calloc(N_E * N_L, sizeof(double));
    double *PQ = (double *)calloc(N_E * N_L, sizeof(double));
    double *PU = (double *)calloc(N_E * N_L, sizeof(double));
    double *cdefl = (double *)malloc(N_R * sizeof(double));
    double *calt = (double *)malloc(N_R * sizeof(double));
    double *sa = (double *)malloc(4 * N_R * sizeof(double));
    double *sl = (double *)malloc(4 * N_R * sizeof(double));
    double *st = (double *)malloc(4 * N_R * sizeof(double));
    double *ai = (double *)malloc(3 * N_L * sizeof(double));
    double *aq = (double *)malloc(3 * N_L * sizeof(double));
    double *au = (double *)malloc(3 * N_L * sizeof(double));
    double *mb = (double *)malloc((3 * N_L + 3) * sizeof(double));
    int terminate = 0;
    for (size_t i = 0; i < S_rows && !terminate; i++) {
        size_t J = 0, j;
        for (j = 0; j < S_cols; j++) if (radiates[i * S_cols + j] == 1) { J = j; break; }
        if (j == S_cols) continue;
        const double *D = deflection + i * N_R, *A = cos_alpha + i * N_R, *L = lag + i * N_R;
        size_t jn = 0;
        while (D[jn] <= hpi) jn++;
        for (size_t k = 0; k < N_R; k++) {
            cdefl[k] = cos(D[N_R - k - 1]);
            calt[k] = A[N_R - k - 1];
        }
        const double *xalt = cdefl + (N_R - jn - 1), *yalt = calt + (N_R - jn - 1);
        const size_t nalt = jn + 1;
        or_steffen_init(xalt, yalt, nalt, st, st + N_R, st + 2 * N_R, st + 3 * N_R);
        or_steffen_init(D, A, N_R, sa, sa + N_R, sa + 2 * N_R, sa + 3 * N_R);
        or_steffen_init(D, L, N_R, sl, sl + N_R, sl + 2 * N_R, sl + 3 * N_R);
        const double radius = radial[i];
        const double Grav_z = sqrt(1.0 - r_s_over_r[i]);
        const double cos_gamma = cos_gammaArray[i];
        const double sin_gamma = sqrt(1.0 - cos_gamma * cos_gamma);
        const double th_i = theta[i * S_cols];
        const double cos_th = cos(th_i), sin_th = sin(th_i), th_over_pi = th_i / M_PI;
        const double beta = radius * omega * sin_th / (C_SI * Grav_z);
        const double Lorentz = sqrt(1.0 - beta * beta);
        double c_alpha = -1.0, deriv = -1.0;
        const double *VEC = cellParams + (i * S_cols + J) * n_params;
        const double *CV = correction ? correction + (i * S_cols + J) * 2 : NULL;
        const int IO = image_order == 0 ? (int)ceil(maxDeflection[i] / M_PI) : image_order;
        for (int I = 0; I < IO; I++) {
            int invis = 2;
            size_t invis_phase = 0;
            for (size_t k = 0; k < leaf_lim; k++) {
                double cos_psi = cos_i * cos_th + sin_i * sin_th * cos(leaves[k]);
                double psi = image_defl(I, acos(cos_psi));
                double sin_psi = sin(psi);
                if (!eq12(psi, 0.0) && eq12(sin_psi, 0.0)) {
                    double ii = (cos_i >= 0.0) ? inclination + inclination * 1.0e-6
                                               : inclination - inclination * 1.0e-6;
                    cos_psi = cos(ii) * cos_th + sin(ii) * sin_th * cos(leaves[k]);
                    psi = image_defl(I, acos(cos_psi));
                    sin_psi = sin(psi);
                }
                int calc = 0;
                double mu = 0.0, sin_alpha = 0.0;
                if (psi <= maxDeflection[i]) {
                    if (psi < D[0] || psi > D[N_R - 1]) {
                        printf("Interpolation error: deflection = %.16e\n", psi);
                        terminate = 1;
                        break;
                    }
                    if (psi <= hpi && cos_psi >= xalt[0])
                        c_alpha = or_steffen_eval(xalt, nalt, st, st + N_R, st + 2 * N_R, st + 3 * N_R, cos_psi);
                    else
                        c_alpha = or_steffen_eval(D, N_R, sa, sa + N_R, sa + 2 * N_R, sa + 3 * N_R, psi);
                    sin_alpha = sqrt(1.0 - c_alpha * c_alpha);
                    mu = c_alpha * cos_gamma;
                    if (!eq12(psi, 0.0)) {
                        double cos_delta = (cos_i - cos_th * cos_psi) / (sin_th * sin_psi);
                        if (th_over_pi < 0.5) mu = mu + sin_alpha * sin_gamma * cos_delta;
                        else mu = mu - sin_alpha * sin_gamma * cos_delta;
                    }
                    if (mu > 0.0)
                        calc = R_in < 1e6 ? or_disk_block(R_in, cos_i, cos_psi, cos_th, r_s_over_r[i], radius,
                                                          sin_alpha, th_over_pi)
                                          : 1;
                }
                if (calc) {
                    if (psi <= hpi && cos_psi >= xalt[0]) {
                        deriv = or_steffen_deriv(xalt, nalt, st, st + N_R, st + 2 * N_R, cos_psi);
                    } else {
                        deriv = or_steffen_deriv(D, N_R, sa, sa + N_R, sa + 2 * N_R, psi);
                        deriv = exp(log(fabs(deriv)) - log(fabs(sin_psi)));
                    }
                    double plag = or_steffen_eval(D, N_R, sl, sl + N_R, sl + 2 * N_R, sl + 3 * N_R, psi);
                    for (int ks = 0; ks < 2; ks++) {
                        int take = (0 < k && k < leaf_lim - 1) || (k == 0 && ks == 0) ||
                                   (k == leaf_lim - 1 && N_L % 2 == 1 && ks == 0) ||
                                   (k == leaf_lim - 1 && N_L % 2 == 0);
                        if (!take) continue;
                        size_t kd = ks == 0 ? k : N_L - 1 - k;
                        double cos_xi, superlum, eta;
                        if (!eq12(psi, 0.0)) {
                            cos_xi = sin_alpha * sin_i * sin(leaves[kd]) / sin_psi;
                            superlum = 1.0 + beta * cos_xi;
                            eta = Lorentz / superlum;
                        } else {
                            cos_xi = 0.0;
                            superlum = 1.0;
                            eta = Lorentz;
                        }
                        double Z = eta * Grav_z;
                        double ABB = mu * eta;
                        double GEOM = mu * fabs(deriv) * Grav_z * eta * eta * eta / superlum;
                        PHASE[kd] = leaves[kd] + plag;
                        double c2 = 0.0, s2 = 0.0;
                        if (pol) {
                            double chi = or_pol_angle(leaves[kd], sin_psi, cos_psi, sin_alpha, c_alpha, sin_th, cos_th,
                                                      sin_i, cos_i, sin_gamma, cos_gamma, Grav_z, mu, eta, beta,
                                                      Lorentz, cos_xi);
                            c2 = cos(2 * chi);
                            s2 = sin(2 * chi);
                        }
                        for (size_t p = 0; p < N_E; p++) {
                            double Ep = energies[p] / Z;
                            double IE = or_hot_I(atm_ext, beam_opt, Ep, ABB, VEC);
                            double cIE = CV ? planck(Ep, CV[0]) * BB_NORM : 0.0;
                            PI_[p * N_L + kd] = (IE * BB_NORM - cIE) * GEOM;
                            if (pol) {
                                double QE = or_hot_Q(atm_ext, Ep, ABB, VEC);
                                PQ[p * N_L + kd] = (QE * c2 * BB_NORM) * GEOM;
                                PU[p * N_L + kd] = (QE * s2 * BB_NORM) * GEOM;
                            }
                        }
                    }
                    if (k == 0) {
                        PHASE[N_L - 1] = PHASE[0] + two_pi;
                        for (size_t p = 0; p < N_E; p++) {
                            PI_[p * N_L + N_L - 1] = PI_[p * N_L];
                            PQ[p * N_L + N_L - 1] = PQ[p * N_L];
                            PU[p * N_L + N_L - 1] = PU[p * N_L];
                        }
                    } else if (invis == 2) {
                        double step = leaves[k] / (double)k;
                        for (size_t m = N_L - k; m < N_L; m++) PHASE[m] = PHASE[m - 1] + step;
                        for (size_t p = 0; p < N_E; p++)
                            for (size_t m = N_L - k; m < N_L; m++)
                                PI_[p * N_L + m] = PQ[p * N_L + m] = PU[p * N_L + m] = 0.0;
                        PHASE[0] = PHASE[N_L - 1] - two_pi;
                        for (size_t m = 1; m < k; m++) PHASE[m] = PHASE[m - 1] + step;
                        for (size_t p = 0; p < N_E; p++) {
                            PI_[p * N_L] = PQ[p * N_L] = PU[p * N_L] = 0.0;
                            for (size_t m = 1; m < k; m++) PI_[p * N_L + m] = PQ[p * N_L + m] = PU[p * N_L + m] = 0.0;
                        }
                    } else if (invis == 1) {
                        double step = (PHASE[k] - PHASE[invis_phase - 1]) / (double)(k - invis_phase + 1);
                        for (size_t m = invis_phase; m < k; m++) PHASE[m] = PHASE[m - 1] + step;
                        step = (PHASE[N_L - invis_phase] - PHASE[N_L - 1 - k]) / (double)(k - invis_phase + 1);
                        for (size_t m = N_L - k; m < N_L - invis_phase; m++) PHASE[m] = PHASE[m - 1] + step;
                    }
                    invis = 0;
                } else if (invis == 0) {
                    double step = (PHASE[N_L - k] - PHASE[k - 1]) / (double)(N_L - 2 * k + 1);
                    for (size_t m = k; m < N_L - k; m++) PHASE[m] = PHASE[m - 1] + step;
                    for (size_t p = 0; p < N_E; p++)
                        for (size_t m = k; m < N_L - k; m++) PI_[p * N_L + m] = PQ[p * N_L + m] = PU[p * N_L + m] = 0.0;
                    invis = 1;
                    invis_phase = k;
                }
            }
            if (terminate) break;
            if (invis == 2) break;
            for (size_t m = 1; m < N_L; m++) {
                if (PHASE[m] <= PHASE[m - 1]) {
                    printf("Interpolation error: phases are not strictly increasing.\n");
                    terminate = 1;
                    break;
                }
            }
            if (terminate) break;
            for (size_t p = 0; p < N_E && !terminate; p++) {
                const double *yi = PI_ + p * N_L, *yq = PQ + p * N_L, *yu = PU + p * N_L;
                or_phase_init(PHASE, yi, N_L, ai, ai + N_L, ai + 2 * N_L, mb);
                if (pol) {
                    or_phase_init(PHASE, yq, N_L, aq, aq + N_L, aq + 2 * N_L, mb);
                    or_phase_init(PHASE, yu, N_L, au, au + N_L, au + 2 * N_L, mb);
                }
                for (size_t jj = 0; jj < S_cols && !terminate; jj++) {
                    if (radiates[i * S_cols + jj] != 1) continue;
                    const double shift = phi[i * S_cols + jj], area = cellArea[i * S_cols + jj];
                    for (size_t k = 0; k < N_P; k++) {
                        double x = phases[k] + shift;
                        if (x > PHASE[N_L - 1]) { while (x > PHASE[N_L - 1]) x -= two_pi; }
                        else if (x < PHASE[0]) { while (x < PHASE[0]) x += two_pi; }
                        if (x < PHASE[0] || x > PHASE[N_L - 1]) {
                            printf("Interpolation error: phase = %.16e\n", x);
                            terminate = 1;
                            break;
                        }
                        double vi = or_akima_eval(PHASE, yi, N_L, ai, ai + N_L, ai + 2 * N_L, x);
                        if (vi > 0.0 || correction) {
                            flux_I[p * N_P + k] += area * vi;
                            if (pol) {
                                flux_Q[p * N_P + k] += area * or_akima_eval(PHASE, yq, N_L, aq, aq + N_L, aq + 2 * N_L, x);
                                flux_U[p * N_P + k] += area * or_akima_eval(PHASE, yu, N_L, au, au + N_L, au + 2 * N_L, x);
                            }
                        }
                    }
                }
            }
            if (terminate) break;
        }
    }
    for (size_t p = 0; p < N_E; p++)
        for (size_t k = 0; k < N_P; k++) {
            flux_I[p * N_P + k] /= energies[p] * KEV;
            if (pol) {
                flux_Q[p * N_P + k] /= energies[p] * KEV;
                flux_U[p * N_P + k] /= energies[p] * KEV;
            }
        }
    free(PHASE); free(PI_); free(PQ); free(PU); free(cdefl); free(calt); free(sa); free(sl); free(st);
    free(ai); free(aq); free(au); free(mb);
    return terminate;
}
