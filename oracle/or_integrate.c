/*
 * Oracle: pulse integration over a hot-region mesh exploiting azimuthal
 * invariance.  TEST INFRASTRUCTURE ONLY.
 * Restates xpsi/cellmesh/integrator_for_azimuthal_invariance.pyx:70-665 with
 * one thread (the reference's recommended setting, Likelihood.py:43-50),
 * the blackbody atmosphere xpsi/surface_radiation_field/hot_BB.pyx:54-98 and
 * the 4-D cubic Lagrange table hot_Num4D.pyx:66-460 (its base-node search
 * keeps state from one call to the next, replayed here in call order).
 * Image deflection: rays.pyx:53-57.  Disk blocking (R_in < 1e6) is not used.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include "xpsi_oracle.h"

static const double C_SI = 2.99792458e8;
static const double KEV = 1.60217662e-16;
static const double K_B = 1.38064852e-23;
static int eq12(double x, double y) { return fabs(x - y) < 1.0e-12; }

/* rays.pyx:53-57 */
static double image_deflection(int order, double psi)
{
    if (order % 2 != 0) return (double)(order + 1) * M_PI + pow(-1.0, (double)order) * psi;
    return (double)order * M_PI + pow(-1.0, (double)order) * psi;
}

/* hot_BB.pyx:54-82, norm :84-98 */
static double bb_I(double E, const double *vec)
{
    double kT = K_B / KEV * pow(10.0, vec[0]);
    return E * E * E / (exp(E / kT) - 1.0);
}
static const double BB_NORM = 1.0e-7 * 5.040366110812353e22;
static const double NUM4D_NORM = 1.0e-7 / 4.135667662e-18;

/* per-call state of the 4-D table interpolator (hot_Num4D.pyx:109-199) */
typedef struct {
    const or_atmosphere *a;
    size_t BN[4];
    double node[8];
    double vcache[4];
    double diff[16];
    double space[16];
} n4_state;

static void n4_space(const double *x, size_t bn, double *s)
{
    s[0] = 1.0 / (x[bn] - x[bn + 1]); s[0] /= x[bn] - x[bn + 2]; s[0] /= x[bn] - x[bn + 3];
    s[1] = 1.0 / (x[bn + 1] - x[bn]); s[1] /= x[bn + 1] - x[bn + 2]; s[1] /= x[bn + 1] - x[bn + 3];
    s[2] = 1.0 / (x[bn + 2] - x[bn]); s[2] /= x[bn + 2] - x[bn + 1]; s[2] /= x[bn + 2] - x[bn + 3];
    s[3] = 1.0 / (x[bn + 3] - x[bn]); s[3] /= x[bn + 3] - x[bn + 1]; s[3] /= x[bn + 3] - x[bn + 2];
}

static void n4_diff(const double *x, size_t bn, double v, double *d)
{
    d[0] = v - x[bn + 1]; d[0] *= v - x[bn + 2]; d[0] *= v - x[bn + 3];
    d[1] = v - x[bn]; d[1] *= v - x[bn + 2]; d[1] *= v - x[bn + 3];
    d[2] = v - x[bn]; d[2] *= v - x[bn + 1]; d[2] *= v - x[bn + 3];
    d[3] = v - x[bn]; d[3] *= v - x[bn + 1]; d[3] *= v - x[bn + 2];
}

static void n4_init(n4_state *s, const or_atmosphere *a)
{
    s->a = a;
    for (int i = 0; i < 4; i++) {
        const double *x = a->params[i];
        s->BN[i] = 0;
        s->vcache[i] = x[1];
        s->node[2 * i] = x[1];
        s->node[2 * i + 1] = x[2];
        n4_space(x, 0, s->space + 4 * i);
        n4_diff(x, 0, x[1], s->diff + 4 * i);
    }
}

/* hot_Num4D.pyx:248-438 */
static double n4_I(n4_state *s, double E, double mu, const double *VEC)
{
    const or_atmosphere *a = s->a;
    double v[4];
    v[0] = VEC[0]; v[1] = VEC[1]; v[2] = mu;
    v[3] = log10(E / (K_B / KEV * pow(10.0, VEC[0])));
    for (int i = 0; i < 4; i++) {
        const double *x = a->params[i];
        size_t N = a->N[i];
        int upd = 0;
        if (v[i] < s->node[2 * i] && s->BN[i] != 0) {
            upd = 1;
            while (v[i] < x[s->BN[i] + 1]) {
                if (s->BN[i] > 0) s->BN[i]--;
                else if (v[i] <= x[0]) { v[i] = x[0]; break; }
                else break;
            }
            s->node[2 * i] = x[s->BN[i] + 1];
            s->node[2 * i + 1] = x[s->BN[i] + 2];
        } else if (v[i] > s->node[2 * i + 1] && s->BN[i] != N - 4) {
            upd = 1;
            while (v[i] > x[s->BN[i] + 2]) {
                if (s->BN[i] < N - 4) s->BN[i]++;
                else if (v[i] >= x[N - 1]) { v[i] = x[N - 1]; break; }
                else break;
            }
            s->node[2 * i] = x[s->BN[i] + 1];
            s->node[2 * i + 1] = x[s->BN[i] + 2];
        }
        if (s->vcache[i] != v[i] || upd) {
            n4_diff(x, s->BN[i], v[i], s->diff + 4 * i);
            s->vcache[i] = v[i];
        }
        if (upd) n4_space(x, s->BN[i], s->space + 4 * i);
    }
    size_t S0 = a->N[1] * a->N[2] * a->N[3], S1 = a->N[2] * a->N[3], S2 = a->N[3];
    double I = 0.0;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++)
                for (int l = 0; l < 4; l++) {
                    double w = s->diff[i] * s->diff[4 + j] * s->diff[8 + k] * s->diff[12 + l];
                    w *= s->space[i] * s->space[4 + j] * s->space[8 + k] * s->space[12 + l];
                    size_t at = (s->BN[0] + i) * S0 + (s->BN[1] + j) * S1 +
                                (s->BN[2] + k) * S2 + s->BN[3] + l;
                    I += w * a->intensity[at];
                }
    if (I < 0.0) return 0.0;
    return I * pow(10.0, 3.0 * v[0]);
}

/* The interpolator over a query sequence, starting from the state each
   integrate() call starts from (hot_Num4D.pyx:109-199): out[q] is
   eval_hot_Num4D(E[q], mu[q], vec2[2q..2q+1]) without the norm. */
void or_num4d_sequence(const or_atmosphere *a, size_t n, const double *E,
                       const double *mu, const double *vec2, double *out)
{
    n4_state s;
    n4_init(&s, a);
    for (size_t q = 0; q < n; q++) out[q] = n4_I(&s, E[q], mu[q], vec2 + 2 * q);
}

int or_integrate(double R, double omega, double r_s, double inclination,
                 size_t S_rows, size_t S_cols, const double *cellArea,
                 const double *radial, const double *r_s_over_r,
                 const double *theta, const double *phi,
                 const double *cellParams, const int *radiates, size_t N_R,
                 const double *deflection, const double *cos_alpha,
                 const double *lag, const double *maxDeflection,
                 const double *cos_gammaArray, size_t N_E,
                 const double *energies, size_t N_L, const double *leaves,
                 size_t N_P, const double *phases, int atm_ext,
                 const or_atmosphere *atm, int image_order_limit,
                 const double *correction, double *flux)
{
    (void)R; (void)r_s;
    const double hpi = 0.5 * M_PI, two_pi = 2.0 * M_PI;
    double sin_i = sin(inclination), cos_i = cos(inclination);
    size_t leaf_lim = (N_L % 2 == 0) ? N_L / 2 : (N_L + 1) / 2;
    int image_order = image_order_limit > 0 ? image_order_limit : 0;
    memset(flux, 0, N_E * N_P * sizeof(double));
    double norm = (atm_ext == 2) ? NUM4D_NORM : BB_NORM;
    int terminate = 0;
    /* Built with OpenMP (the CPU baseline), rings are shared out over threads
       as the reference's prange does, each thread with its own scratch, table
       state and flux, summed at the end (:605-612).  Without it this is the
       one-thread restatement, summing in ring order. */
#pragma omp parallel
    {
    double *PHASE = (double *)malloc(N_L * sizeof(double));
    double *PROF = (double *)calloc(N_E * N_L, sizeof(double));
    double *cdefl = (double *)malloc(N_R * sizeof(double));
    double *calt = (double *)malloc(N_R * sizeof(double));
    double *sa = (double *)malloc(4 * N_R * sizeof(double));   /* psi -> cos a */
    double *sl = (double *)malloc(4 * N_R * sizeof(double));   /* psi -> lag */
    double *st = (double *)malloc(4 * N_R * sizeof(double));   /* cos psi -> cos a */
    double *ab = (double *)malloc(3 * N_L * sizeof(double));
    double *mb = (double *)malloc((3 * N_L + 3) * sizeof(double));
    double *acc = (double *)calloc(N_E * N_P, sizeof(double));
    n4_state n4;
    if (atm_ext == 2) n4_init(&n4, atm);
#pragma omp for schedule(dynamic)
    for (long ir = 0; ir < (long)S_rows; ir++) {
        const size_t i = (size_t)ir;
        if (terminate) continue;
        size_t J = 0, j;
        for (j = 0; j < S_cols; j++) if (radiates[i * S_cols + j] == 1) { J = j; break; }
        if (j == S_cols) continue;
        const double *D = deflection + i * N_R, *A = cos_alpha + i * N_R, *L = lag + i * N_R;
        size_t jn = 0;
        while (D[jn] <= hpi) jn++;
        /* reversed tables over cos psi: increasing cos(deflection) */
        for (size_t k = 0; k < N_R; k++) {
            cdefl[k] = cos(D[N_R - k - 1]);
            calt[k] = A[N_R - k - 1];
        }
        const double *xalt = cdefl + (N_R - jn - 1), *yalt = calt + (N_R - jn - 1);
        size_t nalt = jn + 1;
        or_steffen_init(xalt, yalt, nalt, st, st + N_R, st + 2 * N_R, st + 3 * N_R);
        or_steffen_init(D, A, N_R, sa, sa + N_R, sa + 2 * N_R, sa + 3 * N_R);
        or_steffen_init(D, L, N_R, sl, sl + N_R, sl + 2 * N_R, sl + 3 * N_R);
        double radius = radial[i];
        double Grav_z = sqrt(1.0 - r_s_over_r[i]);
        double cos_gamma = cos_gammaArray[i];
        double sin_gamma = sqrt(1.0 - cos_gamma * cos_gamma);
        double th_i = theta[i * S_cols];
        double cos_th = cos(th_i), sin_th = sin(th_i), th_over_pi = th_i / M_PI;
        double beta = radius * omega * sin_th / (C_SI * Grav_z);
        double Lorentz = sqrt(1.0 - beta * beta);
        double c_alpha = -1.0, deriv = -1.0;
        const double *VEC = cellParams + (i * S_cols + J) * 2;
        int IO = image_order == 0 ? (int)ceil(maxDeflection[i] / M_PI) : image_order;
        for (int I = 0; I < IO; I++) {
            int invis = 2;
            size_t invis_phase = 0;
            for (size_t k = 0; k < leaf_lim; k++) {
                double cos_psi = cos_i * cos_th + sin_i * sin_th * cos(leaves[k]);
                double psi = image_deflection(I, acos(cos_psi));
                double sin_psi = sin(psi);
                if (!eq12(psi, 0.0) && eq12(sin_psi, 0.0)) {
                    double ii = (cos_i >= 0.0) ? inclination + inclination * 1.0e-6
                                               : inclination - inclination * 1.0e-6;
                    cos_psi = cos(ii) * cos_th + sin(ii) * sin_th * cos(leaves[k]);
                    psi = image_deflection(I, acos(cos_psi));
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
                    calc = (mu > 0.0) ? 1 : 0;
                }
                if (calc) {
                    if (psi <= hpi && cos_psi >= xalt[0]) {
                        deriv = or_steffen_deriv(xalt, nalt, st, st + N_R, st + 2 * N_R, cos_psi);
                    } else {
                        deriv = or_steffen_deriv(D, N_R, sa, sa + N_R, sa + 2 * N_R, psi);
                        deriv = exp(log(fabs(deriv)) - log(fabs(sin_psi)));
                    }
                    if (psi < D[0] || psi > D[N_R - 1]) {
                        printf("Interpolation error: deflection = %.16e\n", psi);
                        terminate = 1;
                        break;
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
                            superlum = 1.0;
                            eta = Lorentz;
                        }
                        double Z = eta * Grav_z;
                        double ABB = mu * eta;
                        double GEOM = mu * fabs(deriv) * Grav_z * eta * eta * eta / superlum;
                        PHASE[kd] = leaves[kd] + plag;
                        for (size_t p = 0; p < N_E; p++) {
                            double Ep = energies[p] / Z;
                            double IE = (atm_ext == 2) ? n4_I(&n4, Ep, ABB, VEC) : bb_I(Ep, VEC);
                            /* :469-478: the blackbody elsewhere intensity of the cell subtracted */
                            double cIE = correction ? bb_I(Ep, correction + (i * S_cols + J) * 2) * BB_NORM : 0.0;
                            PROF[p * N_L + kd] = (IE * norm - cIE) * GEOM;
                        }
                    }
                    if (k == 0) {
                        PHASE[N_L - 1] = PHASE[0] + two_pi;
                        for (size_t p = 0; p < N_E; p++) PROF[p * N_L + N_L - 1] = PROF[p * N_L];
                    } else if (invis == 2) {
                        double step = leaves[k] / (double)k;
                        for (size_t m = N_L - k; m < N_L; m++) PHASE[m] = PHASE[m - 1] + step;
                        for (size_t p = 0; p < N_E; p++)
                            for (size_t m = N_L - k; m < N_L; m++) PROF[p * N_L + m] = 0.0;
                        PHASE[0] = PHASE[N_L - 1] - two_pi;
                        for (size_t m = 1; m < k; m++) PHASE[m] = PHASE[m - 1] + step;
                        for (size_t p = 0; p < N_E; p++) {
                            PROF[p * N_L] = 0.0;
                            for (size_t m = 1; m < k; m++) PROF[p * N_L + m] = 0.0;
                        }
                    } else if (invis == 1) {
                        double step = (PHASE[k] - PHASE[invis_phase - 1]) / (double)(k - invis_phase + 1);
                        for (size_t m = invis_phase; m < k; m++) PHASE[m] = PHASE[m - 1] + step;
                        step = (PHASE[N_L - invis_phase] - PHASE[N_L - 1 - k]) / (double)(k - invis_phase + 1);
                        for (size_t m = N_L - k; m < N_L - invis_phase; m++) PHASE[m] = PHASE[m - 1] + step;
                    }
                    invis = 0;
                } else {
                    if (invis == 0) {
                        double step = (PHASE[N_L - k] - PHASE[k - 1]) / (double)(N_L - 2 * k + 1);
                        for (size_t m = k; m < N_L - k; m++) PHASE[m] = PHASE[m - 1] + step;
                        for (size_t p = 0; p < N_E; p++)
                            for (size_t m = k; m < N_L - k; m++) PROF[p * N_L + m] = 0.0;
                        invis = 1;
                        invis_phase = k;
                    }
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
                const double *yp = PROF + p * N_L;
                or_phase_init(PHASE, yp, N_L, ab, ab + N_L, ab + 2 * N_L, mb);
                for (size_t jj = 0; jj < S_cols && !terminate; jj++) {
                    if (radiates[i * S_cols + jj] != 1) continue;
                    double shift = phi[i * S_cols + jj];
                    double area = cellArea[i * S_cols + jj];
                    for (size_t k = 0; k < N_P; k++) {
                        double x = phases[k] + shift;
                        if (x > PHASE[N_L - 1]) { while (x > PHASE[N_L - 1]) x -= two_pi; }
                        else if (x < PHASE[0]) { while (x < PHASE[0]) x += two_pi; }
                        if (x < PHASE[0] || x > PHASE[N_L - 1]) {
                            printf("Interpolation error: phase = %.16e\n", x);
                            terminate = 1;
                            break;
                        }
                        double v = or_akima_eval(PHASE, yp, N_L, ab, ab + N_L, ab + 2 * N_L, x);
                        if (v > 0.0 || correction) acc[p * N_P + k] += area * v;
                    }
                }
            }
            if (terminate) break;
        }
    }
#pragma omp critical
    for (size_t q = 0; q < N_E * N_P; q++) flux[q] += acc[q];
    free(PHASE); free(PROF); free(cdefl); free(calt); free(sa); free(sl);
    free(st); free(ab); free(mb); free(acc);
    }
    for (size_t p = 0; p < N_E; p++)
        for (size_t k = 0; k < N_P; k++) flux[p * N_P + k] /= energies[p] * KEV;
    return terminate;
}

/* xpsi/cellmesh/integrator.pyx:48-667, the integrator without azimuthal
 * invariance, for the blackbody (atm_ext 1): every cell carries its own
 * parameter vector (cellParams [S_rows][S_cols][2], log10 T first) and, with
 * the correction, its own elsewhere temperature (correction, same layout).
 * Per parallel and image the leaves' redshift Z, mu eta (ABB), geometric
 * factor GEOM and phase are formed over all N_L leaves (:316-512; invisible
 * stretches padded with linearly stepped phase, Z and ABB and zero GEOM);
 * then for each radiating cell and model phase GEOM is interpolated with the
 * periodic Akima spline and, where positive, Z with a Steffen spline
 * (:531-592; the blackbody ignores the interpolated ABB, which only beamed
 * atmospheres read), and every energy's intensity at E / Z is added with the cell's
 * area times GEOM.  Flux [N_E][N_P], divided by E keV at the end (:601-606).
 * Rings are summed in order (one thread).  The Stokes form
 * (xpsi/cellmesh/integratorIQU.pyx:51-717) adds Q and U: the atmosphere's Q
 * at the interpolated mu eta, rotated by 2 chi, where chi is the polarisation
 * angle the reference forms inside its cell and phase loop from the leaf
 * loop's last values (:593-596), hence one angle per parallel and image. */
int or_integrate_general_ext(double omega, double inclination, size_t S_rows, size_t S_cols,
                             const double *cellArea, const double *radial, const double *r_s_over_r,
                             const double *theta, const double *phi, const double *cellParams,
                             size_t n_params, const int *radiates, size_t N_R, const double *deflection,
                             const double *cos_alpha, const double *lag, const double *maxDeflection,
                             const double *cos_gammaArray, size_t N_E, const double *energies, size_t N_L,
                             const double *leaves, size_t N_P, const double *phases, int image_order_limit,
                             int atm_ext, int beam_opt, const double *correction, double *flux,
                             double *flux_Q, double *flux_U)
{
    const int pol = flux_Q != NULL && flux_U != NULL;
    if ((atm_ext != 1 && atm_ext != 3) || beam_opt < 0 || beam_opt > 3 || (beam_opt == 3 && atm_ext != 1) ||
        (beam_opt != 0 && n_params < 7) || n_params < 2)
        return 2;
    const double hpi = 0.5 * M_PI, two_pi = 2.0 * M_PI;
    double sin_i = sin(inclination), cos_i = cos(inclination);
    size_t leaf_lim = (N_L % 2 == 0) ? N_L / 2 : (N_L + 1) / 2;
    int image_order = image_order_limit > 0 ? image_order_limit : 0;
    memset(flux, 0, N_E * N_P * sizeof(double));
    double *acc = (double *)calloc(N_E * N_P, sizeof(double));
    double *accQ = (double *)calloc(N_E * N_P, sizeof(double));
    double *accU = (double *)calloc(N_E * N_P, sizeof(double));
    /* the leaf loop's last values, which the Stokes form reads for its
       polarisation angle (integratorIQU.pyx:593-596): the last leaf written
       (_kdx, eta, cos_xi), the last leaf (sin psi, cos psi) and the last leaf
       within the deflection table (sin alpha, cos alpha, mu); they persist
       from parallel to parallel as the reference's variables do */
    size_t l_kdx = 0;
    double l_eta = 0.0, l_cxi = 0.0, l_spsi = 0.0, l_cpsi = 0.0, l_sa = 0.0, l_ca = 0.0, l_mu = 0.0;
    double *PHASE = (double *)calloc(N_L, sizeof(double));
    double *ZZ = (double *)calloc(N_L, sizeof(double));
    double *AB = (double *)calloc(N_L, sizeof(double));
    double *GE = (double *)calloc(N_L, sizeof(double));
    double *cdefl = (double *)malloc(N_R * sizeof(double));
    double *calt = (double *)malloc(N_R * sizeof(double));
    double *sa = (double *)malloc(4 * N_R * sizeof(double));
    double *sl = (double *)malloc(4 * N_R * sizeof(double));
    double *st = (double *)malloc(4 * N_R * sizeof(double));
    double *gb = (double *)malloc(3 * N_L * sizeof(double));   /* Akima, GEOM */
    double *mb = (double *)malloc((3 * N_L + 3) * sizeof(double));
    double *zs = (double *)malloc(4 * N_L * sizeof(double));   /* Steffen, Z */
    double *as = (double *)malloc(4 * N_L * sizeof(double));   /* Steffen, ABB */
    int terminate = 0;
    for (size_t i = 0; i < S_rows && !terminate; i++) {
        size_t j;
        for (j = 0; j < S_cols; j++) if (radiates[i * S_cols + j] == 1) break;
        if (j == S_cols) continue;
        const double *D = deflection + i * N_R, *A = cos_alpha + i * N_R, *L = lag + i * N_R;
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
        or_steffen_init(D, L, N_R, sl, sl + N_R, sl + 2 * N_R, sl + 3 * N_R);
        double radius = radial[i];
        double Grav_z = sqrt(1.0 - r_s_over_r[i]);
        double cos_gamma = cos_gammaArray[i];
        double sin_gamma = sqrt(1.0 - cos_gamma * cos_gamma);
        double th_i = theta[i * S_cols];
        double cos_th = cos(th_i), sin_th = sin(th_i), th_over_pi = th_i / M_PI;
        double beta = radius * omega * sin_th / (C_SI * Grav_z);
        double Lorentz = sqrt(1.0 - beta * beta);
        double c_alpha = -1.0, deriv = -1.0;
        int IO = image_order == 0 ? (int)ceil(maxDeflection[i] / M_PI) : image_order;
        for (int I = 0; I < IO; I++) {
            int invis = 2;
            size_t invis_phase = 0;
            double step, zstep, astep;
            for (size_t k = 0; k < N_L; k++) {
                double cos_psi = cos_i * cos_th + sin_i * sin_th * cos(leaves[k]);
                double psi = image_deflection(I, acos(cos_psi));
                double sin_psi = sin(psi);
                if (!eq12(psi, 0.0) && eq12(sin_psi, 0.0)) {
                    double ii = (cos_i >= 0.0) ? inclination + inclination * 1.0e-6
                                               : inclination - inclination * 1.0e-6;
                    cos_psi = cos(ii) * cos_th + sin(ii) * sin_th * cos(leaves[k]);
                    psi = image_deflection(I, acos(cos_psi));
                    sin_psi = sin(psi);
                }
                l_spsi = sin_psi;
                l_cpsi = cos_psi;
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
                    calc = (mu > 0.0) ? 1 : 0;
                    l_sa = sin_alpha;
                    l_ca = c_alpha;
                    l_mu = mu;
                }
                if (calc) {
                    if (psi <= hpi && cos_psi >= xalt[0]) {
                        deriv = or_steffen_deriv(xalt, nalt, st, st + N_R, st + 2 * N_R, cos_psi);
                    } else {
                        deriv = or_steffen_deriv(D, N_R, sa, sa + N_R, sa + 2 * N_R, psi);
                        deriv = exp(log(fabs(deriv)) - log(fabs(sin_psi)));
                    }
                    if (psi < D[0] || psi > D[N_R - 1]) {
                        printf("Interpolation error: deflection = %.16e\n", psi);
                        terminate = 1;
                        break;
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
                        l_kdx = kd;
                        l_eta = eta;
                        l_cxi = cos_xi;
                        ZZ[kd] = eta * Grav_z;
                        AB[kd] = mu * eta;
                        GE[kd] = mu * fabs(deriv) * Grav_z * eta * eta * eta / superlum;
                        PHASE[kd] = leaves[kd] + plag;
                    }
                    if (k == 0) {
                        PHASE[N_L - 1] = PHASE[0] + two_pi;
                        ZZ[N_L - 1] = ZZ[0];
                        AB[N_L - 1] = AB[0];
                        GE[N_L - 1] = GE[0];
                    } else if (invis == 2) {
                        step = leaves[k] / (double)k;
                        zstep = (ZZ[k] - ZZ[N_L - k - 1]) / (2.0 * (double)k);
                        astep = (AB[k] - AB[N_L - k - 1]) / (2.0 * (double)k);
                        for (size_t m = N_L - k; m < N_L; m++) {
                            PHASE[m] = PHASE[m - 1] + step;
                            ZZ[m] = ZZ[m - 1] + zstep;
                            AB[m] = AB[m - 1] + astep;
                            GE[m] = 0.0;
                        }
                        PHASE[0] = PHASE[N_L - 1] - two_pi;
                        ZZ[0] = ZZ[N_L - 1];
                        AB[0] = AB[N_L - 1];
                        GE[0] = GE[N_L - 1];
                        for (size_t m = 1; m < k; m++) {
                            PHASE[m] = PHASE[m - 1] + step;
                            ZZ[m] = ZZ[m - 1] + zstep;
                            AB[m] = AB[m - 1] + astep;
                            GE[m] = 0.0;
                        }
                    } else if (invis == 1) {
                        step = (PHASE[k] - PHASE[invis_phase - 1]) / (double)(k - invis_phase + 1);
                        zstep = (ZZ[k] - ZZ[invis_phase - 1]) / (double)(k - invis_phase + 1);
                        astep = (AB[k] - AB[invis_phase - 1]) / (double)(k - invis_phase + 1);
                        for (size_t m = invis_phase; m < k; m++) {
                            PHASE[m] = PHASE[m - 1] + step;
                            ZZ[m] = ZZ[m - 1] + zstep;
                            AB[m] = AB[m - 1] + astep;
                        }
                        step = (PHASE[N_L - invis_phase] - PHASE[N_L - 1 - k]) / (double)(k - invis_phase + 1);
                        zstep = (ZZ[N_L - invis_phase] - ZZ[N_L - 1 - k]) / (double)(k - invis_phase + 1);
                        astep = (AB[N_L - invis_phase] - AB[N_L - 1 - k]) / (double)(k - invis_phase + 1);
                        for (size_t m = N_L - k; m < N_L - invis_phase; m++) {
                            PHASE[m] = PHASE[m - 1] + step;
                            ZZ[m] = ZZ[m - 1] + zstep;
                            AB[m] = AB[m - 1] + astep;
                        }
                    }
                    invis = 0;
                } else if (invis == 0) {
                    step = (PHASE[N_L - k] - PHASE[k - 1]) / (double)(N_L - 2 * k + 1);
                    zstep = (ZZ[N_L - k] - ZZ[k - 1]) / (double)(N_L - 2 * k + 1);
                    astep = (AB[N_L - k] - AB[k - 1]) / (double)(N_L - 2 * k + 1);
                    for (size_t m = k; m < N_L - k; m++) {
                        PHASE[m] = PHASE[m - 1] + step;
                        ZZ[m] = ZZ[m - 1] + zstep;
                        AB[m] = AB[m - 1] + astep;
                        GE[m] = 0.0;
                    }
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
            or_phase_init(PHASE, GE, N_L, gb, gb + N_L, gb + 2 * N_L, mb);
            or_steffen_init(PHASE, ZZ, N_L, zs, zs + N_L, zs + 2 * N_L, zs + 3 * N_L);
            or_steffen_init(PHASE, AB, N_L, as, as + N_L, as + 2 * N_L, as + 3 * N_L);
            double c2 = 0.0, s2 = 0.0;
            if (pol) {
                const double chi = or_pol_angle(leaves[l_kdx], l_spsi, l_cpsi, l_sa, l_ca, sin_th, cos_th, sin_i,
                                                cos_i, sin_gamma, cos_gamma, Grav_z, l_mu, l_eta, beta, Lorentz,
                                                l_cxi);
                c2 = cos(2 * chi);
                s2 = sin(2 * chi);
            }
            for (size_t jj = 0; jj < S_cols && !terminate; jj++) {
                if (radiates[i * S_cols + jj] != 1) continue;
                const double shift = phi[i * S_cols + jj], area = cellArea[i * S_cols + jj];
                const double *VEC = cellParams + (i * S_cols + jj) * n_params;
                const double *CV = correction ? correction + (i * S_cols + jj) * 2 : NULL;
                for (size_t k = 0; k < N_P; k++) {
                    double x = phases[k] + shift;
                    if (x > PHASE[N_L - 1]) { while (x > PHASE[N_L - 1]) x -= two_pi; }
                    else if (x < PHASE[0]) { while (x < PHASE[0]) x += two_pi; }
                    if (x < PHASE[0] || x > PHASE[N_L - 1]) {
                        printf("Interpolation error: phase = %.16e\n", x);
                        terminate = 1;
                        break;
                    }
                    const double G = or_akima_eval(PHASE, GE, N_L, gb, gb + N_L, gb + 2 * N_L, x);
                    if (G > 0.0) {
                        const double Zq = or_steffen_eval(PHASE, N_L, zs, zs + N_L, zs + 2 * N_L, zs + 3 * N_L, x);
                        const double Aq = or_steffen_eval(PHASE, N_L, as, as + N_L, as + 2 * N_L, as + 3 * N_L, x);
                        for (size_t p = 0; p < N_E; p++) {
                            const double Ep = energies[p] / Zq;
                            const double IE = (atm_ext == 1 && beam_opt == 0 ? bb_I(Ep, VEC)
                                                                            : or_hot_I(atm_ext, beam_opt, Ep, Aq, VEC)) * BB_NORM;
                            const double cIE = CV ? bb_I(Ep, CV) * BB_NORM : 0.0;
                            acc[p * N_P + k] += area * (IE - cIE) * G;
                            if (pol) {
                                const double QE = or_hot_Q(atm_ext, Ep, Aq, VEC) * BB_NORM;
                                accQ[p * N_P + k] += area * (QE * c2) * G;
                                accU[p * N_P + k] += area * (QE * s2) * G;
                            }
                        }
                    }
                }
            }
            if (terminate) break;
        }
    }
    for (size_t q = 0; q < N_E * N_P; q++) flux[q] = acc[q];
    for (size_t p = 0; p < N_E; p++)
        for (size_t k = 0; k < N_P; k++) flux[p * N_P + k] /= energies[p] * KEV;
    if (pol)
        for (size_t p = 0; p < N_E; p++)
            for (size_t k = 0; k < N_P; k++) {
                flux_Q[p * N_P + k] = accQ[p * N_P + k] / (energies[p] * KEV);
                flux_U[p * N_P + k] = accU[p * N_P + k] / (energies[p] * KEV);
            }
    free(acc); free(accQ); free(accU); free(PHASE); free(ZZ); free(AB); free(GE); free(cdefl); free(calt);
    free(sa); free(sl); free(st); free(gb); free(mb); free(zs); free(as);
    return terminate;
}

int or_integrate_general(double omega, double inclination, size_t S_rows, size_t S_cols,
                         const double *cellArea, const double *radial, const double *r_s_over_r,
                         const double *theta, const double *phi, const double *cellParams,
                         const int *radiates, size_t N_R, const double *deflection,
                         const double *cos_alpha, const double *lag, const double *maxDeflection,
                         const double *cos_gammaArray, size_t N_E, const double *energies, size_t N_L,
                         const double *leaves, size_t N_P, const double *phases, int image_order_limit,
                         const double *correction, double *flux)
{
    return or_integrate_general_ext(omega, inclination, S_rows, S_cols, cellArea, radial, r_s_over_r, theta, phi,
                                    cellParams, 2, radiates, N_R, deflection, cos_alpha, lag, maxDeflection,
                                    cos_gammaArray, N_E, energies, N_L, leaves, N_P, phases, image_order_limit, 1,
                                    0, correction, flux, NULL, NULL);
}
