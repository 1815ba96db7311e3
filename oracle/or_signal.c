/*
 * Oracle: energy-bin integration, phase-bin expected counts and the
 * background-marginalised Poisson likelihood.  TEST INFRASTRUCTURE ONLY.
 * Restates xpsi/tools/energy_integrator.pyx:27-114,
 * xpsi/tools/compute_expected_counts.pyx:66-197 and
 * xpsi/likelihoods/default_background_marginalisation.pyx:38-734.
 * The reference's gsl_integration_cquad over the background is replaced by
 * QAG-GK61 at min(epsrel, OR_MARG_EPSREL); gsl_sf_lnfact(n) by lgamma(n+1).
 * The reference draws llzero-class outputs from rand()/numpy; here they come
 * from a fixed linear congruential sequence (compared as a class in tests).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "xpsi_oracle.h"

static int eq12(double x, double y) { return fabs(x - y) < 1.0e-12; }
static unsigned long long lcg_state = 12345ULL;
static double unit_draw(void)
{
    lcg_state = lcg_state * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(lcg_state >> 11) / 9007199254740992.0;
}

/* energy_integrator.pyx:27-114 (the phase interpolant, periodic Akima by default) */
void or_energy_integrator(size_t N_E, size_t N_P, const double *signal,
                          const double *logE, size_t N_edges,
                          const double *logEdges, double *out)
{
    size_t N_in = N_edges - 1;
    double maxE = logE[N_E - 1];
    double *y = (double *)malloc(N_E * sizeof(double));
    double *c = (double *)malloc(3 * N_E * sizeof(double));
    double *m = (double *)malloc((3 * N_E + 3) * sizeof(double));
    memset(out, 0, N_in * N_P * sizeof(double));
    for (size_t i = 0; i < N_P; i++) {
        for (size_t j = 0; j < N_E; j++)
            y[j] = pow(10.0, logE[j]) * signal[j * N_P + i] * log(10.0);
        or_phase_init(logE, y, N_E, c, c + N_E, c + 2 * N_E, m);
        for (size_t j = 0; j < N_in; j++) {
            double up = logEdges[j + 1] > maxE ? maxE : logEdges[j + 1];
            out[j * N_P + i] = or_akima_integ(logE, y, N_E, c, c + N_E, c + 2 * N_E,
                                              logEdges[j], up);
            if (logEdges[j + 1] > maxE) break;
        }
    }
    free(y); free(c); free(m);
}

/* default_background_marginalisation.pyx:38-68 */
void or_precomputation(size_t N_ch, size_t N_b, const int *counts, double *pre)
{
    for (size_t i = 0; i < N_ch; i++) {
        double s = 0.0;
        for (size_t j = 0; j < N_b; j++) s += lgamma((double)(unsigned int)counts[i * N_b + j] + 1.0);
        pre[i] = -s;
    }
}

typedef struct {
    size_t n;
    double SCALE, T_exp, std, A;
    const double *data;
    double *star;
} margs;

/* default_background_marginalisation.pyx:84-138 */
static double marg_integrand(double B, void *v)
{
    margs *a = (margs *)v;
    double x = 0.0;
    for (size_t j = 0; j < a->n; j++) {
        double c = a->SCALE * (a->star[j] + B);
        if (c > 0.0) x += a->data[j] * log(c) - c;
        else if (eq12(c, 0.0) && eq12(a->data[j], 0.0)) { }
        else return 0.0;
    }
    return exp(x - a->A);
}

/* default_background_marginalisation.pyx:140-173 */
static double newton_delta(double B, margs *a)
{
    double x = 0.0, y = 0.0;
    for (size_t j = 0; j < a->n; j++) {
        y += a->data[j] / (a->star[j] + B);
        x += a->data[j] / pow(a->star[j] + B, 2.0);
    }
    y = 2.0 * a->T_exp - 2.0 * y;
    x *= 2.0;
    a->std = sqrt(2.0 / x);
    return -1.0 * y / x;
}

/* default_background_marginalisation.pyx:176-447; returns 0 on failure */
static int channel_marginal(margs *a, double precomp, double *LL,
                            double av_DATA, double av_STAR, double epsabs,
                            double epsrel, double epsilon, double sigmas,
                            double llzero, const double *support,
                            double *B_ML, double *B_ML_supp)
{
    double T = a->T_exp, B, lower, upper;
    if (eq12(av_DATA, 0.0) && eq12(av_STAR, 0.0)) {
        lower = 0.0;
        if (lower < support[0]) lower = support[0];
        upper = 10.0 / T;
        if (upper > support[1] && support[1] > 0.0) upper = support[1];
        B = 0.0;
        *LL += log((exp(-1.0 * lower * T) - exp(-1.0 * upper * T)) / T);
    } else {
        double B_min = 0.0;
        B = av_DATA - av_STAR;
        if (B <= B_min) {
            double mc = -1.0;
            for (size_t j = 0; j < a->n; j++)
                if (eq12(a->star[j], 0.0)) { mc = -2.0; break; }
            if (eq12(mc, -2.0)) {
                for (size_t j = 0; j < a->n; j++)
                    if ((eq12(mc, -2.0) && a->data[j] > 0.0) ||
                        (0.0 < a->data[j] && a->data[j] < mc))
                        mc = a->data[j];
            }
            if (!eq12(mc, -1.0)) {
                if (eq12(mc, -2.0)) { B = 0.0; B_min = 0.0; }
                else { B = 0.01 * mc / a->SCALE; B_min = 0.1 * B; }
            } else B = B_min;
        }
        double dB = newton_delta(B, a);
        int counter = 0;
        while (fabs(dB) > epsilon * a->std && counter < 2) {
            B += dB;
            if (B < B_min) {
                if (B_min > 0.0) counter += 1; else counter = 2;
                B = B_min;
            }
            dB = newton_delta(B, a);
        }
        double s = 0.0;
        for (size_t j = 0; j < a->n; j++) s += a->data[j] / pow(a->star[j] + B, 2.0);
        s = (s > 0.0) ? sqrt(1.0 / s) : 1e90;
        lower = B - sigmas * s;
        upper = B + sigmas * s;
        if (lower < B_min) lower = B_min;
        double Bi = B;
        if (lower < support[0]) {
            lower = support[0];
            if (upper < support[0] && support[1] > 0.0) { upper = support[1]; Bi = support[0]; }
            else if (upper < support[0]) { upper = support[0] + sigmas * s; Bi = support[0]; }
        }
        if (upper > support[1] && support[1] > 0.0) {
            upper = support[1];
            if (lower > support[1]) { lower = support[0]; Bi = support[1]; }
        }
        if (Bi < lower) Bi = lower;
        else if (Bi > upper) Bi = upper;
        a->A = 0.0;
        for (size_t j = 0; j < a->n; j++) {
            double c = a->SCALE * (a->star[j] + Bi);
            if (c > 0.0) a->A += a->data[j] * log(c) - c;
            else if (eq12(c, 0.0) && eq12(a->data[j], 0.0)) { }
            else a->A += llzero;
        }
        double result, abserr;
        double er = epsrel < OR_MARG_EPSREL ? epsrel : OR_MARG_EPSREL;
        or_qag61(marg_integrand, a, lower, upper, epsabs, er, 100, &result, &abserr);
        if (result > 0.0) {
            *LL += log(result) + a->A + precomp;
        } else {
            *LL = llzero * (0.1 + 0.9 * unit_draw());
            return 0;
        }
    }
    *B_ML = B * T;
    if (B < support[0]) B = support[0];
    else if (B > support[1] && support[1] > 0.0) B = support[1];
    *B_ML_supp = B * T;
    for (size_t j = 0; j < a->n; j++) a->star[j] = a->SCALE * (a->star[j] + B);
    return 1;
}

/* compute_expected_counts.pyx:66-197: channel i's expected count rate in
   each phase bin (shifted, wrapped, negative bin integrals dropped unless
   allowed, then clipped at zero), added into st[N_b] (zeroed by the caller);
   cb [3 (max N_P + 1)] and mb [3 max N_P + 4] are scratch */
static void channel_rate(size_t i, size_t N_b, const double *phases, size_t n_comp,
                         const double *const *comps, const size_t *comp_N_P,
                         const double *const *comp_phases, const double *shifts,
                         const unsigned char *allow_neg, double *cb, double *mb,
                         double *st)
{
    for (size_t c = 0; c < n_comp; c++) {
        size_t NP = comp_N_P[c];
        const double *pulse = comps[c] + i * NP;
        if (NP > 1) {
            const double *ph = comp_phases[c];
            or_phase_init(ph, pulse, NP, cb, cb + NP, cb + 2 * NP, mb);
            for (size_t j = 0; j < N_b; j++) {
                double pa = phases[j] + shifts[c], pb = phases[j + 1] + shifts[c];
                if (eq12(pb - pa, 1.0)) { pa = 0.0; pb = 1.0; }
                else { pa -= floor(pa); pb -= floor(pb); }
                double v;
                if (pa < pb) {
                    v = or_akima_integ(ph, pulse, NP, cb, cb + NP, cb + 2 * NP, pa, pb);
                    if (v > 0.0 || allow_neg[c]) st[j] += v;
                } else {
                    v = or_akima_integ(ph, pulse, NP, cb, cb + NP, cb + 2 * NP, pa, 1.0);
                    if (v > 0.0 || allow_neg[c]) st[j] += v;
                    v = or_akima_integ(ph, pulse, NP, cb, cb + NP, cb + 2 * NP, 0.0, pb);
                    if (v > 0.0 || allow_neg[c]) st[j] += v;
                }
            }
        } else {
            st[0] = pulse[0];
        }
    }
    for (size_t j = 0; j < N_b; j++) if (st[j] < 0.0) st[j] = 0.0;
}

int or_eval_marginal_likelihood(double T_exp, size_t N_b, const double *phases,
                                size_t N_ch, const double *counts,
                                size_t n_comp, const double *const *comps,
                                const size_t *comp_N_P,
                                const double *const *comp_phases,
                                const double *shifts, const double *precomp,
                                const double *support, double epsabs,
                                double epsrel, double epsilon, double sigmas,
                                double llzero, const unsigned char *allow_neg,
                                double slim, const double *background,
                                double *lnl, double *STAR, double *B_ML,
                                double *B_ML_supp)
{
    double n = (double)N_b, SCALE = T_exp / n, LL = 0.0;
    size_t maxP = 0;
    for (size_t c = 0; c < n_comp; c++) if (comp_N_P[c] > maxP) maxP = comp_N_P[c];
    double *cb = (double *)malloc(3 * (maxP + 1) * sizeof(double));
    double *mb = (double *)malloc((3 * maxP + 4) * sizeof(double));
    memset(STAR, 0, N_ch * N_b * sizeof(double));
    memset(B_ML, 0, N_ch * sizeof(double));
    memset(B_ML_supp, 0, N_ch * sizeof(double));
    margs a;
    a.n = N_b; a.SCALE = SCALE; a.T_exp = T_exp;
    for (size_t i = 0; i < N_ch; i++) {
        double *st = STAR + i * N_b;
        channel_rate(i, N_b, phases, n_comp, comps, comp_N_P, comp_phases, shifts, allow_neg, cb, mb, st);
        double av_D = 0.0, av_S = 0.0;
        for (size_t j = 0; j < N_b; j++) {
            if (background) st[j] += background[i * N_b + j];
            st[j] *= n;
            av_S += st[j];
            av_D += counts[i * N_b + j];
        }
        if (!(slim < 0.0)) {
            double limit = av_S * SCALE - slim * sqrt(av_S * SCALE) - av_D;
            if (limit > 0.0) { LL = llzero * (0.1 + 0.9 * unit_draw()); break; }
        }
        av_S /= n;
        av_D /= T_exp;
        a.data = counts + i * N_b;
        a.star = st;
        if (!channel_marginal(&a, precomp[i], &LL, av_D, av_S, epsabs, epsrel,
                              epsilon, sigmas, llzero, support + 2 * i,
                              B_ML + i, B_ML_supp + i))
            break;
    }
    free(cb); free(mb);
    *lnl = LL;
    return 0;
}

/* xpsi/likelihoods/_poisson_likelihood_given_background.pyx:14-113 with
   compute_expected_counts.pyx:199-315: EXPEC = (star rate + background rate)
   x exposure per channel and bin; lnL = sum over channels of
   neg_sum_ln_data_factorial (when given) plus, per bin, d log EXPEC - EXPEC;
   a bin with no expected counts but data adds -1e90 (0.1 + 0.9 U). */
int or_poisson_likelihood_given_background(double T_exp, size_t N_b, const double *phases,
                                           size_t N_ch, const double *counts, size_t n_comp,
                                           const double *const *comps, const size_t *comp_N_P,
                                           const double *const *comp_phases,
                                           const double *shifts, const double *background,
                                           const double *precomp,
                                           const unsigned char *allow_neg,
                                           double *lnl, double *EXPEC)
{
    size_t maxP = 0;
    for (size_t c = 0; c < n_comp; c++) if (comp_N_P[c] > maxP) maxP = comp_N_P[c];
    double *cb = (double *)malloc(3 * (maxP + 1) * sizeof(double));
    double *mb = (double *)malloc((3 * maxP + 4) * sizeof(double));
    memset(EXPEC, 0, N_ch * N_b * sizeof(double));
    double LL = 0.0;
    for (size_t i = 0; i < N_ch; i++) {
        double *ex = EXPEC + i * N_b;
        channel_rate(i, N_b, phases, n_comp, comps, comp_N_P, comp_phases, shifts, allow_neg, cb, mb, ex);
        for (size_t j = 0; j < N_b; j++) {
            ex[j] += background[i * N_b + j];
            ex[j] *= T_exp;
        }
    }
    for (size_t i = 0; i < N_ch; i++) {
        if (precomp) LL += precomp[i];
        for (size_t j = 0; j < N_b; j++) {
            double e = EXPEC[i * N_b + j], d = counts[i * N_b + j];
            if (e > 0.0) {
                LL -= e;
                LL += d * log(e);
            } else if (d == 0.0 && e == 0.0) {
            } else {
                LL += -1.0e90 * (0.1 + 0.9 * unit_draw());
            }
        }
    }
    free(cb); free(mb);
    *lnl = LL;
    return 0;
}

/* xpsi/likelihoods/_gaussian_likelihood_given_background_IQU.pyx:14-91 with
   compute_expected_counts.pyx:199-315: EXPEC as for the Poisson statistic;
   with s2 = errors^2 per channel and bin, lnL is the running
   LOGLIKE -= (EXPEC - d)^2 / (2 s2) + log(2 pi s2) / 2, channel by channel,
   bin by bin. */
int or_gaussian_likelihood_given_background(double T_exp, size_t N_b, const double *phases,
                                            size_t N_ch, const double *counts, const double *errors,
                                            size_t n_comp, const double *const *comps,
                                            const size_t *comp_N_P, const double *const *comp_phases,
                                            const double *shifts, const double *background,
                                            const unsigned char *allow_neg, double *lnl, double *EXPEC)
{
    size_t maxP = 0;
    for (size_t c = 0; c < n_comp; c++) if (comp_N_P[c] > maxP) maxP = comp_N_P[c];
    double *cb = (double *)malloc(3 * (maxP + 1) * sizeof(double));
    double *mb = (double *)malloc((3 * maxP + 4) * sizeof(double));
    memset(EXPEC, 0, N_ch * N_b * sizeof(double));
    for (size_t i = 0; i < N_ch; i++) {
        double *ex = EXPEC + i * N_b;
        channel_rate(i, N_b, phases, n_comp, comps, comp_N_P, comp_phases, shifts, allow_neg, cb, mb, ex);
        for (size_t j = 0; j < N_b; j++) {
            ex[j] += background[i * N_b + j];
            ex[j] *= T_exp;
        }
    }
    double LL = 0.0;
    for (size_t i = 0; i < N_ch; i++)
        for (size_t j = 0; j < N_b; j++) {
            const double s2 = pow(errors[i * N_b + j], 2.0);
            const double norm = 0.5 * log(2.0 * M_PI * s2);
            const double r = EXPEC[i * N_b + j] - counts[i * N_b + j];
            LL -= r * r / (2.0 * s2) + norm;
        }
    free(cb); free(mb);
    *lnl = LL;
    return 0;
}

/* xpsi/likelihoods/_gaussian_likelihood_QnUn.pyx:14-44: counts and errors
   [N_rows, N_b], star [N_b] (the same model row for every data row);
   LOGLIKE -= (star_j - d_ij)^2 / (2 s2_ij) + log(2 pi s2_ij) / 2. */
double or_gaussian_likelihood_QnUn(size_t N_rows, size_t N_b, const double *counts,
                                   const double *errors, const double *star)
{
    double LL = 0.0;
    for (size_t i = 0; i < N_rows; i++)
        for (size_t j = 0; j < N_b; j++) {
            const double s2 = pow(errors[i * N_b + j], 2.0);
            const double norm = 0.5 * log(2.0 * M_PI * s2);
            const double r = star[j] - counts[i * N_b + j];
            LL -= r * r / (2.0 * s2) + norm;
        }
    return LL;
}

/* xpsi/tools/phase_integrator.pyx:23-121: each row's phase interpolant
   integrated over the shifted intervals (the wrap exact at b - a == 1, not
   within 1e-12 as the expected counts take it), a piece kept where positive or
   allowed, the sum not clipped, times the exposure; out [N_rows, N_b]. */
int or_phase_integrator(double T_exp, size_t N_b, const double *phases, size_t N_rows,
                        const double *signal, size_t NP, const double *signal_phases,
                        double shift, int allow_neg, double *out)
{
    if (NP < 2) return 1;
    double *cb = (double *)malloc(3 * (NP + 1) * sizeof(double));
    double *mb = (double *)malloc((3 * NP + 4) * sizeof(double));
    for (size_t i = 0; i < N_rows; i++) {
        const double *pulse = signal + i * NP;
        or_phase_init(signal_phases, pulse, NP, cb, cb + NP, cb + 2 * NP, mb);
        for (size_t j = 0; j < N_b; j++) {
            double a = phases[j] + shift, b = phases[j + 1] + shift, v = 0.0, w;
            if (b - a == 1.0) { a = 0.0; b = 1.0; }
            else { a -= floor(a); b -= floor(b); }
            if (a < b) {
                w = or_akima_integ(signal_phases, pulse, NP, cb, cb + NP, cb + 2 * NP, a, b);
                if (w > 0.0 || allow_neg) v = w;
            } else {
                w = or_akima_integ(signal_phases, pulse, NP, cb, cb + NP, cb + 2 * NP, a, 1.0);
                if (w > 0.0 || allow_neg) v = w;
                w = or_akima_integ(signal_phases, pulse, NP, cb, cb + NP, cb + 2 * NP, 0.0, b);
                if (w > 0.0 || allow_neg) v += w;
            }
            out[i * N_b + j] = v * T_exp;
        }
    }
    free(cb); free(mb);
    return 0;
}
