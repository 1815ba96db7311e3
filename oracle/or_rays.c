/*
 * Oracle: Schwarzschild ray deflection and time-lag tables.
 * TEST INFRASTRUCTURE ONLY.  Restates xpsi/cellmesh/rays.pyx:53-390
 * (integrands :61-88, rayIntegrator :90-247, compute_rays :249-390) with the
 * reference's single-thread semantics: the bad-ray repair counter carries
 * from one parallel to the next, as it does for one OpenMP thread.
 */
#include <math.h>
#include "xpsi_oracle.h"

static const double C_SI = 2.99792458e8;
static int eq12(double x, double y) { return fabs(x - y) < 1.0e-12; }

typedef struct { double a, b; } rp;

/* rays.pyx:61-64 */
static double in_def(double x, void *v)
{
    double rc = ((rp *)v)->a;
    return 1.0 / sqrt(2.0 - x * x - (1.0 - x * x) * (1.0 - x * x) / (rc - 1.0));
}

/* rays.pyx:66-72 */
static double in_lag(double x, void *v)
{
    double rc = ((rp *)v)->a;
    double X = 1.0 / sqrt(2.0 - x * x - (1.0 - x * x) * (1.0 - x * x) / (rc - 1.0));
    return X * X / (x + X);
}

/* rays.pyx:74-79 */
static double out_def(double x, void *v)
{
    double sas = ((rp *)v)->a, Rr = ((rp *)v)->b;
    return x / sqrt(1.0 - sas + x * x * sas * (2.0 - x * x - (1.0 - x * x) * (1.0 - x * x) / (Rr - 1.0)));
}

/* rays.pyx:81-88 */
static double out_lag(double x, void *v)
{
    double sas = ((rp *)v)->a, Rr = ((rp *)v)->b;
    double f = sqrt(1.0 - sas + x * x * sas * (2.0 - x * x - (1.0 - x * x) * (1.0 - x * x) / (Rr - 1.0)));
    return x / (f + f * f);
}

static int quad(or_fn f, rp *p, double a, double b, double *r)
{
    double e;
    return or_qag61(f, p, a, b, 0.0, 1.0e-12, 100, r, &e);
}

/* rays.pyx:90-247: out[0] = deflection, out[1] = lag (seconds) */
static void ray(double cos_alpha, double r_s, double u, double out[2])
{
    const double b_ph = 3.0 * sqrt(3.0) / 2.0;
    double sas = 1.0 - cos_alpha * cos_alpha;
    double sa = sqrt(sas);
    double alpha = acos(cos_alpha);
    double b = sa / (u * sqrt(1.0 - u));
    double d0 = NAN, l0 = NAN;
    int st = OR_SUCCESS;
    if (b <= b_ph) {
        rp p = {sas, 1.0 / u};
        st = quad(out_def, &p, 0.0, 1.0, &d0);
        if (st == OR_EFAILED) d0 = NAN;
        if (st != OR_EFAILED) {
            st = quad(out_lag, &p, 0.0, 1.0, &l0);
            if (st == OR_EFAILED) d0 = NAN;
        }
        if (st != OR_EFAILED) {
            d0 *= 2.0 * b * u;
            l0 *= 2.0 * b * b * u * r_s;
        }
    } else if (alpha <= M_PI / 2.0) {
        double rc, wR;
        if (eq12(sa, 1.0)) { rc = 1.0 / u; wR = 0.0; }
        else {
            rc = 2.0 * b * cos((atan(sqrt(4.0 * b * b / 27.0 - 1.0)) - M_PI) / 3.0) / sqrt(3.0);
            wR = sqrt(1.0 - rc * u);
        }
        rp p = {rc, 0.0};
        if (isnan(wR)) wR = 0.0;
        st = quad(in_def, &p, wR, 1.0, &d0);
        if (st == OR_EFAILED) d0 = NAN;
        if (st != OR_EFAILED) {
            st = quad(in_lag, &p, wR, 1.0, &l0);
            if (st == OR_EFAILED) d0 = NAN;
        }
        if (st != OR_EFAILED) {
            d0 *= 2.0 * b / rc;
            l0 *= 2.0 * b * b * r_s / rc;
        }
    } else {
        double rc = 2.0 * b * cos((atan(sqrt(4.0 * b * b / 27.0 - 1.0)) - M_PI) / 3.0) / sqrt(3.0);
        rp p = {rc, 0.0};
        double wR = sqrt(1.0 - rc * u);
        if (isnan(wR)) wR = 0.0;
        double D1, D0, L1, L0;
        st = quad(in_def, &p, 0.0, 1.0, &D1);
        if (st == OR_EFAILED) d0 = NAN;
        if (st != OR_EFAILED) { st = quad(in_def, &p, wR, 1.0, &D0); if (st == OR_EFAILED) d0 = NAN; }
        if (st != OR_EFAILED) { st = quad(in_lag, &p, 0.0, 1.0, &L1); if (st == OR_EFAILED) d0 = NAN; }
        if (st != OR_EFAILED) { st = quad(in_lag, &p, wR, 1.0, &L0); if (st == OR_EFAILED) d0 = NAN; }
        if (st != OR_EFAILED) {
            d0 = 2.0 * b * (2.0 * D1 - D0) / rc;
            l0 = 2.0 * b * b * r_s * (2.0 * L1 - L0) / rc;
            double cor = 1.0 / u - rc + log((1.0 / u - 1.0) / (rc - 1.0));
            l0 += 2.0 * r_s * cor;
        }
    }
    if (st != OR_EFAILED) l0 /= C_SI;
    out[0] = d0;
    out[1] = l0;
}

/* rays.pyx:249-390 */
int or_compute_rays(size_t N_u, double r_s, const double *u_arr,
                    double *maxAlpha, size_t N_R, double *defl,
                    double *cos_a, double *lag, double *maxDefl)
{
    const double b_ph = 3.0 * sqrt(3.0) / 2.0, eps = 1.0e-8;
    int terminate = 0;
    size_t counter = 0, idx = 0;
    for (size_t i = 0; i < N_u; i++) {
        double u = u_arr[i], extreme = 0.0;
        if (u < 2.0 / 3.0) extreme = M_PI - asin(sqrt(1.0 - u) * b_ph * u);
        else if (eq12(u, 2.0 / 3.0)) extreme = M_PI / 2.0;
        else if (2.0 / 3.0 < u && u < 1.0) extreme = asin(sqrt(1.0 - u) * b_ph * u);
        if (maxAlpha[i] >= (1.0 - eps) * extreme) maxAlpha[i] = (1.0 - eps) * extreme;
        double inc = (1.0 - cos(maxAlpha[i])) / ((double)N_R - 1.0);
        double *D = defl + i * N_R, *A = cos_a + i * N_R, *L = lag + i * N_R;
        for (size_t j = 0; j < N_R; j++) {
            A[j] = 1.0 - (double)j * inc;
            if (j == 0) { A[0] = 1.0; D[0] = 0.0; L[0] = 0.0; continue; }
            if (terminate) continue;
            double o[2];
            ray(A[j], r_s, u, o);
            if (j == N_R - 1 && (isnan(o[0]) || o[0] <= D[j - 1])) {
                terminate = 1;
            } else if (counter > 0 && !isnan(o[0]) && 0.0 <= o[0] &&
                       o[0] < 100.0 * M_PI && o[0] > D[idx]) {
                D[j] = o[0]; L[j] = o[1];
                double delta = (double)counter + 1.0;
                for (size_t k = 1; k <= counter; k++) {
                    D[idx + k] = D[j] * (double)k;
                    D[idx + k] += D[idx] * (delta - (double)k);
                    D[idx + k] /= delta;
                    L[idx + k] = L[j] * (double)k;
                    L[idx + k] += L[idx] * (delta - (double)k);
                    L[idx + k] /= delta;
                }
                counter = 0;
            } else if (isnan(o[0]) || o[0] < 0.0 || o[0] > 100.0 * M_PI) {
                if (counter == 0) idx = j - 1;
                counter++;
            } else {
                D[j] = o[0]; L[j] = o[1];
            }
        }
        maxDefl[i] = D[N_R - 1];
    }
    return terminate;
}
