/*
 * Oracle numerics core: CPU restatement of the GNU Scientific Library subset
 * that the X-PSI pulse-profile likelihood path calls.  TEST INFRASTRUCTURE
 * ONLY: the product library never links or calls anything in oracle/.
 *
 * GSL is a third-party dependency of the reference (xpsi/include/GSL.pxd,
 * version unpinned: conda `gsl`, apt `libgsl-dev`) and is absent from the
 * reference tree, so the published algorithms are restated here:
 *   - interval search: largest i in [lo, hi-1] with x_i <= x (GSL bsearch
 *     convention; x == x_max maps to the last interval);
 *   - Steffen (1990) monotone cubic, one-sided end slopes (GSL steffen);
 *   - Akima (1970) cubic with GSL's degenerate-slope conventions, in the
 *     natural and periodic end-slope variants, plus its exact integral;
 *   - QUADPACK QAG with the 61-point Gauss-Kronrod rule (Piessens et al.
 *     1983), bisecting the interval of largest error estimate.
 * Call sites in the reference: integrator_for_azimuthal_invariance.pyx:194-202,
 * 307-316,372-414,569,592; rays.pyx:126-232; mesh.pyx:48-49,115-120;
 * energy_integrator.pyx:73,89,96; compute_expected_counts.pyx:143-187.
 */
#include <math.h>
#include <float.h>
#include <stddef.h>
#include "xpsi_oracle.h"

size_t or_interval(const double *xa, size_t n, double x)
{
    size_t lo = 0, hi = n - 1;
    while (hi > lo + 1) {
        size_t mid = (hi + lo) / 2;
        if (xa[mid] > x) hi = mid; else lo = mid;
    }
    return lo;
}

static double sgn1(double s)
{
    /* +1 for s >= 0 (zero counts as positive), -1 otherwise */
    return (s < 0.0) ? -1.0 : 1.0;
}

/* Steffen: coefficients per interval, y = d + t(c + t(b + t a)), t = x - x_i */
void or_steffen_init(const double *x, const double *y, size_t n,
                     double *a, double *b, double *c, double *d)
{
    double *yp = d; /* reuse d as slope storage until the final pass */
    double h0 = x[1] - x[0];
    yp[0] = (y[1] - y[0]) / h0;
    for (size_t i = 1; i + 1 < n; i++) {
        double hi = x[i + 1] - x[i], hm = x[i] - x[i - 1];
        double si = (y[i + 1] - y[i]) / hi, sm = (y[i] - y[i - 1]) / hm;
        double p = (sm * hi + si * hm) / (hm + hi);
        double m = fmin(fabs(sm), fmin(fabs(si), 0.5 * fabs(p)));
        yp[i] = (sgn1(sm) + sgn1(si)) * m;
    }
    double ylast = (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
    /* slopes at i and i+1 are needed, so walk forward keeping the next one */
    for (size_t i = 0; i + 1 < n; i++) {
        double hi = x[i + 1] - x[i];
        double si = (y[i + 1] - y[i]) / hi;
        double y0 = yp[i];
        double y1 = (i + 2 < n) ? yp[i + 1] : ylast;
        a[i] = (y0 + y1 - 2.0 * si) / hi / hi;
        b[i] = (3.0 * si - 2.0 * y0 - y1) / hi;
        c[i] = y0;
    }
    for (size_t i = 0; i + 1 < n; i++) d[i] = y[i];
}

double or_steffen_eval(const double *x, size_t n, const double *a,
                       const double *b, const double *c, const double *d,
                       double xq)
{
    size_t i = or_interval(x, n, xq);
    double t = xq - x[i];
    return d[i] + t * (c[i] + t * (b[i] + t * a[i]));
}

double or_steffen_deriv(const double *x, size_t n, const double *a,
                        const double *b, const double *c, double xq)
{
    size_t i = or_interval(x, n, xq);
    double t = xq - x[i];
    return c[i] + t * (2.0 * b[i] + t * 3.0 * a[i]);
}

/* Akima: y = y_i + t(b + t(c + t d)).  m must hold n + 3 doubles. */
void or_akima_init(const double *x, const double *y, size_t n, int periodic,
                   double *b, double *c, double *d, double *mbuf)
{
    double *m = mbuf + 2; /* m[-2], m[-1] addressable */
    for (size_t i = 0; i + 1 < n; i++)
        m[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
    if (periodic) {
        m[-2] = m[n - 3];
        m[-1] = m[n - 2];
        m[n - 1] = m[0];
        m[n] = m[1];
    } else {
        m[-2] = 3.0 * m[0] - 2.0 * m[1];
        m[-1] = 2.0 * m[0] - m[1];
        m[n - 1] = 2.0 * m[n - 2] - m[n - 3];
        m[n] = 3.0 * m[n - 2] - 2.0 * m[n - 3];
    }
    for (long i = 0; i + 1 < (long)n; i++) {
        double w_i = fabs(m[i + 1] - m[i]) + fabs(m[i - 1] - m[i - 2]);
        if (w_i == 0.0) {
            b[i] = m[i];
            c[i] = 0.0;
            d[i] = 0.0;
            continue;
        }
        double h = x[i + 1] - x[i];
        double w_n = fabs(m[i + 2] - m[i + 1]) + fabs(m[i] - m[i - 1]);
        double al = fabs(m[i - 1] - m[i - 2]) / w_i;
        double t_next;
        if (w_n == 0.0) {
            t_next = m[i];
        } else {
            double al_n = fabs(m[i] - m[i - 1]) / w_n;
            t_next = (1.0 - al_n) * m[i] + al_n * m[i + 1];
        }
        b[i] = (1.0 - al) * m[i - 1] + al * m[i];
        c[i] = (3.0 * m[i] - 2.0 * b[i] - t_next) / h;
        d[i] = (b[i] + t_next - 2.0 * m[i]) / (h * h);
    }
}

double or_akima_eval(const double *x, const double *y, size_t n,
                     const double *b, const double *c, const double *d,
                     double xq)
{
    size_t i = or_interval(x, n, xq);
    double t = xq - x[i];
    return y[i] + t * (b[i] + t * (c[i] + t * d[i]));
}

/* exact integral of one Akima piece between lo and hi (offsets from x_i) */
static double piece_integral(double yi, double bi, double ci, double di,
                             double xi, double lo, double hi)
{
    double r1 = lo - xi, r2 = hi - xi;
    double r12 = r1 + r2;
    double bt = 0.5 * bi * r12;
    double ct = (1.0 / 3.0) * ci * (r1 * r1 + r2 * r2 + r1 * r2);
    double dt = 0.25 * di * r12 * (r1 * r1 + r2 * r2);
    return (hi - lo) * (yi + bt + ct + dt);
}

double or_akima_integ(const double *x, const double *y, size_t n,
                      const double *b, const double *c, const double *d,
                      double lo, double hi)
{
    if (lo > hi || lo < x[0] || hi > x[n - 1]) return NAN;
    if (lo == hi) return 0.0;
    size_t ia = or_interval(x, n, lo), ib = or_interval(x, n, hi);
    double s = 0.0;
    for (size_t i = ia; i <= ib; i++) {
        double xl = x[i], xh = x[i + 1], h = xh - xl;
        if (h == 0.0) return 0.0;
        if (i == ia || i == ib) {
            double a1 = (i == ia) ? lo : xl;
            double a2 = (i == ib) ? hi : xh;
            s += piece_integral(y[i], b[i], c[i], d[i], xl, a1, a2);
        } else {
            s += h * (y[i] + h * (0.5 * b[i] + h * (c[i] / 3.0 + 0.25 * d[i] * h)));
        }
    }
    return s;
}

/* Periodic cubic spline (the unique C2 periodic interpolant; y[n-1] closes
 * the period): c_i = y''(x_i) / 2 from the cyclic tridiagonal system
 *   h_{i-1} c_{i-1} + 2 (h_{i-1} + h_i) c_i + h_i c_{i+1}
 *     = 3 ((y_{i+1} - y_i) / h_i - (y_i - y_{i-1}) / h_{i-1}),  i = 0 .. m-1,
 * m = n - 1 unknowns with indices taken modulo m, solved by elimination with
 * the Sherman-Morrison correction for the two corner entries; then per
 * interval b = s_i - h_i (c_{i+1} + 2 c_i) / 3, d = (c_{i+1} - c_i) / (3 h_i),
 * y = y_i + t (b + t (c + t d)).  work: 3 n doubles.  Needs n >= 4. */
void or_cspline_periodic_init(const double *x, const double *y, size_t n, double *b, double *c, double *d,
                              double *work)
{
    const size_t m = n - 1;
    double *dg = work, *rhs = work + n, *z = work + 2 * n;
    /* diagonal, right-hand side; off-diagonal e_i = h_i couples i and i + 1 */
    for (size_t i = 0; i < m; i++) {
        const size_t im = (i + m - 1) % m;
        const double hm = x[im + 1] - x[im], hi = x[i + 1] - x[i];
        dg[i] = 2.0 * (hm + hi);
        rhs[i] = 3.0 * ((y[i + 1] - y[i]) / hi - (y[im + 1] - y[im]) / hm);
    }
    /* Sherman-Morrison: A = T + u v^T with u = (g, 0.., e_last), v = (1, 0.., e_last / g),
       where T has diag[0] - g and diag[m-1] - e_last^2 / g */
    const double e_last = x[m] - x[m - 1];   /* couples m-1 and 0 */
    const double g = -dg[0];
    double *T = b;                            /* modified diagonal */
    for (size_t i = 0; i < m; i++) T[i] = dg[i];
    T[0] -= g;
    T[m - 1] -= e_last * e_last / g;
    /* solve T yv = rhs and T zv = u by the Thomas algorithm (symmetric, e_i = h_i) */
    double *cp = d;                           /* forward-eliminated super-diagonal */
    double *yv = c;
    for (size_t i = 0; i < m; i++) z[i] = 0.0;
    z[0] = g;
    z[m - 1] = e_last;
    {
        double den = T[0];
        cp[0] = (x[1] - x[0]) / den;
        yv[0] = rhs[0] / den;
        z[0] = z[0] / den;
        for (size_t i = 1; i < m; i++) {
            const double em = x[i] - x[i - 1];   /* e_{i-1} */
            den = T[i] - em * cp[i - 1];
            cp[i] = (i + 1 < m) ? (x[i + 1] - x[i]) / den : 0.0;
            yv[i] = (rhs[i] - em * yv[i - 1]) / den;
            z[i] = (z[i] - em * z[i - 1]) / den;
        }
        for (size_t i = m - 1; i-- > 0;) {
            yv[i] -= cp[i] * yv[i + 1];
            z[i] -= cp[i] * z[i + 1];
        }
    }
    const double fac = (yv[0] + yv[m - 1] * e_last / g) / (1.0 + z[0] + z[m - 1] * e_last / g);
    for (size_t i = 0; i < m; i++) c[i] = yv[i] - fac * z[i];
    c[m] = c[0];
    for (size_t i = 0; i < m; i++) {
        const double h = x[i + 1] - x[i];
        b[i] = (y[i + 1] - y[i]) / h - h * (c[i + 1] + 2.0 * c[i]) / 3.0;
        d[i] = (c[i + 1] - c[i]) / (3.0 * h);
    }
}

/* The phase interpolant, a process-wide choice as in the reference
 * (xpsi/tools/core.pyx:21-116): 0 periodic Akima (default), 1 Steffen,
 * 2 periodic cubic spline.  All three give y = y_i + t (b + t (c + t d)),
 * so or_akima_eval and or_akima_integ evaluate and integrate any of them. */
static int g_phase_kind = 0;

int or_set_phase_interpolant(int kind)
{
    if (kind < 0 || kind > 2) return -1;
    g_phase_kind = kind;
    return 0;
}

int or_get_phase_interpolant(void) { return g_phase_kind; }

/* work: 3 n + 3 doubles */
void or_phase_init(const double *x, const double *y, size_t n, double *b, double *c, double *d, double *work)
{
    if (g_phase_kind == 1) {
        /* Steffen's a, b, c (cubic, quadratic, slope) as d, c, b of the shared form */
        double *dd = work;
        or_steffen_init(x, y, n, d, c, b, dd);
        return;
    }
    if (g_phase_kind == 2) {
        or_cspline_periodic_init(x, y, n, b, c, d, work);
        return;
    }
    or_akima_init(x, y, n, 1, b, c, d, work);
}

/* ---------------- QAG with the 61-point Gauss-Kronrod rule ---------------- */
static const double XGK61[31] = {
  0.999484410050490637571325895705810819,
  0.996893484074649540271630050918695283,
  0.991630996870404594858628366109485725,
  0.983668123279747209970032581605662802,
  0.973116322501126268374693868423706885,
  0.960021864968307512216871025581797663,
  0.944374444748559979415831324037439122,
  0.926200047429274325879324277080474004,
  0.90557330769990779854652255892595832,
  0.88256053579205268154311646253022559,
  0.857205233546061098958658510658943857,
  0.829565762382768397442898119732501916,
  0.799727835821839083013668942322683241,
  0.767777432104826194917977340974503132,
  0.733790062453226804726171131369527646,
  0.697850494793315796932292388026640068,
  0.660061064126626961370053668149270753,
  0.620526182989242861140477556431189299,
  0.579345235826361691756024932172540496,
  0.536624148142019899264169793311072794,
  0.492480467861778574993693061207708796,
  0.447033769538089176780609900322854,
  0.400401254830394392535476211542660634,
  0.352704725530878113471037207089373861,
  0.304073202273625077372677107199256554,
  0.254636926167889846439805129817805108,
  0.20452511668230989143895767100202471,
  0.15386991360858354696379467274325592,
  0.102806937966737030147096751318000592,
  0.0514718425553176958330252131667225737,
  0.0
};
static const double WGK61[31] = {
  0.00138901369867700762455159122675969968,
  0.00389046112709988405126720184451550328,
  0.00663070391593129217331982636975016813,
  0.00927327965951776342844114689202436042,
  0.0118230152534963417422328988532505929,
  0.0143697295070458048124514324435800102,
  0.0169208891890532726275722894203220924,
  0.0194141411939423811734089510501284559,
  0.0218280358216091922971674857383389934,
  0.0241911620780806013656863707252320268,
  0.0265099548823331016106017093350754144,
  0.0287540487650412928439787853543342111,
  0.0309072575623877624728842529430922726,
  0.0329814470574837260318141910168539275,
  0.0349793380280600241374996707314678751,
  0.0368823646518212292239110656171359677,
  0.0386789456247275929503486515322810503,
  0.0403745389515359591119952797524681142,
  0.0419698102151642461471475412859697578,
  0.0434525397013560693168317281170732581,
  0.0448148001331626631923555516167232438,
  0.0460592382710069881162717355593735806,
  0.0471855465692991539452614781810994865,
  0.0481858617570871291407794922983045926,
  0.0490554345550297788875281653672381736,
  0.0497956834270742063578115693799423285,
  0.0504059214027823468408930856535850289,
  0.0508817958987496064922974730498046919,
  0.0512215478492587721706562826049442083,
  0.0514261285374590259338628792157812598,
  0.0514947294294515675583404336470993075
};
static const double WG30[15] = {
  0.00796819249616660561546588347467362245,
  0.0184664683110909591423021319120472691,
  0.0287847078833233693497191796112920436,
  0.038799192569627049596801936446347692,
  0.0484026728305940529029381404228075178,
  0.0574931562176190664817216894020561288,
  0.0659742298821804951281285151159623612,
  0.0737559747377052062682438500221907342,
  0.0807558952294202153546949384605297309,
  0.0868997872010829798023875307151257026,
  0.0921225222377861287176327070876187672,
  0.0963687371746442596394686263518098651,
  0.0995934205867952670627802821035694765,
  0.101762389748405504596428952168554045,
  0.102852652893558840341285636705415044
};

static double scale_error(double err, double rabs, double rasc)
{
    err = fabs(err);
    if (rasc != 0.0 && err != 0.0) {
        double s = pow(200.0 * err / rasc, 1.5);
        err = (s < 1.0) ? rasc * s : rasc;
    }
    if (rabs > DBL_MIN / (50.0 * DBL_EPSILON)) {
        double floor_err = 50.0 * DBL_EPSILON * rabs;
        if (floor_err > err) err = floor_err;
    }
    return err;
}

void or_gk61(or_fn f, void *p, double a, double b, double *res,
             double *err, double *rabs, double *rasc)
{
    double fv1[30], fv2[30];
    double ctr = 0.5 * (a + b), hl = 0.5 * (b - a), ahl = fabs(hl);
    double fc = f(ctr, p);
    double rg = 0.0, rk = fc * WGK61[30];
    double ra = fabs(rk);
    for (int j = 0; j < 15; j++) {           /* Gauss nodes: odd indices */
        int k = 2 * j + 1;
        double dx = hl * XGK61[k];
        double f1 = f(ctr - dx, p), f2 = f(ctr + dx, p);
        fv1[k] = f1; fv2[k] = f2;
        rg += WG30[j] * (f1 + f2);
        rk += WGK61[k] * (f1 + f2);
        ra += WGK61[k] * (fabs(f1) + fabs(f2));
    }
    for (int j = 0; j < 15; j++) {           /* Kronrod-only nodes: even */
        int k = 2 * j;
        double dx = hl * XGK61[k];
        double f1 = f(ctr - dx, p), f2 = f(ctr + dx, p);
        fv1[k] = f1; fv2[k] = f2;
        rk += WGK61[k] * (f1 + f2);
        ra += WGK61[k] * (fabs(f1) + fabs(f2));
    }
    double mean = 0.5 * rk;
    double rs = WGK61[30] * fabs(fc - mean);
    for (int k = 0; k < 30; k++)
        rs += WGK61[k] * (fabs(fv1[k] - mean) + fabs(fv2[k] - mean));
    double e = (rk - rg) * hl;
    rk *= hl; ra *= ahl; rs *= ahl;
    *res = rk; *rabs = ra; *rasc = rs;
    *err = scale_error(e, ra, rs);
}

static int too_small(double a1, double a2, double b2)
{
    double tmp = (1.0 + 100.0 * DBL_EPSILON) * (fabs(a2) + 1000.0 * DBL_MIN);
    return fabs(a1) <= tmp && fabs(b2) <= tmp;
}

int or_qag61(or_fn f, void *p, double a, double b, double epsabs,
             double epsrel, int limit, double *result, double *abserr)
{
    double al[OR_QAG_LIMIT], bl[OR_QAG_LIMIT], rl[OR_QAG_LIMIT], el[OR_QAG_LIMIT];
    if (limit > OR_QAG_LIMIT) limit = OR_QAG_LIMIT;
    double r0, e0, ra0, rs0;
    or_gk61(f, p, a, b, &r0, &e0, &ra0, &rs0);
    double tol = fmax(epsabs, epsrel * fabs(r0));
    double roundoff = 50.0 * DBL_EPSILON * ra0;
    *result = r0; *abserr = e0;
    if (e0 <= roundoff && e0 > tol) return OR_EROUND;
    if ((e0 <= tol && e0 != rs0) || e0 == 0.0) return OR_SUCCESS;
    if (limit == 1) return OR_EMAXITER;
    int n = 1;
    al[0] = a; bl[0] = b; rl[0] = r0; el[0] = e0;
    double area = r0, errsum = e0;
    int it = 1, rt1 = 0, rt2 = 0, etype = 0;
    int imax = 0;
    do {
        double ai = al[imax], bi = bl[imax], ri = rl[imax], ei = el[imax];
        double a1 = ai, b1 = 0.5 * (ai + bi), a2 = b1, b2 = bi;
        double r1, e1, ra1, rs1, r2, e2, ra2, rs2;
        or_gk61(f, p, a1, b1, &r1, &e1, &ra1, &rs1);
        or_gk61(f, p, a2, b2, &r2, &e2, &ra2, &rs2);
        double r12 = r1 + r2, e12 = e1 + e2;
        errsum += e12 - ei;
        area += r12 - ri;
        if (rs1 != e1 && rs2 != e2) {
            double delta = ri - r12;
            if (fabs(delta) <= 1.0e-5 * fabs(r12) && e12 >= 0.99 * ei) rt1++;
            if (it >= 10 && e12 > ei) rt2++;
        }
        tol = fmax(epsabs, epsrel * fabs(area));
        if (errsum > tol) {
            if (rt1 >= 6 || rt2 >= 20) etype = 2;
            if (too_small(a1, a2, b2)) etype = 3;
        }
        /* the half with the larger error keeps the parent's slot */
        if (e2 > e1) {
            al[imax] = a2; bl[imax] = b2; rl[imax] = r2; el[imax] = e2;
            al[n] = a1; bl[n] = b1; rl[n] = r1; el[n] = e1;
        } else {
            al[imax] = a1; bl[imax] = b1; rl[imax] = r1; el[imax] = e1;
            al[n] = a2; bl[n] = b2; rl[n] = r2; el[n] = e2;
        }
        n++;
        imax = 0;
        for (int k = 1; k < n; k++) if (el[k] > el[imax]) imax = k;
        it++;
    } while (it < limit && !etype && errsum > tol);
    double s = 0.0;
    for (int k = 0; k < n; k++) s += rl[k];
    *result = s; *abserr = errsum;
    if (errsum <= tol) return OR_SUCCESS;
    if (etype == 2) return OR_EROUND;
    if (etype == 3) return OR_ESING;
    if (it == limit) return OR_EMAXITER;
    return OR_EFAILED;
}

/* Threads of the OpenMP build (the CPU baseline); the checker build has one. */
#ifdef _OPENMP
#include <omp.h>
#endif
int or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
