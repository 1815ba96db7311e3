// Device numerics shared by the X-PSI likelihood kernels (gfx950, wave64).
//
// - interval search with the convention of the reference's interpolation
//   library: largest i in [0, n-2] with x_i <= x (x == x_max -> n-2);
// - Steffen (1990) cubic and Akima (1970) cubic evaluated from a local
//   stencil, so no coefficient tables are needed for one-off lookups;
// - wave-cooperative adaptive quadrature: QUADPACK QAG with the 61-point
//   Gauss-Kronrod rule, one wavefront per integral, lane = Kronrod node,
//   interval list held two slots per lane in VGPRs.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

#define XH_SUCCESS 0
#define XH_EFAILED 5
#define XH_EMAXITER 11
#define XH_EROUND 18
#define XH_ESING 21

namespace xh {

constexpr double PI = 3.14159265358979323846;
constexpr double TWO_PI = 2.0 * PI;
constexpr double HALF_PI = 0.5 * PI;
constexpr double C_SI = 2.99792458e8;
constexpr double KEV = 1.60217662e-16;
constexpr double K_B = 1.38064852e-23;
constexpr double H_KEV = 4.135667662e-18;
constexpr double K_B_OVER_KEV = K_B / KEV;
constexpr double BB_NORM = 1.0e-7 * 5.040366110812353e22;
constexpr double NUM4D_NORM = 1.0e-7 / H_KEV;
constexpr double MESH_EPSREL = 1.0e-10;
constexpr double MARG_EPSREL = 1.0e-11;
constexpr int QAG_LIMIT = 100;

__device__ __forceinline__ bool eq12(double x, double y) { return fabs(x - y) < 1.0e-12; }

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ int wave_or(int v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v |= __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ double wave_bcast(double v, int src) { return __shfl(v, src, 64); }

// largest i in [0, n-2] with x[i] <= xq (x ascending)
__device__ __forceinline__ int interval_of(const double* __restrict__ x, int n, double xq)
{
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) {
        int mid = (hi + lo) >> 1;
        if (x[mid] > xq) hi = mid; else lo = mid;
    }
    return lo;
}

__device__ __forceinline__ double sgn1(double s) { return s < 0.0 ? -1.0 : 1.0; }

// Steffen slope at node i (one-sided at both ends)
__device__ __forceinline__ double steffen_slope(const double* x, const double* y, int n, int i)
{
    if (i == 0) return (y[1] - y[0]) / (x[1] - x[0]);
    if (i == n - 1) return (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
    double hi = x[i + 1] - x[i], hm = x[i] - x[i - 1];
    double si = (y[i + 1] - y[i]) / hi, sm = (y[i] - y[i - 1]) / hm;
    double p = (sm * hi + si * hm) / (hm + hi);
    double m = fmin(fabs(sm), fmin(fabs(si), 0.5 * fabs(p)));
    return (sgn1(sm) + sgn1(si)) * m;
}

// value (and derivative) of the Steffen interpolant through (x, y) at xq
__device__ inline double steffen_eval(const double* x, const double* y, int n, double xq, double* deriv)
{
    int i = interval_of(x, n, xq);
    double h = x[i + 1] - x[i];
    double s = (y[i + 1] - y[i]) / h;
    double y0 = steffen_slope(x, y, n, i);
    double y1 = steffen_slope(x, y, n, i + 1);
    double a = (y0 + y1 - 2.0 * s) / h / h;
    double b = (3.0 * s - 2.0 * y0 - y1) / h;
    double c = y0;
    double t = xq - x[i];
    if (deriv) *deriv = c + t * (2.0 * b + t * 3.0 * a);
    return y[i] + t * (c + t * (b + t * a));
}

// Akima secant slope m[k] for k in [-2, n]; periodic or natural ends
__device__ __forceinline__ double akima_m(const double* x, const double* y, int n, bool periodic, int k)
{
    auto sec = [&](int j) { return (y[j + 1] - y[j]) / (x[j + 1] - x[j]); };
    if (k >= 0 && k <= n - 2) return sec(k);
    if (periodic) {
        if (k == -2) return sec(n - 3);
        if (k == -1) return sec(n - 2);
        if (k == n - 1) return sec(0);
        return sec(1);                       // k == n
    }
    if (k == -2) return 3.0 * sec(0) - 2.0 * sec(1);
    if (k == -1) return 2.0 * sec(0) - sec(1);
    if (k == n - 1) return 2.0 * sec(n - 2) - sec(n - 3);
    return 3.0 * sec(n - 2) - 2.0 * sec(n - 3); // k == n
}

// A cell's phase point wrapped into the leaves' span [lo, hi] as the
// integrator wraps it (integrator_for_azimuthal_invariance.pyx:565-596):
// while (x > hi) x -= 2 pi; while (x < lo) x += 2 pi
__device__ __forceinline__ double wrap_phase(double x, double lo, double hi)
{
    if (x > hi) x -= TWO_PI;
    else if (x < lo) x += TWO_PI;
    if (x > hi || x < lo) {
        int guard = 0;
#pragma unroll 1
        while (x > hi && guard++ < 64) x -= TWO_PI;
#pragma unroll 1
        while (x < lo && guard++ < 64) x += TWO_PI;
    }
    return x;
}

// Akima piece i from slopes m[i-2..i+2]: y = y_i + t(b + t(c + t d))
__device__ __forceinline__ void akima_piece(double mm2, double mm1, double m0, double mp1, double mp2,
                                            double h, double& b, double& c, double& d)
{
    double w = fabs(mp1 - m0) + fabs(mm1 - mm2);
    if (w == 0.0) { b = m0; c = 0.0; d = 0.0; return; }
    double wn = fabs(mp2 - mp1) + fabs(m0 - mm1);
    double al = fabs(mm1 - mm2) / w;
    double tn;
    if (wn == 0.0) tn = m0;
    else {
        double aln = fabs(m0 - mm1) / wn;
        tn = (1.0 - aln) * m0 + aln * mp1;
    }
    b = (1.0 - al) * mm1 + al * m0;
    c = (3.0 * m0 - 2.0 * b - tn) / h;
    d = (b + tn - 2.0 * m0) / (h * h);
}

// Akima (1970) piece from the five surrounding secant slopes, with the
// degenerate-weight branches of the reference's interpolation library
template <typename R>
__device__ __forceinline__ void akima_piece_t(R mm2, R mm1, R m0, R mp1, R mp2, R h, R& b, R& c, R& d)
{
    const R one = (R)1, two = (R)2, three = (R)3;
    R w = fabs(mp1 - m0) + fabs(mm1 - mm2);
    if (w == (R)0) { b = m0; c = (R)0; d = (R)0; return; }
    R wn = fabs(mp2 - mp1) + fabs(m0 - mm1);
    R al = fabs(mm1 - mm2) / w;
    R tn;
    if (wn == (R)0) tn = m0;
    else {
        R aln = fabs(m0 - mm1) / wn;
        tn = (one - aln) * m0 + aln * mp1;
    }
    b = (one - al) * mm1 + al * m0;
    c = (three * m0 - two * b - tn) / h;
    d = (b + tn - two * m0) / (h * h);
}

__device__ inline void akima_coefs(const double* x, const double* y, int n, bool periodic, int i,
                                   double& b, double& c, double& d)
{
    akima_piece(akima_m(x, y, n, periodic, i - 2), akima_m(x, y, n, periodic, i - 1),
                akima_m(x, y, n, periodic, i), akima_m(x, y, n, periodic, i + 1),
                akima_m(x, y, n, periodic, i + 2), x[i + 1] - x[i], b, c, d);
}

__device__ __forceinline__ double akima_piece_integral(double yi, double bi, double ci, double di,
                                                       double xi, double lo, double hi)
{
    double r1 = lo - xi, r2 = hi - xi, r12 = r1 + r2;
    double bt = 0.5 * bi * r12;
    double ct = (1.0 / 3.0) * ci * (r1 * r1 + r2 * r2 + r1 * r2);
    double dt = 0.25 * di * r12 * (r1 * r1 + r2 * r2);
    return (hi - lo) * (yi + bt + ct + dt);
}

// exact integral of the Akima interpolant over [lo, hi] (NaN outside range)
__device__ inline double akima_integ(const double* x, const double* y, int n, bool periodic,
                                     double lo, double hi)
{
    if (lo > hi || lo < x[0] || hi > x[n - 1]) return __builtin_nan("");
    if (lo == hi) return 0.0;
    int ia = interval_of(x, n, lo), ib = interval_of(x, n, hi);
    double s = 0.0;
    for (int i = ia; i <= ib; i++) {
        double xl = x[i], xh = x[i + 1], h = xh - xl;
        if (h == 0.0) return 0.0;
        double b, c, d;
        akima_coefs(x, y, n, periodic, i, b, c, d);
        if (i == ia || i == ib) {
            double a1 = (i == ia) ? lo : xl;
            double a2 = (i == ib) ? hi : xh;
            s += akima_piece_integral(y[i], b, c, d, xl, a1, a2);
        } else {
            s += h * (y[i] + h * (0.5 * b + h * (c / 3.0 + 0.25 * d * h)));
        }
    }
    return s;
}

// ---- the phase interpolant (xpsi/tools/core.pyx:21-116), a process-wide
// choice: 0 periodic Akima (default), 1 Steffen, 2 periodic cubic spline.
// Each gives y = y_i + t (b + t (c + t d)) on interval i, so the evaluation
// and the piece integrals are shared.

// Steffen piece i from the slopes at its two nodes
__device__ inline void steffen_coefs(const double* x, const double* y, int n, int i, double& b, double& c,
                                     double& d)
{
    const double h = x[i + 1] - x[i];
    const double s = (y[i + 1] - y[i]) / h;
    const double y0 = steffen_slope(x, y, n, i), y1 = steffen_slope(x, y, n, i + 1);
    b = y0;
    c = (3.0 * s - 2.0 * y0 - y1) / h;
    d = (y0 + y1 - 2.0 * s) / h / h;
}

// Periodic cubic spline, one thread: the unique C2 periodic interpolant
// (y[n-1] closes the period).  c_i = y''(x_i) / 2 solves the cyclic
// tridiagonal system h_{i-1} c_{i-1} + 2 (h_{i-1} + h_i) c_i + h_i c_{i+1} =
// 3 (s_i - s_{i-1}) over m = n - 1 unknowns (indices modulo m), by the Thomas
// algorithm with the Sherman-Morrison correction for the corner entries;
// then b_i = s_i - h_i (c_{i+1} + 2 c_i) / 3 and d_i = (c_{i+1} - c_i) / (3 h_i).
// b, c, d: n entries each (c[n-1] = c[0]); z: n of work.  Needs n >= 4.
__device__ inline void cspline_periodic(const double* x, const double* y, int n, double* b, double* c, double* d,
                                        double* z)
{
    const int m = n - 1;
    const double e_last = x[m] - x[m - 1];
    double dg0 = 0.0;
    for (int i = 0; i < m; i++) {
        const int im = (i + m - 1) % m;
        const double hm = x[im + 1] - x[im], hi = x[i + 1] - x[i];
        const double dg = 2.0 * (hm + hi);
        if (i == 0) dg0 = dg;
        b[i] = dg;                                              // diagonal
        c[i] = 3.0 * ((y[i + 1] - y[i]) / hi - (y[im + 1] - y[im]) / hm);   // right-hand side
        z[i] = 0.0;
    }
    const double g = -dg0;
    b[0] -= g;
    b[m - 1] -= e_last * e_last / g;
    z[0] = g;
    z[m - 1] = e_last;
    double den = b[0];
    d[0] = (x[1] - x[0]) / den;
    c[0] = c[0] / den;
    z[0] = z[0] / den;
    for (int i = 1; i < m; i++) {
        const double em = x[i] - x[i - 1];
        den = b[i] - em * d[i - 1];
        d[i] = (i + 1 < m) ? (x[i + 1] - x[i]) / den : 0.0;
        c[i] = (c[i] - em * c[i - 1]) / den;
        z[i] = (z[i] - em * z[i - 1]) / den;
    }
    for (int i = m - 2; i >= 0; i--) {
        c[i] -= d[i] * c[i + 1];
        z[i] -= d[i] * z[i + 1];
    }
    const double fac = (c[0] + c[m - 1] * e_last / g) / (1.0 + z[0] + z[m - 1] * e_last / g);
    for (int i = 0; i < m; i++) c[i] = c[i] - fac * z[i];
    c[m] = c[0];
    for (int i = 0; i < m; i++) {
        const double h = x[i + 1] - x[i];
        b[i] = (y[i + 1] - y[i]) / h - h * (c[i + 1] + 2.0 * c[i]) / 3.0;
        d[i] = (c[i + 1] - c[i]) / (3.0 * h);
    }
}

// piece i of the chosen interpolant; kind 2 reads the tables cspline_periodic wrote
__device__ inline void phase_coefs(int kind, const double* x, const double* y, int n, int i, const double* B,
                                   const double* C, const double* D, double& b, double& c, double& d)
{
    if (kind == 2) { b = B[i]; c = C[i]; d = D[i]; }
    else if (kind == 1) steffen_coefs(x, y, n, i, b, c, d);
    else akima_coefs(x, y, n, true, i, b, c, d);
}

// exact integral of the chosen interpolant over [lo, hi] (NaN outside range)
__device__ inline double phase_integ(int kind, const double* x, const double* y, int n, const double* B,
                                     const double* C, const double* D, double lo, double hi)
{
    if (lo > hi || lo < x[0] || hi > x[n - 1]) return __builtin_nan("");
    if (lo == hi) return 0.0;
    int ia = interval_of(x, n, lo), ib = interval_of(x, n, hi);
    double s = 0.0;
    for (int i = ia; i <= ib; i++) {
        double xl = x[i], xh = x[i + 1], h = xh - xl;
        if (h == 0.0) return 0.0;
        double b, c, d;
        phase_coefs(kind, x, y, n, i, B, C, D, b, c, d);
        if (i == ia || i == ib) {
            double a1 = (i == ia) ? lo : xl;
            double a2 = (i == ib) ? hi : xh;
            s += akima_piece_integral(y[i], b, c, d, xl, a1, a2);
        } else {
            s += h * (y[i] + h * (0.5 * b + h * (c / 3.0 + 0.25 * d * h)));
        }
    }
    return s;
}

// 61-point Gauss-Kronrod nodes (positive half, descending; index 30 = centre)
// and weights; the 30-point Gauss weights belong to the odd Kronrod indices.
// Computed from the Stieltjes polynomial of P_30 in 80-digit arithmetic.
__constant__ const double XGK61[31] = {
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
__constant__ const double WGK61[31] = {
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
__constant__ const double WG30[15] = {
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

__device__ __forceinline__ double scale_error(double err, double rabs, double rasc)
{
    err = fabs(err);
    if (rasc != 0.0 && err != 0.0) {
        double s = pow(200.0 * err / rasc, 1.5);
        err = (s < 1.0) ? rasc * s : rasc;
    }
    if (rabs > DBL_MIN / (50.0 * DBL_EPSILON)) {
        double fl = 50.0 * DBL_EPSILON * rabs;
        if (fl > err) err = fl;
    }
    return err;
}

// One 61-point Gauss-Kronrod pass; every lane of the wave must call it and
// every lane receives the same results.  Lanes 0..29: centre - h*x_k,
// lanes 30..59: centre + h*x_k, lane 60: centre, lanes 61..63 idle.
// Out of line: inlined at each QAG step, the integrand's FP64 temporaries
// multiply the caller's register count (and cut its occupancy).
template <class F>
__device__ __noinline__ void wave_gk61(F& f, double a, double b, double& res, double& err,
                                 double& rabs, double& rasc)
{
    const int lane = lane_id();
    const double ctr = 0.5 * (a + b), hl = 0.5 * (b - a), ahl = fabs(hl);
    int k = lane < 30 ? lane : (lane < 60 ? lane - 30 : 30);
    double wk = lane <= 60 ? WGK61[k] : 0.0;
    double wg = (lane < 60 && (k & 1)) ? WG30[k >> 1] : 0.0;
    double x = lane < 30 ? ctr - hl * XGK61[k] : (lane < 60 ? ctr + hl * XGK61[k] : ctr);
    double fv = 0.0;
    if (lane <= 60) fv = f(x);
    double rk = wave_sum(wk * fv);
    double rg = wave_sum(wg * fv);
    double ra = wave_sum(wk * fabs(fv));
    double mean = 0.5 * rk;
    double rs = wave_sum(lane <= 60 ? wk * fabs(fv - mean) : 0.0);
    double e = (rk - rg) * hl;
    res = rk * hl;
    rabs = ra * ahl;
    rasc = rs * ahl;
    err = scale_error(e, rabs, rasc);
}

// Two integrands on the same Gauss-Kronrod nodes (one pass, shared node
// evaluation); results as wave_gk61 would give for each.
template <class F>
__device__ inline void wave_gk61_pair(F& f, double a, double b, double res[2], double err[2],
                                      double rabs[2], double rasc[2])
{
    const int lane = lane_id();
    const double ctr = 0.5 * (a + b), hl = 0.5 * (b - a), ahl = fabs(hl);
    int k = lane < 30 ? lane : (lane < 60 ? lane - 30 : 30);
    double wk = lane <= 60 ? WGK61[k] : 0.0;
    double wg = (lane < 60 && (k & 1)) ? WG30[k >> 1] : 0.0;
    double x = lane < 30 ? ctr - hl * XGK61[k] : (lane < 60 ? ctr + hl * XGK61[k] : ctr);
    double f0 = 0.0, f1 = 0.0;
    if (lane <= 60) f.both(x, f0, f1);
    double rk0 = wave_sum(wk * f0), rk1 = wave_sum(wk * f1);
    double rg0 = wave_sum(wg * f0), rg1 = wave_sum(wg * f1);
    double ra0 = wave_sum(wk * fabs(f0)), ra1 = wave_sum(wk * fabs(f1));
    double m0 = 0.5 * rk0, m1 = 0.5 * rk1;
    double rs0 = wave_sum(lane <= 60 ? wk * fabs(f0 - m0) : 0.0);
    double rs1 = wave_sum(lane <= 60 ? wk * fabs(f1 - m1) : 0.0);
    res[0] = rk0 * hl; res[1] = rk1 * hl;
    rabs[0] = ra0 * ahl; rabs[1] = ra1 * ahl;
    rasc[0] = rs0 * ahl; rasc[1] = rs1 * ahl;
    err[0] = scale_error((rk0 - rg0) * hl, rabs[0], rasc[0]);
    err[1] = scale_error((rk1 - rg1) * hl, rabs[1], rasc[1]);
}

__device__ __forceinline__ bool qag_too_small(double a1, double a2, double b2)
{
    double tmp = (1.0 + 100.0 * DBL_EPSILON) * (fabs(a2) + 1000.0 * DBL_MIN);
    return fabs(a1) <= tmp && fabs(b2) <= tmp;
}

// QUADPACK QAG (bisect the interval of largest error) with GK61.
// Wave-cooperative: all 64 lanes call with identical arguments.
template <class F>
__device__ inline int wave_qag61_from(F& f, double a, double b, double epsabs, double epsrel, int limit,
                                      double r0, double e0, double ra0, double rs0, double& result,
                                      double& abserr);

template <class F>
__device__ inline int wave_qag61(F& f, double a, double b, double epsabs, double epsrel, int limit,
                                 double& result, double& abserr)
{
    double r0, e0, ra0, rs0;
    wave_gk61(f, a, b, r0, e0, ra0, rs0);
    return wave_qag61_from(f, a, b, epsabs, epsrel, limit, r0, e0, ra0, rs0, result, abserr);
}

// QAG continued from the first 61-point pass over [a, b]
template <class F>
__device__ inline int wave_qag61_from(F& f, double a, double b, double epsabs, double epsrel, int limit,
                                      double r0, double e0, double ra0, double rs0, double& result,
                                      double& abserr)
{
    const int lane = lane_id();
    double tol = fmax(epsabs, epsrel * fabs(r0));
    double roundoff = 50.0 * DBL_EPSILON * ra0;
    result = r0;
    abserr = e0;
    if (e0 <= roundoff && e0 > tol) return XH_EROUND;
    if ((e0 <= tol && e0 != rs0) || e0 == 0.0) return XH_SUCCESS;
    if (limit == 1) return XH_EMAXITER;
    if (limit > 128) limit = 128;
    // slot s lives on lane (s & 63), register set (s >> 6)
    double A0 = 0.0, B0 = 0.0, R0 = 0.0, E0 = -1.0;
    double A1 = 0.0, B1 = 0.0, R1 = 0.0, E1 = -1.0;
    if (lane == 0) { A0 = a; B0 = b; R0 = r0; E0 = e0; }
    int n = 1, imax = 0, it = 1, rt1 = 0, rt2 = 0, etype = 0;
    double area = r0, errsum = e0;
    do {
        int src = imax & 63;
        bool hiset = imax >= 64;
        double ai = wave_bcast(hiset ? A1 : A0, src);
        double bi = wave_bcast(hiset ? B1 : B0, src);
        double ri = wave_bcast(hiset ? R1 : R0, src);
        double ei = wave_bcast(hiset ? E1 : E0, src);
        double a1 = ai, b1 = 0.5 * (ai + bi), a2 = b1, b2 = bi;
        double r1, e1, ra1, rs1, r2, e2, ra2, rs2;
        wave_gk61(f, a1, b1, r1, e1, ra1, rs1);
        wave_gk61(f, a2, b2, r2, e2, ra2, rs2);
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
            if (qag_too_small(a1, a2, b2)) etype = 3;
        }
        double ka, kb, kr, ke, na, nb, nr, ne;
        if (e2 > e1) { ka = a2; kb = b2; kr = r2; ke = e2; na = a1; nb = b1; nr = r1; ne = e1; }
        else { ka = a1; kb = b1; kr = r1; ke = e1; na = a2; nb = b2; nr = r2; ne = e2; }
        if (lane == src) {
            if (hiset) { A1 = ka; B1 = kb; R1 = kr; E1 = ke; }
            else { A0 = ka; B0 = kb; R0 = kr; E0 = ke; }
        }
        if (lane == (n & 63)) {
            if (n >= 64) { A1 = na; B1 = nb; R1 = nr; E1 = ne; }
            else { A0 = na; B0 = nb; R0 = nr; E0 = ne; }
        }
        n++;
        // argmax over valid slots, lowest index on ties
        double bv = E0;
        int bi_ = lane;
        if (lane + 64 < n && E1 > bv) { bv = E1; bi_ = lane + 64; }
        if (lane >= n) { bv = -2.0; bi_ = 1 << 20; }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            double ov = __shfl_xor(bv, m, 64);
            int oi = __shfl_xor(bi_, m, 64);
            if (ov > bv || (ov == bv && oi < bi_)) { bv = ov; bi_ = oi; }
        }
        imax = bi_;
        it++;
    } while (it < limit && !etype && errsum > tol);
    double part = (lane < n ? R0 : 0.0) + (lane + 64 < n ? R1 : 0.0);
    result = wave_sum(part);
    abserr = errsum;
    if (errsum <= tol) return XH_SUCCESS;
    if (etype == 2) return XH_EROUND;
    if (etype == 3) return XH_ESING;
    if (it == limit) return XH_EMAXITER;
    return XH_EFAILED;
}

}  // namespace xh
