// Schwarzschild ray deflection and time-lag tables on the GPU.
//
// Reference behaviour: xpsi/cellmesh/rays.pyx:53-390 (integrands :61-88,
// three ray branches :90-247, table construction and repair :249-390) and the
// lag calibration of xpsi/HotRegion.py:880-901.  The QAG-GK61 quadratures
// (epsrel 1e-12, 100 intervals, as the reference) start with one lane per
// ray: most rays are settled by QAG's first Gauss-Kronrod pass, which a lane
// sums alone; the rest are listed and rerun with one wavefront per ray
// through the full bisection.  A last kernel applies the bad-ray repair, the
// last-ray check and the lag calibration per member.
#include "xh_common.hpp"
#include "xh_types.hpp"
#include "xh_kernels.hpp"

namespace xh {

// 1/sqrt(q) and 1/d from the hardware estimates with two Newton steps each
// (within an ulp or two of the correctly rounded values, at a third of the
// cost of the IEEE sequences); non-finite estimates (q or d zero or not
// positive) fall back to the IEEE operations
__device__ __forceinline__ double rsqrt_nr(double q)
{
    double y = __builtin_amdgcn_rsq(q);
    double e = fma(-q, y * y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-q, y * y, 1.0);
    y = fma(0.5 * y, e, y);
    return isfinite(y) && q > 0.0 ? y : 1.0 / sqrt(q);
}
__device__ __forceinline__ double rcp_nr(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    return isfinite(r) && d != 0.0 ? r : 1.0 / d;
}

// integrands of rays.pyx:61-88; k = 1 / (rc - 1) (or 1 / (R/r_s - 1)) is
// formed once per ray
struct InDef {
    double k;
    __device__ double operator()(double x) const
    {
        const double w = 1.0 - x * x;
        return rsqrt_nr(2.0 - x * x - w * w * k);
    }
};
struct InLag {
    double k;
    __device__ double operator()(double x) const
    {
        const double w = 1.0 - x * x;
        const double X = rsqrt_nr(2.0 - x * x - w * w * k);
        return X * X * rcp_nr(x + X);
    }
};
struct InBoth {
    double k;
    __device__ void both(double x, double& d, double& l) const
    {
        const double w = 1.0 - x * x;
        const double X = rsqrt_nr(2.0 - x * x - w * w * k);
        d = X;
        l = X * X * rcp_nr(x + X);
    }
};
struct OutDef {
    double sas, k;
    __device__ double operator()(double x) const
    {
        const double w = 1.0 - x * x;
        return x * rsqrt_nr(1.0 - sas + x * x * sas * (2.0 - x * x - w * w * k));
    }
};
struct OutLag {
    double sas, k;
    __device__ double operator()(double x) const
    {
        const double w = 1.0 - x * x;
        const double q = 1.0 - sas + x * x * sas * (2.0 - x * x - w * w * k);
        const double ri = rsqrt_nr(q), f = q * ri;
        return x * ri * rcp_nr(1.0 + f);
    }
};
struct OutBoth {
    double sas, k;
    __device__ void both(double x, double& d, double& l) const
    {
        const double w = 1.0 - x * x;
        const double q = 1.0 - sas + x * x * sas * (2.0 - x * x - w * w * k);
        const double ri = rsqrt_nr(q), f = q * ri;
        d = x * ri;
        l = x * ri * rcp_nr(1.0 + f);
    }
};

// QAG (epsrel 1e-12, 100 intervals) of the deflection and lag integrands
// over one interval; the first Gauss-Kronrod pass is shared by the two
template <class FB, class FD, class FL>
__device__ void ray_pair(FB& fb, FD& fd, FL& fl, double a, double b, double& d, double& l, int& st_d,
                         int& st_l)
{
    double r[2], e[2], ra[2], rs[2], ed, el;
    wave_gk61_pair(fb, a, b, r, e, ra, rs);
    st_d = wave_qag61_from(fd, a, b, 0.0, 1.0e-12, 100, r[0], e[0], ra[0], rs[0], d, ed);
    st_l = wave_qag61_from(fl, a, b, 0.0, 1.0e-12, 100, r[1], e[1], ra[1], rs[1], l, el);
}

// deflection and lag (seconds) of the ray leaving radius r_s/u at cos(alpha)
__device__ void wave_ray(double cos_alpha, double r_s, double u, double& defl, double& lag)
{
    const double b_ph = 3.0 * sqrt(3.0) / 2.0;
    double sas = 1.0 - cos_alpha * cos_alpha;
    double sa = sqrt(sas);
    double alpha = acos(cos_alpha);
    double b = sa / (u * sqrt(1.0 - u));
    double d0 = __builtin_nan(""), l0 = __builtin_nan("");
    int sd = XH_SUCCESS, sl = XH_SUCCESS;
    if (b <= b_ph) {
        OutBoth fb{sas, 1.0 / (1.0 / u - 1.0)};
        OutDef fd{sas, 1.0 / (1.0 / u - 1.0)};
        OutLag fl{sas, 1.0 / (1.0 / u - 1.0)};
        ray_pair(fb, fd, fl, 0.0, 1.0, d0, l0, sd, sl);
        if (sd == XH_EFAILED || sl == XH_EFAILED) d0 = __builtin_nan("");
        else {
            d0 *= 2.0 * b * u;
            l0 *= 2.0 * b * b * u * r_s;
        }
    } else if (alpha <= PI / 2.0) {
        double rc, wR;
        if (eq12(sa, 1.0)) { rc = 1.0 / u; wR = 0.0; }
        else {
            rc = 2.0 * b * cos((atan(sqrt(4.0 * b * b / 27.0 - 1.0)) - PI) / 3.0) / sqrt(3.0);
            wR = sqrt(1.0 - rc * u);
        }
        if (isnan(wR)) wR = 0.0;
        InBoth fb{1.0 / (rc - 1.0)};
        InDef fd{1.0 / (rc - 1.0)};
        InLag fl{1.0 / (rc - 1.0)};
        ray_pair(fb, fd, fl, wR, 1.0, d0, l0, sd, sl);
        if (sd == XH_EFAILED || sl == XH_EFAILED) d0 = __builtin_nan("");
        else {
            d0 *= 2.0 * b / rc;
            l0 *= 2.0 * b * b * r_s / rc;
        }
    } else {
        double rc = 2.0 * b * cos((atan(sqrt(4.0 * b * b / 27.0 - 1.0)) - PI) / 3.0) / sqrt(3.0);
        double wR = sqrt(1.0 - rc * u);
        if (isnan(wR)) wR = 0.0;
        InBoth fb{1.0 / (rc - 1.0)};
        InDef fd{1.0 / (rc - 1.0)};
        InLag fl{1.0 / (rc - 1.0)};
        double D1 = 0, D0 = 0, L1 = 0, L0 = 0;
        int s1, s2, s3, s4;
        ray_pair(fb, fd, fl, 0.0, 1.0, D1, L1, s1, s3);
        ray_pair(fb, fd, fl, wR, 1.0, D0, L0, s2, s4);
        if (s1 == XH_EFAILED || s2 == XH_EFAILED || s3 == XH_EFAILED || s4 == XH_EFAILED) {
            d0 = __builtin_nan("");
            sd = XH_EFAILED;
        } else {
            d0 = 2.0 * b * (2.0 * D1 - D0) / rc;
            l0 = 2.0 * b * b * r_s * (2.0 * L1 - L0) / rc;
            double cor = 1.0 / u - rc + log((1.0 / u - 1.0) / (rc - 1.0));
            l0 += 2.0 * r_s * cor;
        }
    }
    if (sd != XH_EFAILED && sl != XH_EFAILED) l0 /= C_SI;
    defl = d0;
    lag = l0;
}

__device__ __forceinline__ double clamped_max_alpha(double maxA, double u)
{
    const double b_ph = 3.0 * sqrt(3.0) / 2.0, eps = 1.0e-8;
    double extreme = 0.0;
    if (u < 2.0 / 3.0) extreme = PI - asin(sqrt(1.0 - u) * b_ph * u);
    else if (eq12(u, 2.0 / 3.0)) extreme = PI / 2.0;
    else if (2.0 / 3.0 < u && u < 1.0) extreme = asin(sqrt(1.0 - u) * b_ph * u);
    if (maxA >= (1.0 - eps) * extreme) maxA = (1.0 - eps) * extreme;
    return maxA;
}

// ---------------------------------------------------------------- lane per ray
// The same 61-point Gauss-Kronrod pass as wave_gk61_pair, summed by one lane
// over the nodes (uniform node and weight loads, no cross-lane reductions), in
// one sweep.  QUADPACK's resasc, sum w_k |f_k - m| with m the Kronrod mean, is
// known only after the sweep; the sweep sums w_k |f_k - f_c| against the
// centre value instead, and since the weights sum to 2 the true resasc lies
// within 2 |m - f_c| of that (times the half-length).  resasc enters only
// QAG's error estimate, never the result, so the pass's acceptance is decided
// over that whole range and a ray whose decision the range leaves open goes
// to the full QAG (k_rays).
template <class FB>
__device__ void lane_gk61_pair(const FB& f, double a, double b, double res[2], double eraw[2], double rabs[2],
                               double rasc_lo[2], double rasc_hi[2])
{
    const double ctr = 0.5 * (a + b), hl = 0.5 * (b - a), ahl = fabs(hl);
    double fc0, fc1;
    f.both(ctr, fc0, fc1);
    double rk0 = WGK61[30] * fc0, rk1 = WGK61[30] * fc1;
    double rg0 = 0.0, rg1 = 0.0;
    double ra0 = WGK61[30] * fabs(fc0), ra1 = WGK61[30] * fabs(fc1);
    double rc0 = 0.0, rc1 = 0.0;   // sum w_k |f_k - f_c|
    // the next node's abscissa and weights are read one step ahead, so the
    // sweep does not wait on the scalar loads
    double xk = XGK61[0], wk = WGK61[0];
#pragma unroll 1
    for (int k = 0; k < 30; k++) {
        const double xn = XGK61[k + 1], wn = WGK61[k + 1];
        const double wg = (k & 1) ? WG30[k >> 1] : 0.0;
        const double dx = hl * xk;
        double fa0, fa1, fb0, fb1;
        f.both(ctr - dx, fa0, fa1);
        f.both(ctr + dx, fb0, fb1);
        rk0 += wk * (fa0 + fb0);
        rk1 += wk * (fa1 + fb1);
        ra0 += wk * (fabs(fa0) + fabs(fb0));
        ra1 += wk * (fabs(fa1) + fabs(fb1));
        rc0 += wk * (fabs(fa0 - fc0) + fabs(fb0 - fc0));
        rc1 += wk * (fabs(fa1 - fc1) + fabs(fb1 - fc1));
        if (k & 1) {
            rg0 += wg * (fa0 + fb0);
            rg1 += wg * (fa1 + fb1);
        }
        xk = xn;
        wk = wn;
    }
    const double d0 = 2.0 * fabs(fc0 - 0.5 * rk0), d1 = 2.0 * fabs(fc1 - 0.5 * rk1);
    const double sl0 = 8.0 * DBL_EPSILON * (rc0 + d0), sl1 = 8.0 * DBL_EPSILON * (rc1 + d1);   // rounding slack
    res[0] = rk0 * hl; res[1] = rk1 * hl;
    eraw[0] = (rk0 - rg0) * hl; eraw[1] = (rk1 - rg1) * hl;
    rabs[0] = ra0 * ahl; rabs[1] = ra1 * ahl;
    rasc_lo[0] = fmax(rc0 - d0 - sl0, 0.0) * ahl; rasc_lo[1] = fmax(rc1 - d1 - sl1, 0.0) * ahl;
    rasc_hi[0] = (rc0 + d0 + sl0) * ahl; rasc_hi[1] = (rc1 + d1 + sl1) * ahl;
}

// QAG's acceptance of its first pass (epsabs 0): the pass's result stands
// when its error is within the tolerance (or is zero), or when it is below
// the round-off level (QAG then returns that result with a round-off flag)
__device__ __forceinline__ bool qag_first_pass_stands(double r, double e, double rabs, double rasc, double epsrel)
{
    const double tol = epsrel * fabs(r), roundoff = 50.0 * DBL_EPSILON * rabs;
    if (e <= roundoff && e > tol) return true;
    return (e <= tol && e != rasc) || e == 0.0;
}

// The acceptance over resasc in [lo, hi]: the scaled error is monotone on
// either side of resasc = 200 |E| (where QUADPACK's scaling saturates), so the
// acceptance is settled when it agrees at lo, hi and that point.  Returns
// true only when the pass stands for every resasc in the range.
__device__ __forceinline__ bool first_pass_stands(double r, double eraw, double rabs, double lo, double hi,
                                                  double epsrel)
{
    auto P = [&](double rasc) {
        return qag_first_pass_stands(r, scale_error(eraw, rabs, rasc), rabs, rasc, epsrel);
    };
    if (!P(lo) || !P(hi)) return false;
    const double kink = 200.0 * fabs(eraw);
    if (lo < kink && kink < hi && !P(kink)) return false;
    return true;
}

template <class FB>
__device__ __forceinline__ bool lane_ray_pair(const FB& fb, double a, double b, double& d, double& l)
{
    double r[2], e[2], ra[2], lo[2], hi[2];
    lane_gk61_pair(fb, a, b, r, e, ra, lo, hi);
    d = r[0];
    l = r[1];
    return first_pass_stands(r[0], e[0], ra[0], lo[0], hi[0], 1.0e-12) &&
           first_pass_stands(r[1], e[1], ra[1], lo[1], hi[1], 1.0e-12);
}

// Both integrands of one ray, outgoing (rays.pyx:61-70) or ingoing (:72-88)
// by a lane's own flag: the same operations as OutBoth and InBoth, so lanes of
// either kind share one quadrature sweep
struct AnyBoth {
    double sas, k;
    bool out;
    __device__ void both(double x, double& d, double& l) const
    {
        const double w = 1.0 - x * x;
        const double qi = 2.0 - x * x - w * w * k;
        const double q = out ? 1.0 - sas + x * x * sas * qi : qi;
        const double ri = rsqrt_nr(q);
        const double xr = x * ri;
        d = out ? xr : ri;
        l = (out ? xr : ri * ri) * rcp_nr(out ? 1.0 + q * ri : x + ri);
    }
};

// wave_ray's branches with every quadrature settled by its first pass;
// returns false when some quadrature needs QAG's bisection.  The three
// branches are one sweep with per-lane integrand and limits (a second sweep
// for the rays beyond 90 degrees), so a wavefront whose rays straddle a
// branch boundary does not run each branch's sweep in turn.
__device__ bool lane_ray(double cos_alpha, double r_s, double u, double& defl, double& lag)
{
    const double b_ph = 3.0 * sqrt(3.0) / 2.0;
    double sas = 1.0 - cos_alpha * cos_alpha;
    double sa = sqrt(sas);
    double alpha = acos(cos_alpha);
    double b = sa / (u * sqrt(1.0 - u));
    // branch 0: b <= b_ph; 1: alpha <= pi/2; 2: beyond
    const int br = b <= b_ph ? 0 : (alpha <= PI / 2.0 ? 1 : 2);
    double rc = 1.0 / u, wR = 0.0;
    if (br == 1 && eq12(sa, 1.0)) {
        rc = 1.0 / u;
        wR = 0.0;
    } else if (br != 0) {
        rc = 2.0 * b * cos((atan(sqrt(4.0 * b * b / 27.0 - 1.0)) - PI) / 3.0) / sqrt(3.0);
        wR = sqrt(1.0 - rc * u);
    }
    if (isnan(wR)) wR = 0.0;
    const AnyBoth fb{sas, br == 0 ? 1.0 / (1.0 / u - 1.0) : 1.0 / (rc - 1.0), br == 0};
    const int npass = br == 2 ? 2 : 1;
    double D1 = 0.0, L1 = 0.0, D0 = 0.0, L0 = 0.0;
    bool ok = true;
#pragma unroll 1
    for (int ps = 0; ps < npass; ps++) {
        const double a = (br == 1 || ps == 1) ? wR : 0.0;
        double dd, ll;
        const bool st = lane_ray_pair(fb, a, 1.0, dd, ll);
        ok = ok && st;
        if (ps == 0) { D1 = dd; L1 = ll; }
        else { D0 = dd; L0 = ll; }
    }
    if (!ok) return false;
    double d0, l0;
    if (br == 0) {
        d0 = D1 * (2.0 * b * u);
        l0 = L1 * (2.0 * b * b * u * r_s);
    } else if (br == 1) {
        d0 = D1 * (2.0 * b / rc);
        l0 = L1 * (2.0 * b * b * r_s / rc);
    } else {
        d0 = 2.0 * b * (2.0 * D1 - D0) / rc;
        l0 = 2.0 * b * b * r_s * (2.0 * L1 - L0) / rc;
        double cor = 1.0 / u - rc + log((1.0 / u - 1.0) / (rc - 1.0));
        l0 += 2.0 * r_s * cor;
    }
    defl = d0;
    lag = l0 / C_SI;
    return true;
}

// lane per (ray j >= 1, parallel i, member): rays settled by the first pass
// are written; the others go onto the work list for k_rays
__global__ void __launch_bounds__(64) k_rays_lane(const Member* __restrict__ members, Layout L,
                                                  const double* __restrict__ rr,
                                                  const double* __restrict__ maxAlpha_in,
                                                  double* __restrict__ cos_alpha,
                                                  double* __restrict__ defl_raw,
                                                  double* __restrict__ lag_raw,
                                                  int* __restrict__ work, int* __restrict__ nwork,
                                                  const double* __restrict__ uu)
{
    // the member's (parallel, ray) pairs packed into full wavefronts: lane q
    // takes ray j = q mod (N_R - 1) + 1 of parallel i = q / (N_R - 1)
    const int mi = blockIdx.z;
    const int q = blockIdx.x * 64 + threadIdx.x, nr = L.N_R - 1;
    const int i = q / nr, j = q - i * nr + 1;
    const Member m = members[mi];
    if (!m.active || m.status || i >= m.S) return;
    const size_t ro = (size_t)mi * L.S_max + i;
    const double u = uu ? uu[ro] : m.r_s / rr[ro];
    const double maxA = clamped_max_alpha(maxAlpha_in[ro], u);
    const double inc = (1.0 - cos(maxA)) / ((double)L.N_R - 1.0);
    const double ca = 1.0 - (double)j * inc;
    const size_t o = ro * L.N_R + j;
    cos_alpha[o] = ca;
    double d, l;
    if (lane_ray(ca, m.r_s, u, d, l)) {
        defl_raw[o] = d;
        lag_raw[o] = l;
    } else {
        const int slot = atomicAdd(nwork, 1);
        work[slot] = (int)o;
    }
    if (j == 1) {
        cos_alpha[ro * L.N_R] = 1.0;
        defl_raw[ro * L.N_R] = 0.0;
        lag_raw[ro * L.N_R] = 0.0;
    }
}

// wave per listed ray (grid-stride over the work list): the full QAG with
// bisection; writes raw (unrepaired) values
__global__ void __launch_bounds__(256) k_rays(const Member* __restrict__ members, Layout L,
                                              const double* __restrict__ rr,
                                              const double* __restrict__ maxAlpha_in,
                                              double* __restrict__ defl_raw,
                                              double* __restrict__ lag_raw,
                                              const int* __restrict__ work, const int* __restrict__ nwork,
                                              const double* __restrict__ uu)
{
    const int n = *nwork;
    const int waves = gridDim.x * 4;
    for (int w = blockIdx.x * 4 + (threadIdx.x >> 6); w < n; w += waves) {
        const size_t o = (size_t)work[w];
        const int j = (int)(o % (size_t)L.N_R);
        const size_t ro = o / (size_t)L.N_R;
        const int mi = (int)(ro / (size_t)L.S_max);
        const Member m = members[mi];
        const double u = uu ? uu[ro] : m.r_s / rr[ro];
        const double maxA = clamped_max_alpha(maxAlpha_in[ro], u);
        const double inc = (1.0 - cos(maxA)) / ((double)L.N_R - 1.0);
        const double ca = 1.0 - (double)j * inc;
        double d, l;
        wave_ray(ca, m.r_s, u, d, l);
        if ((threadIdx.x & 63) == 0) {
            defl_raw[o] = d;
            lag_raw[o] = l;
        }
    }
}

__device__ __forceinline__ bool ray_bad(double d) { return isnan(d) || d < 0.0 || d > 100.0 * PI; }

// per member: clamp maxAlpha, repair bad rays (rays.pyx:331-369, one-thread
// order), check the last ray, calibrate the lag (HotRegion.py:880-901)
__global__ void __launch_bounds__(256) k_rays_fix(Member* __restrict__ members, Layout L,
                                                  const double* __restrict__ rr,
                                                  double* __restrict__ maxAlpha,
                                                  double* __restrict__ defl,
                                                  double* __restrict__ lag,
                                                  double* __restrict__ maxDefl, int calibrate,
                                                  const double* __restrict__ uu)
{
    const int mi = blockIdx.x;
    Member* mp = members + mi;
    if (!mp->active || mp->status) return;
    const int S = mp->S, N_R = L.N_R;
    const double r_s = mp->r_s;
    double* D = defl + (size_t)mi * L.S_max * N_R;
    double* Lg = lag + (size_t)mi * L.S_max * N_R;
    __shared__ int any_bad, last_bad;
    if (threadIdx.x == 0) { any_bad = 0; last_bad = 0; }
    __syncthreads();
    for (int q = threadIdx.x; q < S * N_R; q += blockDim.x) {
        int j = q % N_R;
        if (j > 0 && j < N_R - 1 && ray_bad(D[q])) atomicOr(&any_bad, 1);
        if (j == N_R - 1 && (isnan(D[q]) || D[q] <= D[q - 1])) atomicOr(&last_bad, 1);
    }
    __syncthreads();
    if (any_bad) {
        // rare path: the reference's sequential repair, counter kept across parallels
        if (threadIdx.x == 0) {
            int counter = 0, idx = 0, term = 0;
            for (int i = 0; i < S && !term; i++) {
                double* Di = D + (size_t)i * N_R;
                double* Li = Lg + (size_t)i * N_R;
                for (int j = 1; j < N_R; j++) {
                    double o0 = Di[j], o1 = Li[j];
                    if (j == N_R - 1 && (isnan(o0) || o0 <= Di[j - 1])) { term = 1; break; }
                    else if (counter > 0 && !isnan(o0) && 0.0 <= o0 && o0 < 100.0 * PI && o0 > Di[idx]) {
                        Di[j] = o0; Li[j] = o1;
                        double delta = (double)counter + 1.0;
                        for (int k = 1; k <= counter; k++) {
                            Di[idx + k] = Di[j] * (double)k;
                            Di[idx + k] += Di[idx] * (delta - (double)k);
                            Di[idx + k] /= delta;
                            Li[idx + k] = Li[j] * (double)k;
                            Li[idx + k] += Li[idx] * (delta - (double)k);
                            Li[idx + k] /= delta;
                        }
                        counter = 0;
                    } else if (isnan(o0) || o0 < 0.0 || o0 > 100.0 * PI) {
                        if (counter == 0) idx = j - 1;
                        counter++;
                    }
                }
            }
            if (term) mp->status = 1;
        }
    } else if (last_bad) {
        if (threadIdx.x == 0) mp->status = 1;
    }
    __syncthreads();
    if (mp->status) return;
    const double R_i = mp->R_eq / r_s, w = mp->mode_freq * TWO_PI;
    for (int i = threadIdx.x; i < S; i += blockDim.x) {
        size_t ro = (size_t)mi * L.S_max + i;
        double u = uu ? uu[ro] : r_s / rr[ro];
        maxAlpha[ro] = clamped_max_alpha(maxAlpha[ro], u);
        maxDefl[ro] = D[(size_t)i * N_R + N_R - 1];
    }
    if (!calibrate) return;   // compute_rays itself returns the raw lag
    for (int q = threadIdx.x; q < S * N_R; q += blockDim.x) {
        int i = q / N_R;
        double u = r_s / rr[(size_t)mi * L.S_max + i];
        double Cc = (1.0 / u - R_i) + log((1.0 / u - 1.0) / (R_i - 1.0));
        Cc *= r_s / C_SI;
        Lg[q] = (Lg[q] - Cc) * w;
    }
}

}  // namespace xh
