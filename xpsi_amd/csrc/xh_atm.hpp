// Four-dimensional table atmosphere and the reference's query order.
//
// Reference behaviour: xpsi/surface_radiation_field/hot_Num4D.pyx:248-438.
#pragma once
#include "xh_common.hpp"
#include "xh_types.hpp"
#include "xh_kernels.hpp"

namespace xh {

// base node of a cubic Lagrange stencil for an in-table query: interval m
// (x_m <= v < x_{m+1}) -> nodes m-1 .. m+2, kept inside the table
__device__ __forceinline__ int n4_base(const double* x, int N, double v)
{
    if (v > x[N - 1]) return N - 4;
    if (v < x[0]) return 0;
    int m = interval_of(x, N, v) - 1;
    return m < 0 ? 0 : (m > N - 4 ? N - 4 : m);
}

// Cubic Lagrange weights on the four nodes b .. b+3 at q (extrapolating
// when q lies outside them), as hot_Num4D.pyx:356-402 forms DIFF x SPACE.
__device__ __forceinline__ void n4_weights(const double* x, int b, double q, double w[4])
{
    const double x0 = x[b], x1 = x[b + 1], x2 = x[b + 2], x3 = x[b + 3];
    const double D0 = (q - x1) * (q - x2) * (q - x3), D1 = (q - x0) * (q - x2) * (q - x3);
    const double D2 = (q - x0) * (q - x1) * (q - x3), D3 = (q - x0) * (q - x1) * (q - x2);
    const double S0 = 1.0 / (x0 - x1) / (x0 - x2) / (x0 - x3), S1 = 1.0 / (x1 - x0) / (x1 - x2) / (x1 - x3);
    const double S2 = 1.0 / (x2 - x0) / (x2 - x1) / (x2 - x3), S3 = 1.0 / (x3 - x0) / (x3 - x1) / (x3 - x2);
    w[0] = D0 * S0; w[1] = D1 * S1; w[2] = D2 * S2; w[3] = D3 * S3;
}

// Four-dimensional cubic Lagrange interpolation of the atmosphere table
// (hot_Num4D.pyx:248-438).  The reference keeps its base node from call to
// call: a query outside the table is clamped to the edge only when the
// previous query's base node was not already at that edge, and extrapolated
// otherwise.  `prev` is the previous query in the reference's order.
__device__ inline double num4d_I(const FoldArgs& F, const double vec_in[4], const double prev[4], int has_prev)
{
    double v[4], Lw[4][4];
    int bn[4];
    const double* ax = F.atm_axes;
    int S[4];
    S[3] = 1;
    S[2] = F.atm_N[3];
    S[1] = F.atm_N[2] * F.atm_N[3];
    S[0] = F.atm_N[1] * S[1];
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const double* x = ax;
        const int N = F.atm_N[d];
        ax += N;
        double q = vec_in[d];
        int pb = has_prev ? n4_base(x, N, prev[d]) : 0;
        int b;
        if (q > x[N - 1]) {
            b = N - 4;
            if (pb != N - 4) q = x[N - 1];
        } else if (q < x[0]) {
            b = 0;
            if (pb != 0) q = x[0];
        } else b = n4_base(x, N, q);
        v[d] = q;
        bn[d] = b;
        n4_weights(x, b, q, Lw[d]);
    }
    const double* T = F.atm_I + (size_t)bn[0] * S[0] + (size_t)bn[1] * S[1] + (size_t)bn[2] * S[2] + bn[3];
    double I = 0.0;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const double wab = Lw[0][a] * Lw[1][b];
            const double* Tab = T + (size_t)a * S[0] + (size_t)b * S[1];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const double wabc = wab * Lw[2][c];
                const double* Tc = Tab + (size_t)c * S[2];
                I += wabc * (Lw[3][0] * Tc[0] + Lw[3][1] * Tc[1] + Lw[3][2] * Tc[2] + Lw[3][3] * Tc[3]);
            }
        }
    if (I < 0.0) return 0.0;
    return I * pow(10.0, 3.0 * v[0]);
}

// The table contracted over (log T, log g) for one parallel: slab[c][l] =
// sum_ab w_a(log T) w_b(log g) I[bT + a][bg + b][c][l] over every mu node c
// and energy node l, with the base nodes and weights of the parallel's own
// log T and log g (extrapolating outside the table).  A query whose
// predecessor is the same leaf's previous energy has these base nodes and no
// clamp on the first three axes (hot_Num4D.pyx:300-351), so its value is a
// bicubic over (mu, log10(E/kT)) of the slab.
__device__ __forceinline__ void n4_ring_weights(const FoldArgs& F, double logT, double logg, int& b0, int& b1,
                                                double w0[4], double w1[4])
{
    const double* ax0 = F.atm_axes;
    const double* ax1 = ax0 + F.atm_N[0];
    b0 = n4_base(ax0, F.atm_N[0], logT);
    b1 = n4_base(ax1, F.atm_N[1], logg);
    n4_weights(ax0, b0, logT, w0);
    n4_weights(ax1, b1, logg, w1);
}

// Weights as n4_weights with the denominators read from a table (Sb = the
// four of base node b).
__device__ __forceinline__ void n4_weights_pre(const double* x, int b, double q, const double* Sb, double w[4])
{
    const double x0 = x[b], x1 = x[b + 1], x2 = x[b + 2], x3 = x[b + 3];
    w[0] = (q - x1) * (q - x2) * (q - x3) * Sb[0];
    w[1] = (q - x0) * (q - x2) * (q - x3) * Sb[1];
    w[2] = (q - x0) * (q - x1) * (q - x3) * Sb[2];
    w[3] = (q - x0) * (q - x1) * (q - x2) * Sb[3];
}

// interval index i of v (x_i <= v < x_{i+1}, the last interval closed),
// moved up from a previous index at or below it (v only grows)
__device__ __forceinline__ int interval_from(const double* x, int N, double v, int i)
{
    while (i < N - 2 && v >= x[i + 1]) i++;
    return i;
}

// One leaf's run of consecutive energies e0 .. e0 + n - 1 (e0 >= 1) on the
// parallel's slab: q_k = log10(E'_k / kT) increasing.  The energy axis clamps
// as the reference's base-node rule does (a query beyond an edge is clamped
// when its predecessor's base node was not already there).  out[k]: the value
// times 10^(3 log T) (p10), clipped at zero.
template <int RUN>
__device__ __forceinline__ void n4_slab_run(const FoldArgs& F, const double* slab, double mu, const double* q,
                                            double qp, int n, double p10, double* out)
{
    const double* x2 = F.atm_axes + F.atm_N[0] + F.atm_N[1];
    const double* x3 = x2 + F.atm_N[2];
    const int N2 = F.atm_N[2], N3 = F.atm_N[3];
    const double* S2 = F.atm_S + 4 * ((F.atm_N[0] - 3) + (F.atm_N[1] - 3));
    const double* S3 = S2 + 4 * (N2 - 3);
    const int b2 = n4_base(x2, N2, mu);
    double w2[4];
    n4_weights_pre(x2, b2, mu, S2 + 4 * b2, w2);
    const double* rows = slab + (size_t)b2 * N3;
    // predecessor's base node, then this run's interval walk
    int pb = n4_base(x3, N3, qp);
    int iv = qp < x3[0] ? 0 : interval_of(x3, N3, qp);
#pragma unroll
    for (int k = 0; k < RUN; k++) {
        if (k >= n) break;
        double v = q[k];
        int b;
        if (v > x3[N3 - 1]) {
            b = N3 - 4;
            if (pb != N3 - 4) v = x3[N3 - 1];
        } else if (v < x3[0]) {
            b = 0;
            if (pb != 0) v = x3[0];
        } else {
            iv = interval_from(x3, N3, q[k], iv);
            b = iv - 1 < 0 ? 0 : (iv - 1 > N3 - 4 ? N3 - 4 : iv - 1);
        }
        double w3[4];
        n4_weights_pre(x3, b, v, S3 + 4 * b, w3);
        const double* c = rows + b;
        double I = 0.0;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const double* r = c + (size_t)m * N3;
            I += w2[m] * (w3[0] * r[0] + w3[1] * r[1] + w3[2] * r[2] + w3[3] * r[3]);
        }
        out[k] = I < 0.0 ? 0.0 : I * p10;
        // this query's own base node, for the next query's clamp rule
        pb = q[k] > x3[N3 - 1] ? N3 - 4 : (q[k] < x3[0] ? 0 : b);
    }
}

// the query made just before parallel i, image I (the last query of the
// nearest earlier contributing parallel/image), for the table atmosphere
__device__ inline void prev_query_state(const FoldArgs& F, const Layout& L, int mi, int i, int I, double* out)
{
    out[0] = 0.0;
    int ri = i, Ii = I - 1;
    while (ri >= 0) {
        const size_t ro = (size_t)mi * L.S_max + ri;
        if (Ii < 0) {
            // last contributing image of the previous parallel
            ri--;
            if (ri < 0) break;
            const size_t r2 = (size_t)mi * L.S_max + ri;
            Ii = -1;
            for (int q = 0; q < L.IO_max; q++) {
                if (!F.valid[r2 * L.IO_max + q]) break;
                Ii = q;
            }
            continue;
        }
        const size_t vix = ro * L.IO_max + Ii;
        int kd = F.lastq[vix];
        if (kd >= 0) {
            const double* g = F.geo + vix * 5 * L.N_L;
            const double kT = K_B_OVER_KEV * pow(10.0, F.logT[ro]);
            out[0] = 1.0;
            out[1] = F.logT[ro];
            out[2] = F.logg[ro];
            out[3] = g[2 * L.N_L + kd];
            out[4] = log10((F.energies[L.N_E - 1] / g[L.N_L + kd]) / kT);
            return;
        }
        Ii--;
    }
}

}  // namespace xh
