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
        const double x0 = x[b], x1 = x[b + 1], x2 = x[b + 2], x3 = x[b + 3];
        double D0 = (q - x1) * (q - x2) * (q - x3), D1 = (q - x0) * (q - x2) * (q - x3);
        double D2 = (q - x0) * (q - x1) * (q - x3), D3 = (q - x0) * (q - x1) * (q - x2);
        double S0 = 1.0 / (x0 - x1) / (x0 - x2) / (x0 - x3), S1 = 1.0 / (x1 - x0) / (x1 - x2) / (x1 - x3);
        double S2 = 1.0 / (x2 - x0) / (x2 - x1) / (x2 - x3), S3 = 1.0 / (x3 - x0) / (x3 - x1) / (x3 - x2);
        Lw[d][0] = D0 * S0; Lw[d][1] = D1 * S1; Lw[d][2] = D2 * S2; Lw[d][3] = D3 * S3;
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
