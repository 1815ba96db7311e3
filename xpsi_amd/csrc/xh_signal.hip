// Signal registration and the background-marginalised likelihood on the GPU.
//
// Reference behaviour:
//   energy integral  xpsi/tools/energy_integrator.pyx:27-114 (periodic Akima
//                    in log10 E, upper edge clipped at the highest energy);
//   fold             xpsi/Instrument.py:163-197 (matrix @ signal);
//   phase binning    xpsi/tools/compute_expected_counts.pyx:66-197;
//   marginal lnL     xpsi/likelihoods/default_background_marginalisation.pyx
//                    :84-447 (Newton ML background, Gaussian bounds, quadrature).
// The marginal kernel is one wavefront per (channel, request): lanes own phase
// bins for the expected counts and Newton sums, then become Kronrod nodes.
#include "xh_common.hpp"
#include "xh_types.hpp"
#include "xh_kernels.hpp"

namespace xh {

// out[region][j][k] = atten[j] * sum_m integral over bin j of the member's
// specific flux (divided by d^2) in log10 E; one thread per (bin, phase).
__global__ void __launch_bounds__(256) k_energy_integrate(const Member* __restrict__ members,
                                                          EnergyArgs A)
{
    extern __shared__ double wE[];          // [N_E] 10^logE, once per block
    const int reg = blockIdx.y;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int N_in = A.N_edges - 1, N_P = A.N_P, N_E = A.N_E;
    for (int e = threadIdx.x; e < N_E; e += blockDim.x) wE[e] = pow(10.0, A.logE[e]);
    __syncthreads();
    if (q >= N_in * N_P) return;
    const int j = q / N_P, k = q % N_P;
    const double maxE = A.logE[N_E - 1];
    // bins past the first one reaching beyond the highest energy stay zero
    if (j > 0 && A.logEdges[j] > maxE) { A.out[(size_t)reg * N_in * N_P + q] = 0.0; return; }
    const double lo = A.logEdges[j];
    const double hi = A.logEdges[j + 1] > maxE ? maxE : A.logEdges[j + 1];
    double total = 0.0;
    for (int s = 0; s < A.members_per_region; s++) {
        // without a member table (the drop-in): region reg's own signal, d^2 = 1
        int mi = reg;
        double d_sq = 1.0;
        if (A.member_index) {
            mi = A.member_index[reg * A.members_per_region + s];
            if (mi < 0) continue;
            const Member& m = members[mi];
            if (!m.active) continue;
            d_sq = m.d_sq;
        } else if (s > 0) break;
        const double* fl = A.flux + (size_t)mi * N_E * N_P + k;
        auto yv = [&](int e) { return wE[e] * (fl[(size_t)e * N_P] / d_sq) * log(10.0); };
        double v;
        if (lo > hi || lo < A.logE[0] || hi > maxE) v = __builtin_nan("");
        else if (lo == hi) v = 0.0;
        else {
            int ia = interval_of(A.logE, N_E, lo), ib = interval_of(A.logE, N_E, hi);
            v = 0.0;
            for (int i = ia; i <= ib; i++) {
                auto sec = [&](int kk) {
                    int w = kk < 0 ? kk + (N_E - 1) : (kk > N_E - 2 ? kk - (N_E - 1) : kk);
                    return (yv(w + 1) - yv(w)) / (A.logE[w + 1] - A.logE[w]);
                };
                double xl = A.logE[i], xh = A.logE[i + 1], h = xh - xl;
                double b, c, d;
                if (A.interp == 1) {
                    // Steffen: the slopes at the interval's two nodes (one-sided at the ends)
                    auto slope = [&](int kk) {
                        if (kk == 0) return sec(0);
                        if (kk == N_E - 1) return sec(N_E - 2);
                        const double hi_ = A.logE[kk + 1] - A.logE[kk], hm = A.logE[kk] - A.logE[kk - 1];
                        const double si = sec(kk), sm = sec(kk - 1);
                        const double pp = (sm * hi_ + si * hm) / (hm + hi_);
                        return (sgn1(sm) + sgn1(si)) * fmin(fabs(sm), fmin(fabs(si), 0.5 * fabs(pp)));
                    };
                    const double s_ = sec(i), y0 = slope(i), y1 = slope(i + 1);
                    b = y0;
                    c = (3.0 * s_ - 2.0 * y0 - y1) / h;
                    d = (y0 + y1 - 2.0 * s_) / h / h;
                } else {
                    akima_piece(sec(i - 2), sec(i - 1), sec(i), sec(i + 1), sec(i + 2), h, b, c, d);
                }
                double yi = yv(i);
                if (i == ia || i == ib) {
                    double a1 = (i == ia) ? lo : xl;
                    double a2 = (i == ib) ? hi : xh;
                    v += akima_piece_integral(yi, b, c, d, xl, a1, a2);
                } else {
                    v += h * (yi + h * (0.5 * b + h * (c / 3.0 + 0.25 * d * h)));
                }
            }
        }
        total = (s == 0) ? v : total + v;
    }
    if (A.atten) total *= A.atten[j];
    A.out[(size_t)reg * N_in * N_P + q] = total;
}

// The energy integral with the periodic cubic spline (the drop-in; a signal
// with no member table): one block per phase column k, thread 0 solving the
// column's spline into LDS, then one thread per input bin (the integral of
// the spline over the bin, upper edge clipped at the highest energy).
__global__ void __launch_bounds__(256) k_energy_integrate_cs(EnergyArgs A)
{
    extern __shared__ double wE[];
    const int k = blockIdx.x, N_E = A.N_E, N_P = A.N_P, N_in = A.N_edges - 1;
    double* Y = wE;
    double* B = Y + N_E;
    double* C = B + N_E;
    double* D = C + N_E;
    double* Z = D + N_E;
    for (int e = threadIdx.x; e < N_E; e += blockDim.x)
        Y[e] = pow(10.0, A.logE[e]) * A.flux[(size_t)e * N_P + k] * log(10.0);
    __syncthreads();
    if (threadIdx.x == 0) cspline_periodic(A.logE, Y, N_E, B, C, D, Z);
    __syncthreads();
    const double maxE = A.logE[N_E - 1];
    for (int j = threadIdx.x; j < N_in; j += blockDim.x) {
        double v = 0.0;
        if (!(j > 0 && A.logEdges[j] > maxE)) {
            const double lo = A.logEdges[j], hi = A.logEdges[j + 1] > maxE ? maxE : A.logEdges[j + 1];
            v = phase_integ(2, A.logE, Y, N_E, B, C, D, lo, hi);
        }
        if (A.atten) v *= A.atten[j];
        A.out[(size_t)j * N_P + k] = v;
    }
}

// folded[region][c][k] = sum_j M[c][j] * in[region][j][k] on the matrix
// cores (xpsi/Instrument.py:163-197).  A block of four waves owns a 32-channel
// x 32-phase tile, one 16 x 16 sub-tile per wave.  The response matrix is
// shared by every region of the batch and stays in L2.
//
// FP64: v_mfma_f64_16x16x4f64, lane l holds M[c0 + l%16][j + l/16] and
// in[j + l/16][k0 + l%16]; result element r is row 4 r + l/16.
__global__ void __launch_bounds__(256) k_fold_f64(FoldGemmArgs A)
{
    typedef double d4 __attribute__((ext_vector_type(4)));
    const int reg = blockIdx.z, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = blockIdx.y * 32 + (wave >> 1) * 16, k0 = blockIdx.x * 32 + (wave & 1) * 16;
    const int c = c0 + (lane & 15), k = k0 + (lane & 15), jl = lane >> 4;
    const double* X = A.in + (size_t)reg * A.N_in * A.N_P;
    const bool c_ok = c < A.N_ch, k_ok = k < A.N_P;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int j0 = 0; j0 < A.N_in; j0 += 4) {
        const int j = j0 + jl;
        const bool j_ok = j < A.N_in;
        const double a = (c_ok && j_ok) ? A.M[(size_t)c * A.ld_M + j] : 0.0;
        const double b = (k_ok && j_ok) ? X[(size_t)j * A.N_P + k] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    if (!k_ok) return;
    double* out = A.out + (size_t)reg * A.N_ch * A.N_P;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int cr = c0 + 4 * r + jl;
        if (cr < A.N_ch) out[(size_t)cr * A.N_P + k] = acc[r];
    }
}

// Precision form (default): FP32 operands on v_mfma_f32_16x16x4_f32.  On gfx950
// that instruction rounds each product and partial sum to nearest, so every
// 4-deep slice of the sum over input energies starts from zero and its result
// joins an FP64 sum: no FP32 running sum over N_in terms (which drifts) and no
// bf16 dot product (whose sum v_mfma_f32_16x16x32_bf16 truncates toward zero,
// a bias that survives every split) is left in the path.
// Lane l holds M[c0 + l%16][j + l/16] and in[j + l/16][k0 + l%16]; result
// element r is row 4 (l/16) + r.
__global__ void __launch_bounds__(256) k_fold_f32(FoldGemmArgs A, const float* __restrict__ Mf)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int reg = blockIdx.z, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = blockIdx.y * 32 + (wave >> 1) * 16, k0 = blockIdx.x * 32 + (wave & 1) * 16;
    const int c = c0 + (lane & 15), k = k0 + (lane & 15), jl = lane >> 4;
    const double* X = A.in + (size_t)reg * A.N_in * A.N_P;
    const bool c_ok = c < A.N_ch, k_ok = k < A.N_P;
    const f4 z4 = {0.f, 0.f, 0.f, 0.f};
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j0 = 0; j0 < A.N_in; j0 += 4) {
        const int j = j0 + jl;
        const bool j_ok = j < A.N_in;
        const float a = (c_ok && j_ok) ? Mf[(size_t)c * A.ld_M + j] : 0.f;
        const float b = (k_ok && j_ok) ? (float)X[(size_t)j * A.N_P + k] : 0.f;
        const f4 t = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, z4, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r] += (double)t[r];
    }
    if (!k_ok) return;
    double* out = A.out + (size_t)reg * A.N_ch * A.N_P;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int cr = c0 + 4 * (lane >> 4) + r;
        if (cr < A.N_ch) out[(size_t)cr * A.N_P + k] = acc[r];
    }
}

// Precision form, selectable: split-BF16 with FP32 accumulation.  Every operand x
// (rounded to FP32) is a sum of three bf16 terms, x_h = bf16(x),
// x_m = bf16(x - x_h), x_l = bf16(x - x_h - x_m); the response matrix is split
// once on the host (Mh, Mm, Ml), the signal here.  NT = 3 keeps the products
// M_m S_h + M_h S_m + M_h S_h (the two-term split, relative error ~2^-17);
// NT = 6 adds M_l S_h + M_m S_m + M_h S_l (~2^-23), smallest terms first.
// Each instruction truncates its 32-term dot product toward zero (only the
// accumulator add rounds to nearest), so these forms read low by up to an
// FP32 ulp per instruction; the lnL carries that as a one-signed bias.
// v_mfma_f32_16x16x32_bf16: lane l holds M[c0 + l%16][j + 8 (l/16) + v] and
// in[j + 8 (l/16) + v][k0 + l%16], v = 0..7; result element r is row
// 4 (l/16) + r.
template <int NT>
__global__ void __launch_bounds__(256) k_fold_bf16(FoldGemmArgs A, const __bf16* __restrict__ Mh,
                                                   const __bf16* __restrict__ Mm,
                                                   const __bf16* __restrict__ Ml)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
    const int reg = blockIdx.z, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = blockIdx.y * 32 + (wave >> 1) * 16, k0 = blockIdx.x * 32 + (wave & 1) * 16;
    const int c = c0 + (lane & 15), k = k0 + (lane & 15), jl = 8 * (lane >> 4);
    const double* X = A.in + (size_t)reg * A.N_in * A.N_P;
    const bool c_ok = c < A.N_ch, k_ok = k < A.N_P;
    const __bf16 zero = (__bf16)0.0f;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int j0 = 0; j0 < A.N_in; j0 += 32) {
        bf8 ah, am, al, bh, bm, bl;
#pragma unroll
        for (int v = 0; v < 8; v++) {
            const int j = j0 + jl + v;
            const bool j_ok = j < A.N_in;
            const size_t mo = (size_t)c * A.ld_M + j;
            ah[v] = (c_ok && j_ok) ? Mh[mo] : zero;
            am[v] = (c_ok && j_ok) ? Mm[mo] : zero;
            al[v] = (c_ok && j_ok) ? Ml[mo] : zero;
            const float b = (k_ok && j_ok) ? (float)X[(size_t)j * A.N_P + k] : 0.f;
            const __bf16 b1 = (__bf16)b;
            const float r1 = b - (float)b1;
            const __bf16 b2 = (__bf16)r1;
            bh[v] = b1;
            bm[v] = b2;
            bl[v] = (__bf16)(r1 - (float)b2);
        }
        if (NT == 6) {
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
    }
    if (!k_ok) return;
    double* out = A.out + (size_t)reg * A.N_ch * A.N_P;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int cr = c0 + 4 * (lane >> 4) + r;
        if (cr < A.N_ch) out[(size_t)cr * A.N_P + k] = (double)acc[r];
    }
}

template __global__ void k_fold_bf16<3>(FoldGemmArgs, const __bf16*, const __bf16*, const __bf16*);
template __global__ void k_fold_bf16<6>(FoldGemmArgs, const __bf16*, const __bf16*, const __bf16*);

struct MargFn {
    const double* star;
    const double* data;
    int n;
    double SCALE, A;
    __device__ double operator()(double B) const
    {
        double x = 0.0;
        for (int j = 0; j < n; j++) {
            double c = SCALE * (star[j] + B);
            if (c > 0.0) x += data[j] * log(c) - c;
            else if (eq12(c, 0.0) && eq12(data[j], 0.0)) {
            } else return 0.0;
        }
        return exp(x - A);
    }
};

// The expected count rate of request rq in channel ch and phase bin j:
// each component's phase interpolant integrated over the bin shifted by the
// component's phase shift (wrapped; a full cycle taken whole), negative bin
// integrals dropped unless allowed, the total clipped at zero
// (xpsi/tools/compute_expected_counts.pyx:66-197).
__device__ double bin_rate(const MarginalArgs& A, int rq, int ch, int j, const double* tabs)
{
    // the request's signals: components one after another, each [N_ch][its phases]
    int P_tot = A.n_comp * A.N_P;
    if (A.comp_np) {
        P_tot = 0;
        for (int c = 0; c < A.n_comp; c++) P_tot += A.comp_np[c];
    }
    const double* sig = A.signals + (size_t)rq * A.N_ch * P_tot;
    int off = 0;   // phases of the components before c
    double st = 0.0;
    for (int c = 0; c < A.n_comp; c++) {
        const int np = A.comp_np ? A.comp_np[c] : A.N_P;
        const double* pulse = sig + (size_t)A.N_ch * off + (size_t)ch * np;
        const double* cph = A.comp_np ? A.comp_phases + off : A.comp_phases;
        const double shift = A.shifts[rq * A.n_comp + c];
        const bool neg_ok = (A.allow_negative >> c) & 1u;
        if (np > 1) {
            double pa = A.phases[j] + shift, pb = A.phases[j + 1] + shift;
            if (eq12(pb - pa, 1.0)) { pa = 0.0; pb = 1.0; }
            else { pa -= floor(pa); pb -= floor(pb); }
            // the periodic cubic spline's tables of this component (kind 2)
            const double* TB = tabs ? tabs + 3 * off : nullptr;
            const double* TC = tabs ? TB + np : nullptr;
            const double* TD = tabs ? TC + np : nullptr;
            const int kind = A.interp;
            if (pa < pb) {
                double v = phase_integ(kind, cph, pulse, np, TB, TC, TD, pa, pb);
                if (v > 0.0 || neg_ok) st += v;
            } else {
                double v = phase_integ(kind, cph, pulse, np, TB, TC, TD, pa, 1.0);
                if (v > 0.0 || neg_ok) st += v;
                v = phase_integ(kind, cph, pulse, np, TB, TC, TD, 0.0, pb);
                if (v > 0.0 || neg_ok) st += v;
            }
        } else if (j == 0) {
            st = pulse[0];
        }
        off += np;
    }
    return st < 0.0 ? 0.0 : st;
}

// The periodic cubic spline needs each component's whole row solved first:
// lane 0 writes the (b, c, d) tables of request rq's channel ch, component
// after component, into tabs (3 x the components' phase counts, then one
// component's count of work).  Returns tabs, or nullptr for the local
// interpolants.  The caller synchronises the block (one wavefront).
__device__ const double* channel_tables(const MarginalArgs& A, int rq, int ch, double* tabs)
{
    if (A.interp != 2) return nullptr;
    if (threadIdx.x == 0) {
        int P_tot = A.n_comp * A.N_P;
        if (A.comp_np) {
            P_tot = 0;
            for (int c = 0; c < A.n_comp; c++) P_tot += A.comp_np[c];
        }
        const double* sig = A.signals + (size_t)rq * A.N_ch * P_tot;
        double* z = tabs + 3 * P_tot;
        int off = 0;
        for (int c = 0; c < A.n_comp; c++) {
            const int np = A.comp_np ? A.comp_np[c] : A.N_P;
            const double* pulse = sig + (size_t)A.N_ch * off + (size_t)ch * np;
            const double* cph = A.comp_phases + (A.comp_np ? off : 0);
            double* TB = tabs + 3 * off;
            if (np >= 4) cspline_periodic(cph, pulse, np, TB, TB + np, TB + 2 * np, z);
            off += np;
        }
    }
    return tabs;
}

// wave per (channel, request)
__global__ void __launch_bounds__(64) k_marginal(MarginalArgs A)
{
    const int ch = blockIdx.x, rq = blockIdx.y, lane = threadIdx.x;
    const int N_b = A.N_b;
    extern __shared__ double sm[];
    double* star = sm;              // N_b
    double* data = star + N_b;      // N_b
    const double n = (double)N_b, T = A.exposure, SCALE = T / n;
    const double* counts = A.counts + (size_t)ch * N_b;
    const double* tabs = channel_tables(A, rq, ch, data + N_b);
    __syncthreads();
    // expected count rate from each region, phase-binned (shifted, wrapped)
    for (int j = lane; j < N_b; j += 64) {
        double st = bin_rate(A, rq, ch, j, tabs);
        if (A.background) st += A.background[(size_t)ch * N_b + j];
        st *= n;
        star[j] = st;
        data[j] = counts[j];
    }
    __syncthreads();
    double av_S = 0.0, av_D = 0.0;
    for (int j = lane; j < N_b; j += 64) { av_S += star[j]; av_D += data[j]; }
    av_S = wave_sum(av_S);
    av_D = wave_sum(av_D);
    const size_t ro = (size_t)rq * A.N_ch + ch;
    int status = 0;
    double ll = 0.0, B = 0.0;
    if (!(A.slim < 0.0)) {
        double limit = av_S * SCALE - A.slim * sqrt(av_S * SCALE) - av_D;
        if (limit > 0.0) status = 1;
    }
    const double* sup = A.support + 2 * ch;
    if (!status) {
        av_S /= n;
        av_D /= T;
        if (eq12(av_D, 0.0) && eq12(av_S, 0.0)) {
            double lower = 0.0;
            if (lower < sup[0]) lower = sup[0];
            double upper = 10.0 / T;
            if (upper > sup[1] && sup[1] > 0.0) upper = sup[1];
            B = 0.0;
            ll = log((exp(-1.0 * lower * T) - exp(-1.0 * upper * T)) / T);
        } else {
            double B_min = 0.0;
            B = av_D - av_S;
            if (B <= B_min) {
                double mc = -1.0;
                for (int j = 0; j < N_b; j++) if (eq12(star[j], 0.0)) { mc = -2.0; break; }
                if (eq12(mc, -2.0)) {
                    for (int j = 0; j < N_b; j++)
                        if ((eq12(mc, -2.0) && data[j] > 0.0) || (0.0 < data[j] && data[j] < mc)) mc = data[j];
                }
                if (!eq12(mc, -1.0)) {
                    if (eq12(mc, -2.0)) { B = 0.0; B_min = 0.0; }
                    else { B = 0.01 * mc / SCALE; B_min = 0.1 * B; }
                } else B = B_min;
            }
            auto newton = [&](double Bv, double& stdv) {
                double x = 0.0, y = 0.0;
                for (int j = lane; j < N_b; j += 64) {
                    y += data[j] / (star[j] + Bv);
                    x += data[j] / pow(star[j] + Bv, 2.0);
                }
                x = wave_sum(x);
                y = wave_sum(y);
                y = 2.0 * T - 2.0 * y;
                x *= 2.0;
                stdv = sqrt(2.0 / x);
                return -1.0 * y / x;
            };
            double stdv;
            double dB = newton(B, stdv);
            int counter = 0, iters = 0;
            // the reference iterates without a cap; 10000 steps bound pathological inputs
            while (fabs(dB) > A.epsilon * stdv && counter < 2 && iters++ < 10000) {
                B += dB;
                if (B < B_min) {
                    if (B_min > 0.0) counter += 1; else counter = 2;
                    B = B_min;
                }
                dB = newton(B, stdv);
            }
            double s = 0.0;
            for (int j = lane; j < N_b; j += 64) s += data[j] / pow(star[j] + B, 2.0);
            s = wave_sum(s);
            s = s > 0.0 ? sqrt(1.0 / s) : 1e90;
            double lower = B - A.sigmas * s, upper = B + A.sigmas * s;
            if (lower < B_min) lower = B_min;
            double Bi = B;
            if (lower < sup[0]) {
                lower = sup[0];
                if (upper < sup[0] && sup[1] > 0.0) { upper = sup[1]; Bi = sup[0]; }
                else if (upper < sup[0]) { upper = sup[0] + A.sigmas * s; Bi = sup[0]; }
            }
            if (upper > sup[1] && sup[1] > 0.0) {
                upper = sup[1];
                if (lower > sup[1]) { lower = sup[0]; Bi = sup[1]; }
            }
            if (Bi < lower) Bi = lower;
            else if (Bi > upper) Bi = upper;
            double Aval = 0.0;
            for (int j = lane; j < N_b; j += 64) {
                double c = SCALE * (star[j] + Bi);
                if (c > 0.0) Aval += data[j] * log(c) - c;
                else if (eq12(c, 0.0) && eq12(data[j], 0.0)) {
                } else Aval += A.llzero;
            }
            Aval = wave_sum(Aval);
            MargFn f{star, data, N_b, SCALE, Aval};
            double result, abserr;
            double er = A.epsrel < MARG_EPSREL ? A.epsrel : MARG_EPSREL;
            wave_qag61(f, lower, upper, A.epsabs, er, QAG_LIMIT, result, abserr);
            if (result > 0.0) ll = log(result) + Aval + A.precomp[ch];
            else status = 2;
        }
    }
    if (lane == 0) {
        A.ll[ro] = ll;
        A.status[ro] = status;
        if (!status) {
            A.B_ML[ro] = B * T;
            double Bs = B;
            if (B < sup[0]) Bs = sup[0];
            else if (B > sup[1] && sup[1] > 0.0) Bs = sup[1];
            A.B_ML_supp[ro] = Bs * T;
        }
    }
    if (!status && A.STAR) {
        double Bs = B;
        if (B < sup[0]) Bs = sup[0];
        else if (B > sup[1] && sup[1] > 0.0) Bs = sup[1];
        for (int j = lane; j < N_b; j += 64) A.STAR[ro * N_b + j] = SCALE * (star[j] + Bs);
    }
}

__global__ void __launch_bounds__(256) k_fill(double* p, size_t n, double v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(256) k_over(double* out, const double* x, size_t n, double a)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a / x[i];
}

// rows [rows][cols] of a per-row (by_row) or per-column value: out[i][j] = v[i] or v[j]
__global__ void __launch_bounds__(256) k_spread(double* out, const double* v, int rows, int cols, int by_row)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < (size_t)rows * cols) out[q] = by_row ? v[q / cols] : v[q % cols];
}

__global__ void __launch_bounds__(256) k_fill_int(int* p, size_t n, int v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ordered sum over channels; a slim or failed channel makes the request
// llzero-class (the reference's random draw near llzero)
__global__ void __launch_bounds__(64) k_lnl_reduce(MarginalArgs A, const Member* __restrict__ members,
                                                   int members_per_request, double* __restrict__ lnl)
{
    const int rq = blockIdx.x;
    if (threadIdx.x != 0) return;
    double s = 0.0;
    int bad = 0;
    for (int c = 0; c < A.N_ch; c++) {
        size_t ro = (size_t)rq * A.N_ch + c;
        if (A.status[ro]) { bad = 1; break; }
        s += A.ll[ro];
    }
    if (A.request_status && A.request_status[rq]) bad = 1;
    if (members)
        for (int q = 0; q < members_per_request; q++) {
            const Member& m = members[(size_t)rq * members_per_request + q];
            if (m.active && m.status) bad = 1;
        }
    if (bad || s <= A.llzero) {
        // splitmix64 of (launch seed, request): a fresh draw per launch and request
        const unsigned long long sd = A.seed_ptr ? *A.seed_ptr : A.seed;
        unsigned long long h = sd * 0x100000001B3ULL + (unsigned long long)(rq + 1) * 0x9E3779B97F4A7C15ULL;
        h ^= h >> 30;
        h *= 0xBF58476D1CE4E5B9ULL;
        h ^= h >> 27;
        h *= 0x94D049BB133111EBULL;
        h ^= h >> 31;
        double u = (double)(h >> 11) / 9007199254740992.0;
        s = A.llzero * (0.1 + 0.9 * u);
    }
    lnl[rq] = s;
}

// Poisson likelihood given a background (the expected counts are the star's
// plus the background's count rate times the exposure), wave per (channel,
// request): lanes form each bin's expected counts and its term, then one lane
// sums the channel in bin order (from the channel's -sum ln d! when given) and
// counts the bins with data but no expected counts.  ll [request][channel],
// status [request][channel] = such bins; STAR holds the expected counts.
// (xpsi/likelihoods/_poisson_likelihood_given_background.pyx:14-113)
__global__ void __launch_bounds__(64) k_poisson_bg(MarginalArgs A)
{
    const int ch = blockIdx.x, rq = blockIdx.y, lane = threadIdx.x;
    const int N_b = A.N_b;
    extern __shared__ double sm[];
    double* ex = sm;                // N_b expected counts
    double* tm = ex + N_b;          // N_b terms d log(ex) - ex
    const double* counts = A.counts + (size_t)ch * N_b;
    const size_t ro = (size_t)rq * A.N_ch + ch;
    const double* tabs = channel_tables(A, rq, ch, tm + N_b);
    __syncthreads();
    int bad = 0;
    for (int j = lane; j < N_b; j += 64) {
        const double e = (bin_rate(A, rq, ch, j, tabs) + A.background[(size_t)ch * N_b + j]) * A.exposure;
        const double d = counts[j];
        ex[j] = e;
        if (A.STAR) A.STAR[ro * N_b + j] = e;
        if (e > 0.0) tm[j] = d * log(e);
        else {
            tm[j] = 0.0;
            if (!(d == 0.0 && e == 0.0)) bad++;
        }
    }
    bad = (int)wave_sum((double)bad);
    __syncthreads();
    if (lane == 0) {
        double s = A.precomp ? A.precomp[ch] : 0.0;
        for (int j = 0; j < N_b; j++) {
            if (ex[j] > 0.0) {
                s -= ex[j];
                s += tm[j];
            }
        }
        A.ll[ro] = s;
        A.status[ro] = bad;
    }
}

// ordered sum over channels; every bin with data but no expected counts adds
// its own draw llzero (0.1 + 0.9 U) (the reference's -1e90 (0.1 + 0.9 U))
__global__ void __launch_bounds__(64) k_poisson_reduce(MarginalArgs A, double* __restrict__ lnl)
{
    const int rq = blockIdx.x;
    if (threadIdx.x != 0) return;
    const unsigned long long sd = A.seed_ptr ? *A.seed_ptr : A.seed;
    double s = 0.0;
    unsigned long long k = 0;
    for (int c = 0; c < A.N_ch; c++) {
        const size_t ro = (size_t)rq * A.N_ch + c;
        s += A.ll[ro];
        for (int b = 0; b < A.status[ro]; b++, k++) {
            unsigned long long h = sd * 0x100000001B3ULL + (unsigned long long)(rq + 1) * 0x9E3779B97F4A7C15ULL +
                                   k * 0xD1B54A32D192ED03ULL;
            h ^= h >> 30;
            h *= 0xBF58476D1CE4E5B9ULL;
            h ^= h >> 27;
            h *= 0x94D049BB133111EBULL;
            h ^= h >> 31;
            s += A.llzero * (0.1 + 0.9 * ((double)(h >> 11) / 9007199254740992.0));
        }
    }
    lnl[rq] = s;
}


// Gaussian likelihood given a background
// (xpsi/likelihoods/_gaussian_likelihood_given_background_IQU.pyx:14-91):
// wave per channel; each bin's expected counts (the star's plus the
// background's count rate, times the exposure, as the Poisson statistic forms
// them) into STAR, and its term (EXPEC - d)^2 / (2 s^2) + ln(2 pi s^2) / 2,
// s the bin's error, into tm [N_ch][N_b]
__global__ void __launch_bounds__(64) k_gauss_bg(MarginalArgs A, const double* __restrict__ errors,
                                                 double* __restrict__ tm)
{
    const int ch = blockIdx.x, lane = threadIdx.x;
    const int N_b = A.N_b;
    extern __shared__ double sm[];
    const double* tabs = channel_tables(A, 0, ch, sm);
    __syncthreads();
    for (int j = lane; j < N_b; j += 64) {
        const size_t q = (size_t)ch * N_b + j;
        const double e = (bin_rate(A, 0, ch, j, tabs) + A.background[q]) * A.exposure;
        if (A.STAR) A.STAR[q] = e;
        const double s2 = errors[q] * errors[q], r = e - A.counts[q];
        tm[q] = r * r / (2.0 * s2) + 0.5 * log(2.0 * PI * s2);
    }
}

// xpsi/likelihoods/_gaussian_likelihood_QnUn.pyx:14-44: the term of data row
// i and bin j against the one model row star [N_b]
__global__ void __launch_bounds__(256) k_gauss_qnun(int n_rows, int N_b, const double* __restrict__ counts,
                                                    const double* __restrict__ errors,
                                                    const double* __restrict__ star, double* __restrict__ tm)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)n_rows * N_b) return;
    const double s2 = errors[q] * errors[q], r = star[q % N_b] - counts[q];
    tm[q] = r * r / (2.0 * s2) + 0.5 * log(2.0 * PI * s2);
}

// the reference's running sum, LOGLIKE -= term, row by row and bin by bin
__global__ void __launch_bounds__(64) k_gauss_reduce(const double* __restrict__ tm, size_t n,
                                                     double* __restrict__ lnl)
{
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (size_t q = 0; q < n; q++) s -= tm[q];
    *lnl = s;
}


// xpsi/tools/phase_integrator.pyx:23-121: block per row; each row's phase
// interpolant integrated over the shifted intervals (the wrap exact at
// b - a == 1), a piece kept where positive or allowed, the sum not clipped,
// times the exposure, into out [rows][N_b].  A.signals [rows][N_P] on the grid
// A.comp_phases, one component.
__global__ void __launch_bounds__(64) k_phase_integrate(MarginalArgs A, double shift, double* __restrict__ out)
{
    const int row = blockIdx.x;
    extern __shared__ double sm[];
    const double* tabs = channel_tables(A, 0, row, sm);
    __syncthreads();
    const int np = A.N_P;
    const double* pulse = A.signals + (size_t)row * np;
    const double* TB = tabs;
    const double* TC = tabs ? TB + np : nullptr;
    const double* TD = tabs ? TC + np : nullptr;
    const bool neg_ok = A.allow_negative & 1u;
    for (int j = threadIdx.x; j < A.N_b; j += 64) {
        double a = A.phases[j] + shift, b = A.phases[j + 1] + shift, v = 0.0;
        if (b - a == 1.0) { a = 0.0; b = 1.0; }
        else { a -= floor(a); b -= floor(b); }
        if (a < b) {
            const double w = phase_integ(A.interp, A.comp_phases, pulse, np, TB, TC, TD, a, b);
            if (w > 0.0 || neg_ok) v = w;
        } else {
            const double w1 = phase_integ(A.interp, A.comp_phases, pulse, np, TB, TC, TD, a, 1.0);
            if (w1 > 0.0 || neg_ok) v = w1;
            const double w2 = phase_integ(A.interp, A.comp_phases, pulse, np, TB, TC, TD, 0.0, b);
            if (w2 > 0.0 || neg_ok) v += w2;
        }
        out[(size_t)row * A.N_b + j] = v * A.exposure;
    }
}


// ---------------------------------------------------------------- synthesis
// Poisson realisations of the expected counts
// (xpsi/tools/synthesise.pyx:40-276).  Random numbers: splitmix64 (Steele,
// Lea and Flood 2014) of a per-bin stream (seed, bin index); Poisson variates
// by sequential inversion below mean 30 and by Hormann's transformed
// rejection with squeeze (PTRS, Hormann 1993) above.
struct SplitMix {
    unsigned long long s;
    __device__ unsigned long long next()
    {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    __device__ double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

__device__ double poisson_draw(double mu, SplitMix& r)
{
    if (!(mu > 0.0)) return 0.0;
    if (mu < 30.0) {
        // inversion: walk the cumulative distribution from zero
        double p = exp(-mu), c = p, u = r.uniform();
        double k = 0.0;
        while (u > c && k < 1000.0) {
            k += 1.0;
            p *= mu / k;
            c += p;
        }
        return k;
    }
    // PTRS: a hat from the transformed uniform U, the squeeze accepting most
    // draws at once, the exact log-density test for the rest
    const double smu = sqrt(mu), lmu = log(mu);
    const double b = 0.931 + 2.53 * smu, a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4), v_r = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < 1000; it++) {
        const double U = r.uniform() - 0.5, V = r.uniform();
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + mu + 0.43);
        if (us >= 0.07 && V <= v_r) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -mu + k * lmu - lgamma(k + 1.0)) return k;
    }
    return floor(mu + 0.5);   // not reached in practice (acceptance above 0.9 per try)
}

// the clipped star count rate per (channel, bin) into out [N_ch][N_b]
__global__ void __launch_bounds__(64) k_star_rate(MarginalArgs A, double* __restrict__ out)
{
    const int ch = blockIdx.x;
    extern __shared__ double sm[];
    const double* tabs = channel_tables(A, 0, ch, sm);
    __syncthreads();
    for (int j = threadIdx.x; j < A.N_b; j += 64) out[(size_t)ch * A.N_b + j] = bin_rate(A, 0, ch, j, tabs);
}

// ordered totals (the reference's sequential loops): scal[0] background sum,
// scal[1] sum of the star counts rate x mul
__global__ void __launch_bounds__(64) k_synth_totals(const double* __restrict__ rate,
                                                     const double* __restrict__ bg, int n, double mul, double* scal)
{
    if (threadIdx.x != 0) return;
    double b = 0.0, st = 0.0;
    for (int q = 0; q < n; q++) { st += rate[q] * mul; b += bg[q]; }
    scal[0] = b;
    scal[1] = st;
}

// expected counts and one Poisson draw per (channel, bin).  mode 0
// (synthesise_exposure): EXPEC = (rate + bg * sb / T) * T with
// sb = n_bg / sum(bg); mode 1 (synthesise_given_total_count_number):
// EXPEC = (rate * 1e5) * ss + bg * sb, ss = n_star / sum(rate * 1e5).
// scal[2], scal[3] receive sb and ss.
__global__ void __launch_bounds__(256) k_synth_draw(int mode, const double* __restrict__ rate,
                                                    const double* __restrict__ bg, int n, double T, double n_star,
                                                    double n_bg, double* scal, unsigned long long seed,
                                                    double* __restrict__ expec, double* __restrict__ synth)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const double B = scal[0];
    const double sb = B == 0.0 ? 0.0 : n_bg / B;
    double ss = 0.0;
    if (mode == 1) ss = n_star / scal[1];
    if (q == 0) { scal[2] = sb; scal[3] = ss; }
    if (q >= n) return;
    double e;
    if (mode == 0) {
        e = rate[q];
        e += bg[q] * sb / T;
        e *= T;
    } else {
        e = (rate[q] * 1.0e5) * ss + bg[q] * sb;
    }
    expec[q] = e;
    SplitMix r{seed ^ (0xD1B54A32D192ED03ULL * (unsigned long long)(q + 1))};
    synth[q] = poisson_draw(e, r);
}

}  // namespace xh
