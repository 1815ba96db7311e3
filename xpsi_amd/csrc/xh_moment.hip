// Moment form of the cell x phase x energy fold, on the matrix cores.
//
// Reference behaviour: the energy x cell x phase spline loop of
// xpsi/cellmesh/integrator_for_azimuthal_invariance.pyx:565-596 and the
// blackbody intensity of xpsi/surface_radiation_field/hot_BB.pyx:54-98.
//
// For one parallel and image order the reference evaluates, for every energy
// k, the periodic Akima interpolant s_k of the profile at x = phase_p + phi_j
// for every radiating cell j and model phase p, keeps values > 0 and adds
// cellArea_j * s_k(x).  The interval i(j, p) holding x and the offset
// t = x - PHASE[i] do not depend on the energy.  Where s_k is positive over
// the whole interval the clipping is inactive and
//
//     sum_j A_j s_k(x_jp) = sum_i sum_m coef_k[i][m] * MOM[i][m][p],
//     MOM[i][m][p]       = sum over cells j with i(j, p) = i of A_j t^m,
//
// the product of an [energy x (interval, power)] matrix with an
// [(interval, power) x phase] matrix: one 16x16x4 MFMA per interval and
// 16 x 16 tile.  Intervals where s_k changes sign are evaluated point by point
// with the clipping, exactly as the reference does; intervals where s_k <= 0
// throughout contribute nothing.
//
// One block of four waves owns 16 energies x 64 phases of one member and
// walks the parallels of its ring block.  Per parallel and image: profile
// (intensity x geometry) into LDS, secant slopes, Akima coefficients with a
// sign class per (interval, energy); then per 16-phase tile, moments from the
// (cell, phase) queries by LDS atomics, direct evaluation of sign-changing
// pieces, and the MFMA sweep, intervals dealt round-robin to the waves.  Each
// parallel's tile is added into FP64 accumulators, so the float form sums
// parallels in FP64.
//
// R = double: FP64 throughout (v_mfma_f64_16x16x4f64).
// R = float: the precision form.  Blackbody intensity, spline coefficients,
// moments and products in FP32 (v_mfma_f32_16x16x4f32); the intensity's
// arguments (E' = E/Z and E'/kT) and the phase offsets t are formed in FP64
// and rounded once, the per-parallel partial sums are accumulated in FP64 and
// the blackbody norm is applied in FP64 at the end.
#include "xh_common.hpp"
#include "xh_types.hpp"
#include "xh_kernels.hpp"
#include "xh_atm.hpp"

namespace xh {

template <typename R> struct Mfma;

// lane l: A[l % 16][l / 16], B[l / 16][l % 16]; result element r of lane l is
// D[4 r + l / 16][l % 16]
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * r + (l >> 4); }
};

// lane l: A[l % 16][l / 16], B[l / 16][l % 16]; result element r of lane l is
// D[4 (l / 16) + r][l % 16]
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

template <typename R> struct Eps;
template <> struct Eps<double> { static constexpr double v = DBL_EPSILON; };
template <> struct Eps<float> { static constexpr float v = FLT_EPSILON; };

// Sign class of f(t) = a + t (b + t (c + t d)) on [0, h] from its cubic
// Bernstein coefficients (the curve lies in their convex hull):
// 1 = every Bernstein coefficient above the evaluation-error margin, so f > 0
//     at every point and the clipping cannot act;
// 0 = identically zero, or every coefficient below minus the margin (f < 0);
// 2 = otherwise, evaluated point by point with the clipping (conservative: a
//     positive piece whose hull dips below the margin is evaluated exactly).
template <typename R>
__device__ __forceinline__ int cubic_class(R a, R b, R c, R d, R h)
{
    if (a == (R)0 && b == (R)0 && c == (R)0 && d == (R)0) return 0;
    const R bh = b * h, ch = c * h * h, dh = d * h * h * h;
    const R scale = fabs(a) + fabs(bh) + fabs(ch) + fabs(dh);
    const R tol = (R)64 * Eps<R>::v * scale;
    const R b0 = a, b1 = a + bh / (R)3, b2 = a + (R)2 * bh / (R)3 + ch / (R)3, b3 = a + bh + ch + dh;
    const R lo = fmin(fmin(b0, b1), fmin(b2, b3)), hi = fmax(fmax(b0, b1), fmax(b2, b3));
    if (lo > tol) return 1;
    if (hi < -tol) return 0;
    return 2;
}

// One block of four waves owns 16 energies x TPB * 16 phases of one member.
// Per parallel and image it builds the profile, Akima coefficients and sign
// classes once, then walks its phase tiles: moments of the tile's queries
// (one 16-phase moment buffer in LDS), direct sums of the sign-changing
// pieces, and the MFMA sweep with those pieces masked out of the A operand.
template <typename R, int ATM, int TPB>
__global__ void __launch_bounds__(256) k_moment_fold(const Member* __restrict__ members, Layout L, FoldArgs F)
{
    constexpr int PCB = 16 * TPB;           // phases per block
    extern __shared__ __align__(16) double smd[];
    const int mi = blockIdx.z, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const Member m = members[mi];
    if (!m.active || m.status) return;
    const int S = m.S, N_L = L.N_L, N_P = L.N_P, N_E = L.N_E, NI = N_L - 1;
    const int n_ec = (N_E + 15) >> 4;
    const int e0 = (blockIdx.x % n_ec) * 16, p0 = (blockIdx.x / n_ec) * PCB;
    const int r0 = blockIdx.y * F.ring_block;
    if (r0 >= S) return;
    const int r1 = min(S, r0 + F.ring_block);
    const int ntile = min(TPB, (N_P - p0 + 15) >> 4);

    double* PH = smd;                       // [N_L] leaf phases
    double* Zi = PH + N_L;                  // [N_L] 1 / redshift factor
    double* As = Zi + N_L;                  // [N_L] mu * eta
    double* Gs = As + N_L;                  // [N_L] geometric factor
    double* PV = Gs + N_L;                  // [N_L] previous queried leaf
    double* cphi = PV + N_L;                // [S_max]
    double* carea = cphi + L.S_max;         // [S_max], NaN = not radiating
    double* ST = carea + L.S_max;           // [8] previous table query; area slots
    double* D64 = ST + 8;                   // [16][PCB] direct sums, FP64
    // [16][PCB] direct sums of one parallel (the FP64 form sums into D64 itself)
    R* DIR = sizeof(R) == 8 ? (R*)D64 : (R*)(D64 + 16 * PCB);
    unsigned* MIX = (unsigned*)(DIR + 16 * PCB);  // [N_L (even)] energy bits of sign-changing pieces
    R* CO = (R*)(MIX + ((N_L + 1) & ~1));   // [N_L][4][16] coefficient m of energy k
    R* MOM = CO + N_L * 64;                 // [NI][4][16] moment m at phase p of one tile
    R* SL = MOM;                            // [NI][16] secant slopes (before the moments)

    // FP32 form: cell areas enter the moments relative to the ring's largest
    // one.  The mesh is equal-area, so interior cells become exactly 1 and
    // their common area multiplies each parallel's sums in FP64 (one FP32
    // rounding of it would scale every bin alike).  Two slots alternate with
    // the parallel's parity.
    unsigned long long* ABITS = (unsigned long long*)(ST + 6);
    if (tid < 2) ABITS[tid] = 0ull;
    for (int q = tid; q < 16 * PCB; q += 256) { DIR[q] = (R)0; D64[q] = 0.0; }   // (one array for FP64)
    double tot[TPB][4];
#pragma unroll
    for (int t = 0; t < TPB; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) tot[t][r] = 0.0;
    const size_t cbase = (size_t)mi * L.S_max * L.S_max;
    const int p = tid & 15, part = tid >> 4;
    const int ja = (part * S) >> 4, jb = ((part + 1) * S) >> 4;

    for (int i = r0; i < r1; i++) {
        const size_t ro = (size_t)mi * L.S_max + i;
        __syncthreads();   // the previous parallel is done with the cell arrays
        if (tid == 0) ABITS[(i + 1) & 1] = 0ull;
        for (int j = tid; j < S; j += 256) {
            size_t c = cbase + (size_t)i * S + j;
            double a = F.cellArea[c];
            bool rad = F.radiates ? (F.radiates[c] == 1) : (a > 0.0);
            carea[j] = rad ? a : __builtin_nan("");
            cphi[j] = F.phi[(size_t)mi * L.S_max * (F.phi_row_stride ? L.S_max : 1) +
                            (size_t)i * F.phi_row_stride + j];
            if (sizeof(R) == 4 && rad && a > 0.0) atomicMax(&ABITS[i & 1], (unsigned long long)__double_as_longlong(a));
        }
        const double kT = K_B_OVER_KEV * pow(10.0, F.logT[ro]);
        const double kT_else = K_B_OVER_KEV * pow(10.0, m.else_logT);
        for (int I = 0; I < L.IO_max; I++) {
            const size_t vix = ro * L.IO_max + I;
            if (!F.valid[vix]) break;
            __syncthreads();   // the previous image's MFMA sweep is done with CO/MOM
            const double* g = F.geo + vix * 5 * N_L;
            for (int q = tid; q < N_L; q += 256) {
                PH[q] = g[q]; Zi[q] = 1.0 / g[N_L + q]; As[q] = g[2 * N_L + q]; Gs[q] = g[3 * N_L + q];
                PV[q] = g[4 * N_L + q];
                MIX[q] = 0u;
            }
            if (ATM == 2 && tid == 0) prev_query_state(F, L, mi, i, I, ST);
            const double a_ref = (sizeof(R) == 4 && ABITS[i & 1]) ? __longlong_as_double((long long)ABITS[i & 1]) : 1.0;
            if (sizeof(R) == 4 && I == 0)
                for (int j = tid; j < S; j += 256) carea[j] = carea[j] / a_ref;
            __syncthreads();
            // profile: specific intensity x geometry, zero where the leaf is hidden
            for (int q = tid; q < N_L * 16; q += 256) {
                const int mm = q >> 4, k = q & 15, e = e0 + k;
                R v = (R)0;
                const double geom = Gs[mm];
                if (geom != 0.0 && e < N_E && PV[mm] > -1.5) {
                    const double Ep = F.energies[e] * Zi[mm];
                    if (ATM == 2) {
                        double vec[4], prev[4];
                        int has_prev = 1;
                        vec[0] = F.logT[ro]; vec[1] = F.logg[ro]; vec[2] = As[mm];
                        vec[3] = log10(Ep / kT);
                        if (e > 0) {
                            prev[0] = vec[0]; prev[1] = vec[1]; prev[2] = vec[2];
                            prev[3] = log10((F.energies[e - 1] * Zi[mm]) / kT);
                        } else if (PV[mm] > -0.5) {
                            int kd = (int)PV[mm];
                            prev[0] = vec[0]; prev[1] = vec[1]; prev[2] = As[kd];
                            prev[3] = log10((F.energies[N_E - 1] * Zi[kd]) / kT);
                        } else if (ST[0] > 0.5) {
                            prev[0] = ST[1]; prev[1] = ST[2]; prev[2] = ST[3]; prev[3] = ST[4];
                        } else has_prev = 0;
                        double corr = 0.0;
                        if (m.correct) corr = Ep * Ep * Ep / (exp(Ep / kT_else) - 1.0) * BB_NORM;
                        v = (R)((num4d_I(F, vec, prev, has_prev) * NUM4D_NORM - corr) * geom);
                    } else if (sizeof(R) == 8) {
                        double I_E = Ep * Ep * Ep / (exp(Ep / kT) - 1.0);
                        double corr = 0.0;
                        if (m.correct) corr = Ep * Ep * Ep / (exp(Ep / kT_else) - 1.0) * BB_NORM;
                        v = (R)((I_E * BB_NORM - corr) * geom);
                    } else {
                        // FP32 intensity from arguments rounded once each: a shared
                        // rounding (of E, kT or the norm) would bias every bin alike,
                        // so E' and E'/kT are formed in FP64 and the norm is applied
                        // in FP64 after the sums
                        const float Ef = (float)Ep;
                        const float Ep3 = Ef * Ef * Ef;
                        float I_E = Ep3 / (expf((float)(Ep / kT)) - 1.0f);
                        if (m.correct) I_E -= Ep3 / (expf((float)(Ep / kT_else)) - 1.0f);
                        v = (R)(I_E * (float)geom);
                    }
                }
                CO[mm * 64 + k] = v;
            }
            __syncthreads();
            if (tid < 16 && PV[N_L - 1] < -1.5) CO[(N_L - 1) * 64 + tid] = CO[tid];
            __syncthreads();
            // secant slopes
            for (int q = tid; q < NI * 16; q += 256) {
                const int mm = q >> 4, k = q & 15;
                SL[q] = (CO[(mm + 1) * 64 + k] - CO[mm * 64 + k]) / (R)(PH[mm + 1] - PH[mm]);
            }
            __syncthreads();
            // periodic Akima coefficients and the sign class of each piece
            for (int q = tid; q < NI * 16; q += 256) {
                const int mm = q >> 4, k = q & 15;
                auto sl = [&](int kk) {
                    kk = kk < 0 ? kk + NI : (kk >= NI ? kk - NI : kk);
                    return SL[kk * 16 + k];
                };
                const R h = (R)(PH[mm + 1] - PH[mm]);
                R a = CO[mm * 64 + k], b, c, d;
                akima_piece_t<R>(sl(mm - 2), sl(mm - 1), sl(mm), sl(mm + 1), sl(mm + 2), h, b, c, d);
                // with the elsewhere correction every value is kept (:593)
                const int cls = m.correct ? 1 : cubic_class<R>(a, b, c, d, h);
                if (cls == 0) { a = (R)0; b = (R)0; c = (R)0; d = (R)0; }
                else if (cls == 2) atomicOr(&MIX[mm], 1u << k);
                CO[mm * 64 + k] = a;
                CO[mm * 64 + 16 + k] = b;
                CO[mm * 64 + 32 + k] = c;
                CO[mm * 64 + 48 + k] = d;
            }
            const double lo = PH[0], hi = PH[N_L - 1];
            for (int t = 0; t < ntile; t++) {
                __syncthreads();   // slopes / the previous tile's sweep are done with MOM
                for (int q = tid; q < NI * 64; q += 256) MOM[q] = (R)0;
                __syncthreads();
                // moments of the tile's (cell, phase) queries; sign-changing pieces directly
                const int pt = t * 16 + p, pg = p0 + pt;
                if (pg < N_P) {
                    const double ph = F.phases[pg];
                    int id = -1, cur = -1;
                    R m0 = (R)0, m1 = (R)0, m2 = (R)0, m3 = (R)0;
                    for (int j = ja; j < jb; j++) {
                        const double area = carea[j];
                        if (isnan(area)) continue;
                        double x = ph + cphi[j];
                        int guard = 0;
                        if (x > hi) { while (x > hi && guard++ < 64) x -= TWO_PI; }
                        else if (x < lo) { while (x < lo && guard++ < 64) x += TWO_PI; }
                        if (id < 0 || x < PH[id]) id = interval_of(PH, N_L, x);
                        else while (id < NI - 1 && PH[id + 1] <= x) id++;
                        const R tq = (R)(x - PH[id]), ar = (R)area;
                        if (id != cur) {
                            if (cur >= 0) {
                                R* mo = MOM + cur * 64 + p;
                                atomicAdd(mo, m0);
                                atomicAdd(mo + 16, m1);
                                atomicAdd(mo + 32, m2);
                                atomicAdd(mo + 48, m3);
                            }
                            cur = id;
                            m0 = (R)0; m1 = (R)0; m2 = (R)0; m3 = (R)0;
                        }
                        R at = ar;
                        m0 += at; at *= tq; m1 += at; at *= tq; m2 += at; at *= tq; m3 += at;
                        unsigned mx = MIX[id];
                        while (mx) {
                            const int k = __ffs(mx) - 1;
                            mx &= mx - 1u;
                            const R* co = CO + id * 64 + k;
                            const R v = co[0] + tq * (co[16] + tq * (co[32] + tq * co[48]));
                            if (v > (R)0) atomicAdd(&DIR[k * PCB + pt], ar * v);
                        }
                    }
                    if (cur >= 0) {
                        R* mo = MOM + cur * 64 + p;
                        atomicAdd(mo, m0);
                        atomicAdd(mo + 16, m1);
                        atomicAdd(mo + 32, m2);
                        atomicAdd(mo + 48, m3);
                    }
                }
                __syncthreads();
                // the sweep: sign-changing pieces masked out of the A operand
                typename Mfma<R>::v4 acc = {(R)0, (R)0, (R)0, (R)0};
                const int kb = lane & 15;
                for (int idx = wave; idx < NI; idx += 4) {
                    R a = CO[idx * 64 + lane];
                    if ((MIX[idx] >> kb) & 1u) a = (R)0;
                    acc = Mfma<R>::step(a, MOM[idx * 64 + lane], acc);
                }
#pragma unroll
                for (int tt = 0; tt < TPB; tt++)
                    if (tt == t)
#pragma unroll
                        for (int r = 0; r < 4; r++) tot[tt][r] += (double)acc[r] * a_ref;
            }
            if (sizeof(R) == 4) {
                __syncthreads();
                for (int q = tid; q < 16 * PCB; q += 256) { D64[q] += (double)DIR[q] * a_ref; DIR[q] = (R)0; }
            }
        }
    }
    double* RED = (double*)MOM;   // [4 waves][16][16]; the host sizes LDS to hold it
    for (int t = 0; t < ntile; t++) {
        __syncthreads();
#pragma unroll
        for (int tt = 0; tt < TPB; tt++)
            if (tt == t)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    RED[wave * 256 + Mfma<R>::row(lane, r) * 16 + (lane & 15)] = tot[tt][r];
        __syncthreads();
        const int k = tid >> 4, pp = t * 16 + (tid & 15), e = e0 + k, pq = p0 + pp;
        if (e < N_E && pq < N_P) {
            double s = D64[k * PCB + pp] + RED[tid] + RED[256 + tid] + RED[512 + tid] + RED[768 + tid];
            if (sizeof(R) == 4 && ATM == 1) s *= BB_NORM;
            const double E = F.energies[e];
            double* fo = F.flux + (size_t)mi * N_E * N_P + (size_t)e * N_P + pq;
            if (gridDim.y == 1) *fo = s / (E * KEV);
            else atomicAdd(fo, s * (1.0 / (E * KEV)));
        }
    }
}

#define XH_MOMENT_INST(R, TPB)                                                                 \
    template __global__ void k_moment_fold<R, 1, TPB>(const Member*, Layout, FoldArgs);        \
    template __global__ void k_moment_fold<R, 2, TPB>(const Member*, Layout, FoldArgs);
XH_MOMENT_INST(double, 1) XH_MOMENT_INST(double, 2) XH_MOMENT_INST(double, 4)
XH_MOMENT_INST(float, 1) XH_MOMENT_INST(float, 2) XH_MOMENT_INST(float, 4)

}  // namespace xh
