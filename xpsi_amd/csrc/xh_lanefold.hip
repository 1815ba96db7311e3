// Energy-lane form of the cell x phase x energy fold.
//
// Reference behaviour: the energy x cell x phase spline loop of
// xpsi/cellmesh/integrator_for_azimuthal_invariance.pyx:565-596, blackbody
// intensity xpsi/surface_radiation_field/hot_BB.pyx:54-98, the 4-D table
// hot_Num4D.pyx:248-438 and the elsewhere correction (:469-478, :593).
//
// For one parallel and image, every (cell, phase) query x = phase + phi_cell
// falls in one interval of the leaf phases with offset t; both depend on the
// cell and phase only.  Lanes therefore hold 64 energies and a wave walks the
// cells in order with its 16 phases as independent queries per cell: the
// interval search and t are wave-uniform, each lane reads the interval's four
// coefficients (one 16- or 32-byte LDS read), does the reference's Horner
// evaluation, keeps values > 0 (all values under the elsewhere correction)
// and accumulates cellArea * value in a per-phase register.
//
// One block of four waves owns 64 energies x 4 NQ phases of one member and walks
// the parallels of its ring block.  Per parallel and image: the profile
// (intensity x geometry) for the block's energies, then the periodic Akima
// coefficients, each wave taking a quarter of the intervals with a sliding
// window of secant slopes per lane, then the fold.  Each parallel's per-phase
// sums are added into FP64 accumulators.
//
// R = double: FP64 throughout, the reference's arithmetic per evaluation.
// R = float: the precision form, as in the moment form: intensity arguments
// formed in FP64 and rounded once, cell areas relative to the ring's largest
// (interior cells exactly 1), the blackbody norm applied in FP64 at the end.
#include "xh_common.hpp"
#include "xh_types.hpp"
#include "xh_kernels.hpp"
#include "xh_atm.hpp"

namespace xh {

template <typename R> struct alignas(4 * sizeof(R)) Co4 { R a, b, c, d; };

template <typename R, int ATM, int NQ>
__global__ void __launch_bounds__(256) k_lane_fold(const Member* __restrict__ members, Layout L, FoldArgs F)
{
    constexpr int PB = 4 * NQ;              // phases per block (NQ per wave)
    extern __shared__ __align__(16) double smd[];
    const int mi = blockIdx.z, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const Member m = members[mi];
    if (!m.active || m.status) return;
    const int S = m.S, N_L = L.N_L, N_P = L.N_P, N_E = L.N_E, NI = N_L - 1;
    const int n_eg = (N_E + 63) >> 6;
    const int e0 = (blockIdx.x % n_eg) * 64, pw = (blockIdx.x / n_eg) * PB + wave * NQ;
    const int r0 = blockIdx.y * F.ring_block;
    if (r0 >= S) return;
    const int r1 = min(S, r0 + F.ring_block);

    double* PH = smd;                       // [N_L] leaf phases
    double* Zi = PH + N_L;                  // [N_L] 1 / redshift factor
    double* As = Zi + N_L;                  // [N_L] mu * eta
    double* Gs = As + N_L;                  // [N_L] geometric factor
    double* PV = Gs + N_L;                  // [N_L] previous queried leaf
    double* cphi = PV + N_L;                // [S_max]
    double* carea = cphi + L.S_max;         // [S_max], NaN = not radiating
    double* ST = carea + L.S_max;           // [8] previous table query; area slots
    R* YL = (R*)(smd + ((5 * N_L + 2 * L.S_max + 8 + 3) & ~3));   // [64] profile at the last leaf (32-byte aligned)
    Co4<R>* CO = (Co4<R>*)(YL + 64);        // [NI][64] piece coefficients; .a = profile

    unsigned long long* ABITS = (unsigned long long*)(ST + 6);
    if (tid < 2) ABITS[tid] = 0ull;
    double tot[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) tot[q] = 0.0;
    const size_t cbase = (size_t)mi * L.S_max * L.S_max;
    const int e = e0 + lane;
    // each wave's share of the intervals for the Akima pass
    const int ia = (wave * NI) >> 2, ib = ((wave + 1) * NI) >> 2;

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
            __syncthreads();   // the previous image's fold is done with CO
            const double* g = F.geo + vix * 5 * N_L;
            for (int q = tid; q < N_L; q += 256) {
                PH[q] = g[q]; Zi[q] = 1.0 / g[N_L + q]; As[q] = g[2 * N_L + q]; Gs[q] = g[3 * N_L + q];
                PV[q] = g[4 * N_L + q];
            }
            if (ATM == 2 && tid == 0) prev_query_state(F, L, mi, i, I, ST);
            const double a_ref = (sizeof(R) == 4 && ABITS[i & 1]) ? __longlong_as_double((long long)ABITS[i & 1]) : 1.0;
            if (sizeof(R) == 4 && I == 0)
                for (int j = tid; j < S; j += 256) carea[j] = carea[j] / a_ref;
            __syncthreads();
            // profile: specific intensity x geometry, zero where the leaf is hidden
            for (int q = tid; q < N_L * 64; q += 256) {
                const int mm = q >> 6, k = q & 63, ek = e0 + k;
                R v = (R)0;
                const double geom = Gs[mm];
                if (geom != 0.0 && ek < N_E && PV[mm] > -1.5) {
                    const double Ep = F.energies[ek] * Zi[mm];
                    if (ATM == 2) {
                        double vec[4], prev[4];
                        int has_prev = 1;
                        vec[0] = F.logT[ro]; vec[1] = F.logg[ro]; vec[2] = As[mm];
                        vec[3] = log10(Ep / kT);
                        if (ek > 0) {
                            prev[0] = vec[0]; prev[1] = vec[1]; prev[2] = vec[2];
                            prev[3] = log10((F.energies[ek - 1] * Zi[mm]) / kT);
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
                        // one rounding per argument; the norm is applied in FP64 at the end
                        const float Ef = (float)Ep;
                        const float Ep3 = Ef * Ef * Ef;
                        float I_E = Ep3 / (expf((float)(Ep / kT)) - 1.0f);
                        if (m.correct) I_E -= Ep3 / (expf((float)(Ep / kT_else)) - 1.0f);
                        v = (R)(I_E * (float)geom);
                    }
                }
                if (mm < NI) CO[mm * 64 + k].a = v;
                else YL[k] = v;
            }
            __syncthreads();
            if (tid < 64 && PV[N_L - 1] < -1.5) YL[tid] = CO[tid].a;
            __syncthreads();
            // periodic Akima coefficients: lane = energy, a sliding window of
            // the secant slopes m[idx-2 .. idx+2] (m[-2] = m[NI-2], m[-1] =
            // m[NI-1], m[NI] = m[0], m[NI+1] = m[1])
            {
                auto yv = [&](int k) { return k < NI ? CO[k * 64 + lane].a : YL[lane]; };
                auto slope = [&](int k) {
                    k = k < 0 ? k + NI : (k >= NI ? k - NI : k);
                    return (yv(k + 1) - yv(k)) / (R)(PH[k + 1] - PH[k]);
                };
                R w0 = slope(ia - 2), w1 = slope(ia - 1), w2 = slope(ia), w3 = slope(ia + 1), w4;
                for (int idx = ia; idx < ib; idx++) {
                    w4 = slope(idx + 2);
                    R b, c, d;
                    akima_piece_t<R>(w0, w1, w2, w3, w4, (R)(PH[idx + 1] - PH[idx]), b, c, d);
                    Co4<R>& co = CO[idx * 64 + lane];
                    co.b = b; co.c = c; co.d = d;
                    w0 = w1; w1 = w2; w2 = w3; w3 = w4;
                }
            }
            __syncthreads();
            // the fold: cells in order, the wave's 16 phases as independent
            // queries per cell (per-phase interval state in registers)
            const double lo = PH[0], hi = PH[N_L - 1];
            double ph[NQ];
            int cur[NQ];
            R acc[NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                ph[q] = (pw + q < N_P) ? F.phases[pw + q] : 0.0;
                cur[q] = -1;
                acc[q] = (R)0;
            }
            const int nq = min(NQ, N_P - pw);
            for (int j = 0; j < S; j++) {
                const double area = carea[j];
                if (isnan(area)) continue;
                const double phi = cphi[j];
                const R ar = (R)area;
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    if (q >= nq) break;
                    double x = ph[q] + phi;
                    int guard = 0;
                    if (x > hi) { while (x > hi && guard++ < 64) x -= TWO_PI; }
                    else if (x < lo) { while (x < lo && guard++ < 64) x += TWO_PI; }
                    int c = cur[q];
                    if (c < 0 || x < PH[c]) c = interval_of(PH, N_L, x);
                    else while (c < NI - 1 && PH[c + 1] <= x) c++;
                    cur[q] = c;
                    const double cl = PH[c];
                    const Co4<R> c4 = CO[c * 64 + lane];
                    const R t = (R)(x - cl);
                    const R v = c4.a + t * (c4.b + t * (c4.c + t * c4.d));
                    if (v > (R)0 || m.correct) acc[q] += ar * v;
                }
            }
#pragma unroll
            for (int q = 0; q < NQ; q++) tot[q] += (double)acc[q] * a_ref;
        }
    }
    if (e >= N_E) return;
    const double E = F.energies[e];
    double* fo = F.flux + (size_t)mi * N_E * N_P + (size_t)e * N_P;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int p = pw + q;
        if (p >= N_P) break;
        double s = tot[q];
        if (sizeof(R) == 4 && ATM == 1) s *= BB_NORM;
        if (gridDim.y == 1) fo[p] = s / (E * KEV);
        else atomicAdd(fo + p, s * (1.0 / (E * KEV)));
    }
}

#define XH_LANE_INST(R, NQ)                                                                     \
    template __global__ void k_lane_fold<R, 1, NQ>(const Member*, Layout, FoldArgs);            \
    template __global__ void k_lane_fold<R, 2, NQ>(const Member*, Layout, FoldArgs);
XH_LANE_INST(double, 8) XH_LANE_INST(double, 16)
XH_LANE_INST(float, 8) XH_LANE_INST(float, 16)

}  // namespace xh
