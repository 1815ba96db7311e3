// Pulse integration over a hot-region mesh exploiting azimuthal invariance.
//
// Reference behaviour: xpsi/cellmesh/integrator_for_azimuthal_invariance.pyx
// :70-665 (the leaf x energy loop :342-549 and the energy x cell x phase
// spline loop :565-596), blackbody intensity hot_BB.pyx:54-98.
//
// k_geometry   one block per (parallel, member), image orders in turn:
//              ray-table lookups (Steffen), Doppler factors and the
//              invisible-phase padding of the leaf phases; writes
//              PHASE/Z/mu/GEOM per leaf.
// k_profile_fold one block per (energy chunk, parallel block, member): the
//              specific-intensity profile and its periodic Akima coefficients
//              are staged in LDS (interval-major, energy-minor) and every lane
//              folds one energy over the parallel's cells at its phases,
//              accumulating in registers.
#include "xh_common.hpp"
#include "xh_types.hpp"
#include "xh_kernels.hpp"
#include "xh_atm.hpp"
#include "xh_leaf.hpp"

namespace xh {

// One block per (parallel, member): the ring's ray tables are staged once,
// then the image orders are taken in turn.  An order contributes only when
// every lower order has a visible leaf (the reference's image loop), so the
// walk stops at the first order without one; later orders are marked invalid.
// OPT = 1 adds the hot-region options of the Stokes integrator (disk
// blocking, polarisation angles); OPT = 0 is the plain integrator's form,
// which keeps fewer values live per leaf.
template <int OPT>
__global__ void __launch_bounds__(128) k_geometry(Member* __restrict__ members, Layout L, RingArgs A)
{
    extern __shared__ double sm[];
    const int i = blockIdx.x, mi = blockIdx.z, tid = threadIdx.x, nt = blockDim.x;
    Member* mp = members + mi;
    const Member m = *mp;
    if (!m.active || m.status) return;
    if (i >= m.S) return;
    const size_t ro = (size_t)mi * L.S_max + i;
    const size_t vbase = ro * L.IO_max;
    if (!A.ring_ok[ro]) {
        for (int I = tid; I < L.IO_max; I += nt) A.valid[vbase + I] = 0;
        return;
    }
    const double maxD = A.maxDefl[ro];
    const int IO = L.image_limit > 0 ? L.image_limit : (int)ceil(maxD / PI);
    if (tid == 0 && IO > L.IO_max) atomicMax(&mp->status, 4);
    for (int I = tid; I < L.IO_max; I += nt)
        if (I >= IO) A.valid[vbase + I] = 0;
    const int N_R = L.N_R, N_L = L.N_L;
    double* D = sm;
    double* Ac = D + N_R;
    double* Lg = Ac + N_R;
    double* xalt = Lg + N_R;
    double* yalt = xalt + N_R;
    double* PH = yalt + N_R;            // N_L
    double* Zs = PH + N_L;              // N_L
    double* As = Zs + N_L;              // N_L
    double* Gs = As + N_L;              // N_L
    double* PV = Gs + N_L;              // N_L
    double* aPH = PV + N_L;             // N_L: leaf phases by slot, copied in at their step
    double* C2 = aPH + N_L;             // N_L: cos 2chi (Stokes outputs only)
    double* S2 = C2 + N_L;              // N_L: sin 2chi
    int* vis = (int*)(S2 + N_L);        // N_L/2+1
    __shared__ int jn_s, term, shown;
    const size_t tb = ro * N_R;
    for (int q = tid; q < N_R; q += nt) {
        D[q] = A.defl[tb + q];
        Ac[q] = A.cos_alpha[tb + q];
        Lg[q] = A.lag[tb + q];
    }
    if (tid == 0) jn_s = N_R - 1;
    __syncthreads();
    for (int q = tid; q < N_R; q += nt)
        if (D[q] > HALF_PI) atomicMin(&jn_s, q);
    __syncthreads();
    const int jn = jn_s, nalt = jn + 1;
    for (int q = tid; q < nalt; q += nt) {
        xalt[q] = cos(D[jn - q]);
        yalt[q] = Ac[jn - q];
    }
    const double incl = m.inclination, cos_i = cos(incl), sin_i = sin(incl);
    const double th = A.theta[ro], cos_th = cos(th), sin_th = sin(th), thp = th / PI;
    const double cg = A.cos_gamma[ro], sg = sqrt(1.0 - cg * cg);
    const double radius = A.r[ro], Grav_z = sqrt(1.0 - A.u[ro]);
    const double beta = radius * m.Omega * sin_th / (C_SI * Grav_z);
    const double Lorentz = sqrt(1.0 - beta * beta);
    const int leaf_lim = (N_L % 2 == 0) ? N_L / 2 : (N_L + 1) / 2;
    const double* leaves = A.leaves;
    const int n_orders = min(IO, L.IO_max);
    for (int I = 0; I < n_orders; I++) {
        const size_t vix = vbase + I;
        __syncthreads();   // the tables are staged; the previous order's outputs are written
        for (int k = tid; k < N_L; k += nt) { Gs[k] = 0.0; Zs[k] = 1.0; As[k] = 0.0; PV[k] = -3.0; C2[k] = 0.0; S2[k] = 0.0; }
        if (tid == 0) { term = 0; shown = 0; }
        __syncthreads();
        for (int k = tid; k < leaf_lim; k += nt) {
            LeafGeo g = leaf_geometry(I, leaves[k], incl, cos_i, sin_i, cos_th, sin_th, thp, cg, sg, maxD,
                                      D, Ac, Lg, N_R, xalt, yalt, nalt, true, (OPT && A.disk) ? A.R_in : 1.0e6,
                                      A.u[ro], radius);
            vis[k] = g.calc;
            if (!g.calc) continue;
            for (int ks = 0; ks < 2; ks++) {
                bool take = (0 < k && k < leaf_lim - 1) || (k == 0 && ks == 0) ||
                            (k == leaf_lim - 1 && N_L % 2 == 1 && ks == 0) ||
                            (k == leaf_lim - 1 && N_L % 2 == 0);
                if (!take) continue;
                int kd = ks == 0 ? k : N_L - 1 - k;
                double superlum, eta, cos_xi = 0.0;
                if (!eq12(g.psi, 0.0)) {
                    cos_xi = g.sin_alpha * sin_i * sin(leaves[kd]) / g.sin_psi;
                    superlum = 1.0 + beta * cos_xi;
                    eta = Lorentz / superlum;
                } else {
                    superlum = 1.0;
                    eta = Lorentz;
                }
                Zs[kd] = eta * Grav_z;
                As[kd] = g.mu * eta;
                Gs[kd] = g.mu * fabs(g.deriv) * Grav_z * eta * eta * eta / superlum;
                aPH[kd] = leaves[kd] + g.phase_lag;
                if (OPT && A.pol) {
                    const double chi = pol_angle(leaves[kd], g.sin_psi, g.cos_psi, g.sin_alpha, g.cos_alpha, sin_th,
                                                 cos_th, sin_i, cos_i, sg, cg, Grav_z, g.mu, eta, beta, Lorentz,
                                                 cos_xi);
                    C2[kd] = cos(2 * chi);
                    S2[kd] = sin(2 * chi);
                }
            }
        }
        __syncthreads();
        if (tid == 0) {
            // invisible-phase padding, in the reference's leaf order (:480-549);
            // a visible leaf's phases enter at its own step, so a padding of an
            // earlier step is overwritten where the reference overwrites it
            int invis = 2, invis_phase = 0;
            for (int k = 0; k < leaf_lim; k++) {
                if (vis[k]) {
                    for (int ks = 0; ks < 2; ks++) {
                        bool take = (0 < k && k < leaf_lim - 1) || (k == 0 && ks == 0) ||
                                    (k == leaf_lim - 1 && N_L % 2 == 1 && ks == 0) ||
                                    (k == leaf_lim - 1 && N_L % 2 == 0);
                        if (take) PH[ks == 0 ? k : N_L - 1 - k] = aPH[ks == 0 ? k : N_L - 1 - k];
                    }
                    if (k == 0) {
                        PH[N_L - 1] = PH[0] + TWO_PI;
                        Zs[N_L - 1] = Zs[0]; As[N_L - 1] = As[0]; Gs[N_L - 1] = Gs[0];
                        C2[N_L - 1] = C2[0]; S2[N_L - 1] = S2[0];
                    } else if (invis == 2) {
                        double step = leaves[k] / (double)k;
                        for (int q = N_L - k; q < N_L; q++) PH[q] = PH[q - 1] + step;
                        PH[0] = PH[N_L - 1] - TWO_PI;
                        for (int q = 1; q < k; q++) PH[q] = PH[q - 1] + step;
                    } else if (invis == 1) {
                        double step = (PH[k] - PH[invis_phase - 1]) / (double)(k - invis_phase + 1);
                        for (int q = invis_phase; q < k; q++) PH[q] = PH[q - 1] + step;
                        step = (PH[N_L - invis_phase] - PH[N_L - 1 - k]) / (double)(k - invis_phase + 1);
                        for (int q = N_L - k; q < N_L - invis_phase; q++) PH[q] = PH[q - 1] + step;
                    }
                    invis = 0;
                } else if (invis == 0) {
                    double step = (PH[N_L - k] - PH[k - 1]) / (double)(N_L - 2 * k + 1);
                    for (int q = k; q < N_L - k; q++) PH[q] = PH[q - 1] + step;
                    invis = 1;
                    invis_phase = k;
                }
            }
            // order of the intensity queries (leaf k ascending, then its mirror):
            // PV[kd] = previous queried leaf, -1 = last query of the previous
            // parallel/image, -2 = copied from leaf 0 (no query)
            int prevq = -1;
            for (int k = 0; k < leaf_lim; k++) {
                if (!vis[k]) continue;
                for (int ks = 0; ks < 2; ks++) {
                    bool take = (0 < k && k < leaf_lim - 1) || (k == 0 && ks == 0) ||
                                (k == leaf_lim - 1 && N_L % 2 == 1 && ks == 0) ||
                                (k == leaf_lim - 1 && N_L % 2 == 0);
                    if (!take) continue;
                    int kd = ks == 0 ? k : N_L - 1 - k;
                    PV[kd] = (double)prevq;
                    prevq = kd;
                }
                if (k == 0) PV[N_L - 1] = -2.0;
            }
            A.lastq[vix] = prevq;
            if (invis == 2) {
                A.valid[vix] = 0;
            } else {
                shown = 1;
                for (int q = 1; q < N_L; q++)
                    if (PH[q] <= PH[q - 1]) { term = 1; break; }
                A.valid[vix] = term ? 0 : 1;
                if (term) atomicMax(&mp->status, 2);
            }
        }
        __syncthreads();
        double* out = A.geo + vix * 5 * N_L;
        for (int q = tid; q < N_L; q += nt) {
            out[q] = PH[q];
            out[N_L + q] = Zs[q];
            out[2 * N_L + q] = As[q];
            out[3 * N_L + q] = Gs[q];
            out[4 * N_L + q] = PV[q];
        }
        if (OPT && A.pol) {
            double* po = A.pol + vix * 2 * N_L;
            for (int q = tid; q < N_L; q += nt) { po[q] = C2[q]; po[N_L + q] = S2[q]; }
        }
        if (!shown) {
            // no visible leaf at this order: the higher orders do not contribute
            for (int q = I + 1 + tid; q < n_orders; q += nt) A.valid[vbase + q] = 0;
            break;
        }
    }
}


template <int EC, int MAXK, int ATM>
__global__ void __launch_bounds__(256, 4) k_profile_fold(const Member* __restrict__ members, Layout L, FoldArgs F)
{
    extern __shared__ double sm[];
    const int mi = blockIdx.z, tid = threadIdx.x;
    const Member m = members[mi];
    if (!m.active || m.status) return;
    const int S = m.S, N_L = L.N_L, N_P = L.N_P, N_E = L.N_E;
    const int e0 = blockIdx.x * EC;
    const int r0 = blockIdx.y * F.ring_block;
    if (r0 >= S) return;
    const int r1 = min(S, r0 + F.ring_block);
    constexpr int PPAR = 256 / EC;
    const int NI = N_L - 1;               // intervals
    double* Y = sm;                       // [N_L][EC]  profile
    double* Bc = Y + N_L * EC;            // [NI][EC]   Akima b
    double* Cc = Bc + NI * EC;            // [NI][EC]   Akima c
    double* Dc = Cc + NI * EC;            // [NI][EC]   Akima d
    double* SL = Dc + NI * EC;            // [NI][EC]   secant slopes
    double* PH = SL + NI * EC;            // N_L
    double* Zs = PH + N_L;
    double* As = Zs + N_L;
    double* Gs = As + N_L;
    double* PV = Gs + N_L;
    double* cphi = PV + N_L;              // S_max
    double* carea = cphi + L.S_max;       // S_max
    const int p = tid % EC, slot = tid / EC;
    const int pe = e0 + p;
    const bool p_ok = pe < N_E;
    const double E = p_ok ? F.energies[pe] : 1.0;
    double ph[MAXK], acc[MAXK];
#pragma unroll
    for (int q = 0; q < MAXK; q++) {
        int k = slot + q * PPAR;
        ph[q] = k < N_P ? F.phases[k] : 0.0;
        acc[q] = 0.0;
    }
    const size_t cbase = (size_t)mi * L.S_max * L.S_max;
    __shared__ double prevq[5];   // has-previous flag, then logT, logg, mu, log10(E/kT)
    for (int i = r0; i < r1; i++) {
        const size_t ro = (size_t)mi * L.S_max + i;
        __syncthreads();   // the previous parallel's fold is done with LDS
        for (int j = tid; j < S; j += 256) {
            size_t c = cbase + (size_t)i * S + j;
            double a = F.cellArea[c];
            bool rad = F.radiates ? (F.radiates[c] == 1) : (a > 0.0);
            carea[j] = rad ? a : __builtin_nan("");
            cphi[j] = F.phi[(size_t)mi * L.S_max * (F.phi_row_stride ? L.S_max : 1) +
                            (size_t)i * F.phi_row_stride + j];
        }
        const double kT = K_B_OVER_KEV * pow(10.0, F.logT[ro]);
        const double kT_else = K_B_OVER_KEV * pow(10.0, m.else_logT);
        for (int I = 0; I < L.IO_max; I++) {
            const size_t vix = ro * L.IO_max + I;
            if (!F.valid[vix]) break;
            __syncthreads();
            const double* g = F.geo + vix * 5 * N_L;
            for (int q = tid; q < N_L; q += 256) {
                PH[q] = g[q]; Zs[q] = g[N_L + q]; As[q] = g[2 * N_L + q]; Gs[q] = g[3 * N_L + q];
                PV[q] = g[4 * N_L + q];
            }
            if (ATM == 2 && tid == 0) prev_query_state(F, L, mi, i, I, prevq);
            __syncthreads();
            // specific-intensity profile, zero where the leaf is hidden
            for (int q = tid; q < N_L * EC; q += 256) {
                int mm = q / EC, pp = q % EC;
                double v = 0.0;
                double geom = Gs[mm];
                if (geom != 0.0 && e0 + pp < N_E && PV[mm] > -1.5) {
                    double Ep = F.energies[e0 + pp] / Zs[mm];
                    if (ATM == 2) {
                        double vec[4], prev[4];
                        int has_prev = 1;
                        vec[0] = F.logT[ro]; vec[1] = F.logg[ro]; vec[2] = As[mm];
                        vec[3] = log10(Ep / kT);
                        if (e0 + pp > 0) {
                            prev[0] = vec[0]; prev[1] = vec[1]; prev[2] = vec[2];
                            prev[3] = log10((F.energies[e0 + pp - 1] / Zs[mm]) / kT);
                        } else if (PV[mm] > -0.5) {
                            int kd = (int)PV[mm];
                            prev[0] = vec[0]; prev[1] = vec[1]; prev[2] = As[kd];
                            prev[3] = log10((F.energies[N_E - 1] / Zs[kd]) / kT);
                        } else if (prevq[0] > 0.5) {
                            prev[0] = prevq[1]; prev[1] = prevq[2]; prev[2] = prevq[3]; prev[3] = prevq[4];
                        } else has_prev = 0;
                        double corr = 0.0;
                        if (m.correct) corr = Ep * Ep * Ep / (exp(Ep / kT_else) - 1.0) * BB_NORM;
                        v = (num4d_I(F, vec, prev, has_prev) * NUM4D_NORM - corr) * geom;
                    } else {
                        double I_E = Ep * Ep * Ep / (exp(Ep / kT) - 1.0);
                        double corr = 0.0;
                        if (m.correct) corr = Ep * Ep * Ep / (exp(Ep / kT_else) - 1.0) * BB_NORM;
                        v = (I_E * BB_NORM - corr) * geom;
                    }
                }
                Y[q] = v;
            }
            __syncthreads();
            for (int q = tid; q < EC; q += 256)
                if (PV[N_L - 1] < -1.5) Y[(N_L - 1) * EC + q] = Y[q];
            __syncthreads();
            // secant slopes, one division per interval and energy
            for (int q = tid; q < NI * EC; q += 256) {
                int mm = q / EC;
                SL[q] = (Y[q + EC] - Y[q]) / (PH[mm + 1] - PH[mm]);
            }
            __syncthreads();
            // periodic Akima coefficients (slopes wrap: m[-2] = m[NI-2], ...)
            for (int q = tid; q < NI * EC; q += 256) {
                int mm = q / EC, pp = q % EC;
                auto sl = [&](int k) {
                    int kk = k < 0 ? k + NI : (k >= NI ? k - NI : k);
                    return SL[kk * EC + pp];
                };
                double b, c, d;
                akima_piece(sl(mm - 2), sl(mm - 1), sl(mm), sl(mm + 1), sl(mm + 2), PH[mm + 1] - PH[mm], b, c, d);
                Bc[q] = b;
                Cc[q] = c;
                Dc[q] = d;
            }
            __syncthreads();
            const double lo = PH[0], hi = PH[N_L - 1];
            int idx[MAXK];
#pragma unroll
            for (int q = 0; q < MAXK; q++) idx[q] = 0;
            for (int j = 0; j < S; j++) {
                const double area = carea[j];
                if (isnan(area)) continue;
                const double sh = cphi[j];
#pragma unroll
                for (int q = 0; q < MAXK; q++) {
                    int k = slot + q * PPAR;
                    if (k >= N_P) break;
                    double x = ph[q] + sh;
                    int guard = 0;
                    if (x > hi) { while (x > hi && guard++ < 64) x -= TWO_PI; }
                    else if (x < lo) { while (x < lo && guard++ < 64) x += TWO_PI; }
                    // interval search: x grows with the cell index between wraps
                    int id = idx[q];
                    if (x < PH[id]) id = 0;
                    while (id < NI - 1 && PH[id + 1] <= x) id++;
                    idx[q] = id;
                    double t = x - PH[id];
                    int o = id * EC + p;
                    double v = Y[o] + t * (Bc[o] + t * (Cc[o] + t * Dc[o]));
                    if (v > 0.0 || m.correct) acc[q] += area * v;
                }
            }
        }
    }
    if (!p_ok) return;
    double* fo = F.flux + (size_t)mi * N_E * N_P + (size_t)pe * N_P;
    const double inv = 1.0 / (E * KEV);
#pragma unroll
    for (int q = 0; q < MAXK; q++) {
        int k = slot + q * PPAR;
        if (k >= N_P) break;
        if (gridDim.y == 1) fo[k] = acc[q] / (E * KEV);
        else atomicAdd(fo + k, acc[q] * inv);
    }
}

// per member: sum over parallels and contributing image orders of the number
// of radiating cells (the cell x phase x energy evaluation count / (N_E N_P))
__global__ void __launch_bounds__(256) k_work_count(const Member* __restrict__ members, Layout L,
                                                    const double* __restrict__ cellArea,
                                                    const int* __restrict__ valid, double* __restrict__ work)
{
    const int mi = blockIdx.x;
    const Member m = members[mi];
    __shared__ double acc;
    if (threadIdx.x == 0) acc = 0.0;
    __syncthreads();
    if (m.active && !m.status) {
        double w = 0.0;
        for (int i = threadIdx.x; i < m.S; i += blockDim.x) {
            const double* ca = cellArea + (size_t)mi * L.S_max * L.S_max + (size_t)i * m.S;
            int c = 0;
            for (int j = 0; j < m.S; j++) c += ca[j] > 0.0;
            int imgs = 0;
            for (int I = 0; I < L.IO_max; I++) {
                if (!valid[((size_t)mi * L.S_max + i) * L.IO_max + I]) break;
                imgs++;
            }
            w += (double)c * imgs;
        }
        atomicAdd(&acc, w);
    }
    __syncthreads();
    if (threadIdx.x == 0) work[mi] = acc;
}

#define XH_FOLD_INST(EC, K)                                                                   \
    template __global__ void k_profile_fold<EC, K, 1>(const Member*, Layout, FoldArgs);        \
    template __global__ void k_profile_fold<EC, K, 2>(const Member*, Layout, FoldArgs);
XH_FOLD_INST(64, 4) XH_FOLD_INST(64, 8) XH_FOLD_INST(64, 16)
XH_FOLD_INST(32, 4) XH_FOLD_INST(32, 8) XH_FOLD_INST(32, 16)
XH_FOLD_INST(16, 4) XH_FOLD_INST(16, 8) XH_FOLD_INST(16, 16)
XH_FOLD_INST(8, 4) XH_FOLD_INST(8, 8) XH_FOLD_INST(8, 16)

template __global__ void k_geometry<0>(Member*, Layout, RingArgs);
template __global__ void k_geometry<1>(Member*, Layout, RingArgs);

}  // namespace xh
