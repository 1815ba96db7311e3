// Pulse integration over a hot-region mesh without azimuthal invariance:
// every cell carries its own parameters (xpsi/cellmesh/integrator.pyx:48-667;
// its Stokes form xpsi/cellmesh/integratorIQU.pyx:51-717), the blackbody or
// burst blackbody with the beaming options (xh_hot.hpp), optionally minus a
// per-cell blackbody elsewhere intensity (the correction, integrator.pyx:580-590).
//
// k_general_geometry  one block per parallel, image orders in turn: leaf
//                     geometry in parallel, then the reference's leaf loop in
//                     leaf order on one thread, which pads the stretches where
//                     the parallel is hidden with linearly stepped phase,
//                     redshift and mu eta and a zero geometric factor.
//                     Writes PHASE / Z / ABB / GEOM per leaf.
// k_general_fold      one block per (parallel, image, energy tile x phase
//                     tile): the image's GEOM (periodic Akima), Z and mu eta
//                     (Steffen) interpolants are built in LDS; each lane
//                     owns one model phase and eight energies and walks the
//                     parallel's cells, interpolating GEOM and Z at the
//                     cell's phase point and adding area * GEOM * I(E / Z) of
//                     the cell's temperature (or the burst atmosphere and
//                     beaming at the interpolated mu eta; the Stokes form adds
//                     Q and U rotated by the parallel and image's angle).
//                     The block's sums, divided by E keV, meet in the flux by
//                     FP64 atomic additions.
//
// The bound is FP64 issue: per (cell, phase, energy) one exp and one
// division (two of each with the correction); the interpolation is shared by
// the lane's eight energies.
#include "xh_common.hpp"
#include "xh_types.hpp"
#include "xh_kernels.hpp"
#include "xh_leaf.hpp"
#include "xh_hot.hpp"

namespace xh {

namespace {

__device__ __forceinline__ bool takes_slot(int k, int ks, int leaf_lim, int N_L)
{
    return (0 < k && k < leaf_lim - 1) || (k == 0 && ks == 0) || (k == leaf_lim - 1 && N_L % 2 == 1 && ks == 0) ||
           (k == leaf_lim - 1 && N_L % 2 == 0);
}

// The leaf loop of integrator.pyx:316-512 for one parallel and image, in leaf
// order over all N_L leaves.  A visible leaf of the first half copies its
// precomputed values (aP/aZ/aA/aG, indexed by destination slot) in at its own
// step, so a padding written at an earlier step is overwritten exactly where
// the reference overwrites it.  vis[k] bit 0: leaf k is seen; bit 1: its
// deflection lies in the table.  Returns the final visibility flag (2: the
// image was never visible).
__device__ int leaf_pass(int N_L, int leaf_lim, const double* __restrict__ leaves, const unsigned char* vis,
                         const double* aP, const double* aZ, const double* aA, const double* aG, double* P,
                         double* Z, double* A, double* G)
{
    int invis = 2, ip = 0;
#pragma unroll 1
    for (int k = 0; k < N_L; k++) {
        if (vis[k] & 1) {
            if (k < leaf_lim) {
                for (int ks = 0; ks < 2; ks++) {
                    if (!takes_slot(k, ks, leaf_lim, N_L)) continue;
                    const int kd = ks == 0 ? k : N_L - 1 - k;
                    P[kd] = aP[kd]; Z[kd] = aZ[kd]; A[kd] = aA[kd]; G[kd] = aG[kd];
                }
            }
            if (k == 0) {
                P[N_L - 1] = P[0] + TWO_PI;
                Z[N_L - 1] = Z[0]; A[N_L - 1] = A[0]; G[N_L - 1] = G[0];
            } else if (invis == 2) {
                const double st = leaves[k] / (double)k;
                const double zs = (Z[k] - Z[N_L - k - 1]) / (2.0 * (double)k);
                const double as = (A[k] - A[N_L - k - 1]) / (2.0 * (double)k);
                for (int m = N_L - k; m < N_L; m++) {
                    P[m] = P[m - 1] + st; Z[m] = Z[m - 1] + zs; A[m] = A[m - 1] + as; G[m] = 0.0;
                }
                P[0] = P[N_L - 1] - TWO_PI;
                Z[0] = Z[N_L - 1]; A[0] = A[N_L - 1]; G[0] = G[N_L - 1];
                for (int m = 1; m < k; m++) {
                    P[m] = P[m - 1] + st; Z[m] = Z[m - 1] + zs; A[m] = A[m - 1] + as; G[m] = 0.0;
                }
            } else if (invis == 1) {
                const double n = (double)(k - ip + 1);
                double st = (P[k] - P[ip - 1]) / n, zs = (Z[k] - Z[ip - 1]) / n, as = (A[k] - A[ip - 1]) / n;
                for (int m = ip; m < k; m++) { P[m] = P[m - 1] + st; Z[m] = Z[m - 1] + zs; A[m] = A[m - 1] + as; }
                st = (P[N_L - ip] - P[N_L - 1 - k]) / n;
                zs = (Z[N_L - ip] - Z[N_L - 1 - k]) / n;
                as = (A[N_L - ip] - A[N_L - 1 - k]) / n;
                for (int m = N_L - k; m < N_L - ip; m++) {
                    P[m] = P[m - 1] + st; Z[m] = Z[m - 1] + zs; A[m] = A[m - 1] + as;
                }
            }
            invis = 0;
        } else if (invis == 0) {
            const double n = (double)(N_L - 2 * k + 1);
            const double st = (P[N_L - k] - P[k - 1]) / n, zs = (Z[N_L - k] - Z[k - 1]) / n,
                         as = (A[N_L - k] - A[k - 1]) / n;
            for (int m = k; m < N_L - k; m++) {
                P[m] = P[m - 1] + st; Z[m] = Z[m - 1] + zs; A[m] = A[m - 1] + as; G[m] = 0.0;
            }
            invis = 1;
            ip = k;
        }
    }
    return invis;
}

}  // namespace

__global__ void __launch_bounds__(128) k_general_geometry(Member* __restrict__ members, Layout L, RingArgs A)
{
    extern __shared__ double sm[];
    const int i = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    Member* mp = members;
    const Member m = *mp;
    if (i >= m.S) return;
    const size_t vbase = (size_t)i * L.IO_max;
    if (!A.ring_ok[i]) {
        for (int I = tid; I < L.IO_max; I += nt) A.valid[vbase + I] = 0;
        return;
    }
    const double maxD = A.maxDefl[i];
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
    double* aP = yalt + N_R;
    double* aZ = aP + N_L;
    double* aA = aZ + N_L;
    double* aG = aA + N_L;
    double* P = aG + N_L;
    double* Z = P + N_L;
    double* Ab = Z + N_L;
    double* G = Ab + N_L;
    unsigned char* vis = (unsigned char*)(G + N_L);
    __shared__ int jn_s, shown;
    const size_t tb = (size_t)i * N_R;
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
    for (int k = tid; k < N_L; k += nt) { P[k] = 0.0; Z[k] = 0.0; Ab[k] = 0.0; G[k] = 0.0; }
    const double incl = m.inclination, cos_i = cos(incl), sin_i = sin(incl);
    const double th = A.theta[i], cos_th = cos(th), sin_th = sin(th), thp = th / PI;
    const double cg = A.cos_gamma[i], sg = sqrt(1.0 - cg * cg);
    const double radius = A.r[i], Grav_z = sqrt(1.0 - A.u[i]);
    const double beta = radius * m.Omega * sin_th / (C_SI * Grav_z);
    const double Lorentz = sqrt(1.0 - beta * beta);
    const int leaf_lim = (N_L % 2 == 0) ? N_L / 2 : (N_L + 1) / 2;
    const double* leaves = A.leaves;
    const int n_orders = min(IO, L.IO_max);
    for (int I = 0; I < n_orders; I++) {
        const size_t vix = vbase + I;
        __syncthreads();   // tables staged; the previous order's outputs written
        for (int k = tid; k < N_L; k += nt) {
            const bool first = k < leaf_lim;
            LeafGeo g = leaf_geometry(I, leaves[k], incl, cos_i, sin_i, cos_th, sin_th, thp, cg, sg, maxD, D, Ac,
                                      Lg, N_R, xalt, yalt, nalt, first);
            vis[k] = (unsigned char)(g.calc | (g.psi <= maxD ? 2 : 0));
            if (!g.calc || !first) continue;
            for (int ks = 0; ks < 2; ks++) {
                if (!takes_slot(k, ks, leaf_lim, N_L)) continue;
                const int kd = ks == 0 ? k : N_L - 1 - k;
                double superlum, eta;
                if (!eq12(g.psi, 0.0)) {
                    const double cos_xi = g.sin_alpha * sin_i * sin(leaves[kd]) / g.sin_psi;
                    superlum = 1.0 + beta * cos_xi;
                    eta = Lorentz / superlum;
                } else {
                    superlum = 1.0;
                    eta = Lorentz;
                }
                aZ[kd] = eta * Grav_z;
                aA[kd] = g.mu * eta;
                aG[kd] = g.mu * fabs(g.deriv) * Grav_z * eta * eta * eta / superlum;
                aP[kd] = leaves[kd] + g.phase_lag;
            }
        }
        __syncthreads();
        if (tid == 0) {
            const int invis = leaf_pass(N_L, leaf_lim, leaves, vis, aP, aZ, aA, aG, P, Z, Ab, G);
            shown = invis != 2;
            if (shown && A.pol) {
                // the Stokes form's polarisation angle (integratorIQU.pyx:593-596),
                // formed from the leaf loop's last values: the last leaf written
                // (eta, cos xi), the last leaf (psi) and the last leaf within the
                // deflection table (alpha, mu)
                int k1 = -1, k2 = -1;
                for (int k = 0; k < N_L; k++) {
                    if (vis[k] & 2) k1 = k;
                    if ((vis[k] & 1) && k < leaf_lim) k2 = k;
                }
                double c2 = 1.0, s2 = 0.0;
                if (k1 >= 0 && k2 >= 0) {
                    const LeafGeo g1 = leaf_geometry(I, leaves[k1], incl, cos_i, sin_i, cos_th, sin_th, thp, cg, sg,
                                                     maxD, D, Ac, Lg, N_R, xalt, yalt, nalt, false);
                    const LeafGeo gN = leaf_geometry(I, leaves[N_L - 1], incl, cos_i, sin_i, cos_th, sin_th, thp, cg,
                                                     sg, maxD, D, Ac, Lg, N_R, xalt, yalt, nalt, false);
                    const LeafGeo g2 = leaf_geometry(I, leaves[k2], incl, cos_i, sin_i, cos_th, sin_th, thp, cg, sg,
                                                     maxD, D, Ac, Lg, N_R, xalt, yalt, nalt, false);
                    const int kd = takes_slot(k2, 1, leaf_lim, N_L) ? N_L - 1 - k2 : k2;
                    double cos_xi = 0.0, eta = Lorentz;
                    if (!eq12(g2.psi, 0.0)) {
                        cos_xi = g2.sin_alpha * sin_i * sin(leaves[kd]) / g2.sin_psi;
                        eta = Lorentz / (1.0 + beta * cos_xi);
                    }
                    const double chi = pol_angle(leaves[kd], gN.sin_psi, gN.cos_psi, g1.sin_alpha, g1.cos_alpha,
                                                 sin_th, cos_th, sin_i, cos_i, sg, cg, Grav_z, g1.mu, eta, beta,
                                                 Lorentz, cos_xi);
                    c2 = cos(2 * chi);
                    s2 = sin(2 * chi);
                }
                A.pol[vix * 2] = c2;
                A.pol[vix * 2 + 1] = s2;
            }
            if (!shown) {
                A.valid[vix] = 0;
            } else {
                int term = 0;
                for (int q = 1; q < N_L; q++)
                    if (P[q] <= P[q - 1]) { term = 1; break; }
                A.valid[vix] = term ? 0 : 1;
                if (term) atomicMax(&mp->status, 2);
            }
        }
        __syncthreads();
        double* out = A.geo + vix * 4 * N_L;
        for (int q = tid; q < N_L; q += nt) {
            out[q] = P[q];
            out[N_L + q] = Z[q];
            out[2 * N_L + q] = Ab[q];
            out[3 * N_L + q] = G[q];
        }
        if (!shown) {
            // no visible leaf at this order: the higher orders do not contribute
            for (int q = I + 1 + tid; q < n_orders; q += nt) A.valid[vbase + q] = 0;
            break;
        }
    }
}

// grid (parallel, image, energy tile + phase tile * n_etiles); 256 lanes:
// lane & 63 = phase within the 64-phase tile, lane >> 6 = energy group, the
// group's energies e0 + g + 4q for q < 8 (a 32-energy tile)
__global__ void __launch_bounds__(256) k_general_fold(Member* __restrict__ members, Layout L, GeneralArgs F)
{
    extern __shared__ double sm[];
    const int i = blockIdx.x, I = blockIdx.y, tid = threadIdx.x;
    const int S = members->S;
    if (i >= S || I >= L.IO_max || !F.valid[(size_t)i * L.IO_max + I]) return;
    const int N_L = L.N_L, N_P = L.N_P, N_E = L.N_E;
    const int n_et = (N_E + 31) / 32;
    const int e0 = (blockIdx.z % n_et) * 32, k0 = (blockIdx.z / n_et) * 64;
    double* PH = sm;
    double* GY = PH + N_L;
    double* Gb = GY + N_L;
    double* Gc = Gb + N_L;
    double* Gd = Gc + N_L;
    double* ZY = Gd + N_L;
    double* Za = ZY + N_L;
    double* Zb = Za + N_L;
    double* Zc = Zb + N_L;
    double* AY = Zc + N_L;      // mu eta: read by the burst atmosphere and beaming
    double* Aa = AY + N_L;
    double* Ab = Aa + N_L;
    double* Ac = Ab + N_L;
    const bool need_abb = F.atm != 1 || F.beam != 0;
    const size_t vix = (size_t)i * L.IO_max + I;
    const double* geo = F.geo + vix * 4 * N_L;
    for (int q = tid; q < N_L; q += blockDim.x) {
        PH[q] = geo[q];
        ZY[q] = geo[N_L + q];
        AY[q] = geo[2 * N_L + q];
        GY[q] = geo[3 * N_L + q];
    }
    __syncthreads();
    // Steffen coefficients: y = y_i + t (c + t (b + t a))
    auto steffen = [&](const double* Y, int q, double* a_, double* b_, double* c_) {
        const double h = PH[q + 1] - PH[q];
        const double s = (Y[q + 1] - Y[q]) / h;
        const double y0 = steffen_slope(PH, Y, N_L, q), y1 = steffen_slope(PH, Y, N_L, q + 1);
        a_[q] = (y0 + y1 - 2.0 * s) / h / h;
        b_[q] = (3.0 * s - 2.0 * y0 - y1) / h;
        c_[q] = y0;
    };
    // GEOM's phase interpolant (periodic Akima, Steffen, or the periodic cubic
    // spline solved by one thread); Z and mu eta are Steffen throughout
    if (F.interp == 2 && tid == 0) cspline_periodic(PH, GY, N_L, Gb, Gc, Gd, Ac + N_L);
    for (int q = tid; q < N_L - 1; q += blockDim.x) {
        if (F.interp != 2) {
            double b, c, d;
            phase_coefs(F.interp, PH, GY, N_L, q, nullptr, nullptr, nullptr, b, c, d);
            Gb[q] = b; Gc[q] = c; Gd[q] = d;
        }
        steffen(ZY, q, Za, Zb, Zc);
        if (need_abb) steffen(AY, q, Aa, Ab, Ac);
    }
    __syncthreads();
    const int k = k0 + (tid & 63), g = tid >> 6;
    const bool kin = k < N_P;
    const bool pol = F.flux_Q != nullptr;
    const double c2 = pol ? F.pol[vix * 2] : 0.0, s2 = pol ? F.pol[vix * 2 + 1] : 0.0;
    double E[8], acc[8], accQ[8], accU[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int p = e0 + g + 4 * q;
        E[q] = p < N_E ? F.energies[p] : 1.0;
        acc[q] = 0.0;
        accQ[q] = 0.0;
        accU[q] = 0.0;
    }
    const double x0 = kin ? F.phases[k] : 0.0;
    const double lo = PH[0], hi = PH[N_L - 1];
    const size_t row = (size_t)i * S;
    int bad = 0;
#pragma unroll 1
    for (int j = 0; j < S; j++) {
        if (F.radiates[row + j] != 1) continue;
        const double x = wrap_phase(x0 + F.phi[row + j], lo, hi);
        if (!(x >= lo && x <= hi)) { bad = 1; continue; }
        const int iv = interval_of(PH, N_L, x);
        const double t = x - PH[iv];
        const double Gx = GY[iv] + t * (Gb[iv] + t * (Gc[iv] + t * Gd[iv]));
        if (!(Gx > 0.0) || !kin) continue;
        const double Zx = ZY[iv] + t * (Zc[iv] + t * (Zb[iv] + t * Za[iv]));
        const double wA = F.cellArea[row + j] * Gx;
        const double* VEC = F.params + (row + j) * F.n_params;
        if (need_abb || pol) {
            // burst atmosphere, beaming or Stokes: the atmosphere at the
            // interpolated mu eta
            const double Ax = AY[iv] + t * (Ac[iv] + t * (Ab[iv] + t * Aa[iv]));
            const double* CV = F.correction ? F.correction + (row + j) * 2 : nullptr;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const double Ep = E[q] / Zx;
                const double IE = hot_I(F.atm, F.beam, Ep, Ax, VEC) * BB_NORM;
                const double cIE = CV ? planck(Ep, CV[0]) * BB_NORM : 0.0;
                acc[q] += wA * (IE - cIE);
                if (pol) {
                    const double QE = hot_Q(F.atm, Ep, Ax, VEC) * BB_NORM;
                    accQ[q] += wA * (QE * c2);
                    accU[q] += wA * (QE * s2);
                }
            }
            continue;
        }
        const double kT = K_B_OVER_KEV * pow(10.0, VEC[0]);
        if (F.correction) {
            const double kTc = K_B_OVER_KEV * pow(10.0, F.correction[(row + j) * 2]);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const double Ep = E[q] / Zx, E3 = Ep * Ep * Ep;
                const double IE = E3 / (exp(Ep / kT) - 1.0) * BB_NORM;
                const double cIE = E3 / (exp(Ep / kTc) - 1.0) * BB_NORM;
                acc[q] += wA * (IE - cIE);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const double Ep = E[q] / Zx;
                acc[q] += wA * (Ep * Ep * Ep / (exp(Ep / kT) - 1.0) * BB_NORM);
            }
        }
    }
    if (bad) atomicMax(&members->status, 2);
    if (!kin) return;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int p = e0 + g + 4 * q;
        if (p >= N_E) continue;
        unsafeAtomicAdd(&F.flux[(size_t)p * N_P + k], acc[q] / (E[q] * KEV));
        if (pol) {
            unsafeAtomicAdd(&F.flux_Q[(size_t)p * N_P + k], accQ[q] / (E[q] * KEV));
            unsafeAtomicAdd(&F.flux_U[(size_t)p * N_P + k], accU[q] / (E[q] * KEV));
        }
    }
}

}  // namespace xh
