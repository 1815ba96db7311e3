// Stokes (I, Q, U) pulse integration exploiting azimuthal invariance, and the
// hot-region options the plain integrator shares with it
// (xpsi/cellmesh/integratorIQU_for_azimuthal_invariance.pyx:68-746,
// integrator_for_azimuthal_invariance.pyx:70-665):
//  - the blackbody and burst blackbody atmospheres and the beaming options
//    (xh_hot.hpp);
//  - the per-parallel blackbody elsewhere correction;
//  - disk blocking and the polarisation angle come from k_geometry.
//
// k_stokes_fold: one block per (parallel, image, energy x phase tile).  The
// image's I (and Q, U) profile over the leaves is formed in LDS from the
// geometry (redshift, mu eta, geometric factor, cos 2chi, sin 2chi), each
// with its periodic Akima coefficients; each thread owns one model phase and
// walks the parallel's cells in order, keeping a cell's (I, Q, U) where the
// interpolated I is positive (or always, with the correction), and adds the
// sums divided by E keV to the fluxes with FP64 atomics.  Bound: latency of a
// small launch; per (cell, phase) one interval search and one to three cubics.
#include "xh_common.hpp"
#include "xh_types.hpp"
#include "xh_kernels.hpp"
#include "xh_hot.hpp"

namespace xh {


// grid (parallel, image, energy + N_E * phase tile); blockDim.x phases per tile
__global__ void __launch_bounds__(256) k_stokes_fold(Member* __restrict__ members, Layout L, StokesArgs F)
{
    extern __shared__ double sm[];
    const int i = blockIdx.x, I = blockIdx.y, tid = threadIdx.x, nt = blockDim.x;
    const int S = members->S;
    if (i >= S || I >= L.IO_max) return;
    const size_t vix = (size_t)i * L.IO_max + I;
    if (!F.valid[vix]) return;
    const int N_L = L.N_L, N_P = L.N_P, N_E = L.N_E, NI = N_L - 1;
    const int e = blockIdx.z % N_E, k = (blockIdx.z / N_E) * nt + tid;
    const bool pol = F.flux_Q != nullptr;
    const int nst = pol ? 3 : 1;
    double* PH = sm;              // [N_L]
    double* Y = PH + N_L;         // [3][N_L] profiles I, Q, U
    double* Bc = Y + 3 * N_L;     // [3][N_L] Akima b, c, d
    double* Cc = Bc + 3 * N_L;
    double* Dc = Cc + 3 * N_L;     // then [3][N_L] of work for the cubic spline
    __shared__ int J_s;
    const size_t row = (size_t)i * S;
    if (tid == 0) {
        int J = -1;
        for (int j = 0; j < S; j++)
            if (F.radiates[row + j] == 1) { J = j; break; }
        J_s = J;
    }
    __syncthreads();
    const int J = J_s;
    if (J < 0) return;
    // the parallel's parameters are those of its first radiating cell (:303-306)
    const double* VEC = F.params + (row + J) * F.n_params;
    const double* geo = F.geo + vix * 5 * N_L;
    const double* pg = pol ? F.pol + vix * 2 * N_L : nullptr;
    const double E = F.energies[e];
    for (int q = tid; q < N_L; q += nt) {
        PH[q] = geo[q];
        const double Z = geo[N_L + q], abb = geo[2 * N_L + q], G = geo[3 * N_L + q];
        double vI = 0.0, vQ = 0.0, vU = 0.0;
        if (G != 0.0) {
            const double Ep = E / Z;
            const double IE = hot_I(F.atm, F.beam, Ep, abb, VEC);
            const double cIE = F.correction ? planck(Ep, F.correction[(row + J) * 2]) * BB_NORM : 0.0;
            vI = (IE * BB_NORM - cIE) * G;
            if (pol) {
                const double QE = hot_Q(F.atm, Ep, abb, VEC);
                vQ = (QE * pg[q] * BB_NORM) * G;
                vU = (QE * pg[N_L + q] * BB_NORM) * G;
            }
        }
        Y[q] = vI;
        Y[N_L + q] = vQ;
        Y[2 * N_L + q] = vU;
    }
    __syncthreads();
    // the profiles' phase interpolant (periodic Akima, Steffen, or the
    // periodic cubic spline, whose system one thread per profile solves)
    if (F.interp == 2) {
        double* Zw = Dc + 3 * N_L;   // [3][N_L] work
        if (tid < nst) cspline_periodic(PH, Y + tid * N_L, N_L, Bc + tid * N_L, Cc + tid * N_L, Dc + tid * N_L,
                                        Zw + tid * N_L);
    } else {
        for (int w = tid; w < nst * NI; w += nt) {
            const int s = w / NI, q = w % NI;
            double b, c, d;
            phase_coefs(F.interp, PH, Y + s * N_L, N_L, q, nullptr, nullptr, nullptr, b, c, d);
            Bc[s * N_L + q] = b;
            Cc[s * N_L + q] = c;
            Dc[s * N_L + q] = d;
        }
    }
    __syncthreads();
    if (k >= N_P) return;
    const double lo = PH[0], hi = PH[N_L - 1];
    const double x0 = F.phases[k];
    double aI = 0.0, aQ = 0.0, aU = 0.0;
    int bad = 0;
#pragma unroll 1
    for (int j = 0; j < S; j++) {
        if (F.radiates[row + j] != 1) continue;
        const double x = wrap_phase(x0 + F.phi[row + j], lo, hi);
        if (!(x >= lo && x <= hi)) { bad = 1; break; }
        const int iv = interval_of(PH, N_L, x);
        const double t = x - PH[iv];
        const double vI = Y[iv] + t * (Bc[iv] + t * (Cc[iv] + t * Dc[iv]));
        if (vI > 0.0 || F.correction) {
            const double area = F.cellArea[row + j];
            aI += area * vI;
            if (pol) {
                aQ += area * (Y[N_L + iv] + t * (Bc[N_L + iv] + t * (Cc[N_L + iv] + t * Dc[N_L + iv])));
                aU += area * (Y[2 * N_L + iv] + t * (Bc[2 * N_L + iv] + t * (Cc[2 * N_L + iv] + t * Dc[2 * N_L + iv])));
            }
        }
    }
    if (bad) atomicMax(&members->status, 2);
    const double den = E * KEV;
    const size_t o = (size_t)e * N_P + k;
    unsafeAtomicAdd(F.flux_I + o, aI / den);
    if (pol) {
        unsafeAtomicAdd(F.flux_Q + o, aQ / den);
        unsafeAtomicAdd(F.flux_U + o, aU / den);
    }
}

}  // namespace xh
