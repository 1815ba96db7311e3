// Time-invariant spectrum of the star outside its hot regions ("elsewhere").
//
// Reference behaviour: xpsi/cellmesh/integrator_for_time_invariance.pyx:59-338
// (blackbody, elsewhere_wrapper.pyx:23-81 -> hot_BB.pyx:54-98) and the
// addition of the spectrum to the first component of the first hot region at
// every phase (xpsi/Photosphere.py:590-592).
//
// k_else_integrate  one block per (ring, request).  The ring's ray tables go
//                   to LDS; thread j traces cell j through its image orders
//                   and parks (redshift, geometric factor) in fixed slots;
//                   then each thread sums one energy over the slots in order
//                   and writes the ring's partial spectrum.
// k_else_add        per (request, energy): sums the ring partials in ring
//                   order, scales by cell area, norm and 1/(E keV), and adds
//                   the result to every phase of the first member's flux.
#include "xh_common.hpp"
#include "xh_types.hpp"
#include "xh_kernels.hpp"

namespace xh {

constexpr int ELSE_IO_MAX = 8;   // image orders traced per cell

__device__ __forceinline__ double else_image_deflection(int order, double psi)
{
    if (order % 2 != 0) return (double)(order + 1) * PI - psi;
    return (double)order * PI + psi;
}

__global__ void __launch_bounds__(256) k_else_integrate(Member* __restrict__ members, Layout Le,
                                                        const double* __restrict__ colat,
                                                        const double* __restrict__ rr,
                                                        const double* __restrict__ cos_gamma,
                                                        const double* __restrict__ maxDefl,
                                                        const double* __restrict__ defl,
                                                        const double* __restrict__ cosa,
                                                        const double* __restrict__ energies, int N_E,
                                                        const double* __restrict__ phi, int phi_row_stride,
                                                        const double* __restrict__ params, int param_stride,
                                                        double* __restrict__ part)
{
    extern __shared__ double sm[];
    const int i = blockIdx.x, mi = blockIdx.y, tid = threadIdx.x;
    Member* mp = members + mi;
    const Member m = *mp;
    if (!m.active || m.status) return;
    const int S = m.S, N_R = Le.N_R;
    if (i >= S) return;
    double* D = sm;                         // [N_R] deflection
    double* Ac = D + N_R;                   // [N_R] cos alpha
    double* xalt = Ac + N_R;                // [N_R] cos(deflection), ascending
    double* yalt = xalt + N_R;              // [N_R]
    double* Zs = yalt + N_R;                // [S_max * ELSE_IO_MAX] redshift factor
    double* Gs = Zs + Le.S_max * ELSE_IO_MAX;   // geometric factor, 0 = no image
    double* Ks = Gs + Le.S_max * ELSE_IO_MAX;   // [S_max] kT of the cell
    __shared__ int jn_s, term;
    const size_t ro = (size_t)mi * Le.S_max + i;
    const size_t tb = ro * N_R;
    for (int q = tid; q < N_R; q += 256) { D[q] = defl[tb + q]; Ac[q] = cosa[tb + q]; }
    if (tid == 0) { jn_s = N_R - 1; term = 0; }
    __syncthreads();
    for (int q = tid; q < N_R; q += 256)
        if (D[q] > HALF_PI) atomicMin(&jn_s, q);
    __syncthreads();
    const int jn = jn_s, nalt = jn + 1;
    for (int q = tid; q < nalt; q += 256) { xalt[q] = cos(D[jn - q]); yalt[q] = Ac[jn - q]; }
    for (int q = tid; q < S * ELSE_IO_MAX; q += 256) Gs[q] = 0.0;
    __syncthreads();

    const double incl = m.inclination, cos_i = cos(incl), sin_i = sin(incl);
    const double th = colat[ro], cos_th = cos(th), sin_th = sin(th), thp = th / PI;
    const double cg = cos_gamma[ro], sg = sqrt(1.0 - cg * cg);
    const double radius = rr[ro], Grav_z = sqrt(1.0 - m.r_s / radius);
    const double beta = radius * m.Omega * sin_th / (C_SI * Grav_z);
    const double Lorentz = sqrt(1.0 - beta * beta);
    const double maxD = maxDefl[ro];
    const int IO = Le.image_limit > 0 ? Le.image_limit : (int)ceil(maxD / PI);
    if (tid == 0 && IO > ELSE_IO_MAX) atomicMax(&mp->status, 4);
    const double dphi = TWO_PI / (double)S;
    for (int j = tid; j < S; j += 256) {
        // cell azimuths: given, or the closed mesh's linspace(-pi + dphi/2, pi - dphi/2, S)
        double ph;
        if (phi) ph = phi[(size_t)mi * S * (phi_row_stride ? S : 1) + (size_t)i * phi_row_stride + j];
        else if (j == S - 1) ph = PI - 0.5 * dphi;
        else ph = (double)j * (((PI - 0.5 * dphi) - (-PI + 0.5 * dphi)) / (double)(S - 1)) + (-PI + 0.5 * dphi);
        const double lt = params ? params[((size_t)i * S + j) * param_stride] : m.logT;
        Ks[j] = K_B_OVER_KEV * pow(10.0, lt);
        const double c_psi0 = cos_i * cos_th + sin_i * sin_th * cos(ph);
        const double psi0 = acos(c_psi0);
        for (int I = 0; I < IO && I < ELSE_IO_MAX; I++) {
            double cos_psi = c_psi0;
            double psi = else_image_deflection(I, psi0);
            double sin_psi = sin(psi);
            if (!eq12(psi, 0.0) && eq12(sin_psi, 0.0)) {
                double ii = cos_i >= 0.0 ? incl + incl * 1.0e-6 : incl - incl * 1.0e-6;
                cos_psi = cos(ii) * cos_th + sin(ii) * sin_th * cos(ph);
                psi = else_image_deflection(I, acos(cos_psi));
                sin_psi = sin(psi);
            }
            if (psi > maxD) break;
            if (psi < D[0] || psi > D[N_R - 1]) { term = 1; break; }
            const bool alt = (psi <= HALF_PI && cos_psi >= xalt[0]);
            const double ca = alt ? steffen_eval(xalt, yalt, nalt, cos_psi, nullptr)
                                  : steffen_eval(D, Ac, N_R, psi, nullptr);
            const double sin_alpha = sqrt(1.0 - ca * ca);
            double mu = ca * cg;
            if (!eq12(psi, 0.0)) {
                const double cos_delta = (cos_i - cos_th * cos_psi) / (sin_th * sin_psi);
                if (thp < 0.5) mu = mu + sin_alpha * sg * cos_delta;
                else mu = mu - sin_alpha * sg * cos_delta;
            }
            if (!(mu > 0.0)) continue;
            double eta;
            if (!eq12(sin_psi, 0.0)) {
                const double cos_xi = sin_alpha * sin_i * sin(ph) / sin_psi;
                eta = Lorentz / (1.0 + beta * cos_xi);
            } else eta = Lorentz;
            double deriv;
            if (alt) steffen_eval(xalt, yalt, nalt, cos_psi, &deriv);
            else {
                steffen_eval(D, Ac, N_R, psi, &deriv);
                deriv = exp(log(fabs(deriv)) - log(fabs(sin_psi)));
            }
            Zs[j * ELSE_IO_MAX + I] = eta * Grav_z;
            Gs[j * ELSE_IO_MAX + I] = mu * fabs(deriv) * Grav_z * eta * eta * eta;
        }
    }
    __syncthreads();
    if (term) {
        if (tid == 0) atomicMax(&mp->status, 2);
        return;
    }
    double* out = part + ro * N_E;
    for (int e = tid; e < N_E; e += 256) {
        const double E = energies[e];
        double acc = 0.0;
        for (int j = 0; j < S; j++) {
            const double kT = Ks[j];
            for (int I = 0; I < ELSE_IO_MAX; I++) {
                const double G = Gs[j * ELSE_IO_MAX + I];
                if (G == 0.0) continue;
                const double Ep = E / Zs[j * ELSE_IO_MAX + I];
                acc += Ep * Ep * Ep / (exp(Ep / kT) - 1.0) * G;
            }
        }
        out[e] = acc;
    }
}

// drop-in finalisation: spectrum[e] = sum over rings (in order) x cell area x
// norm / (E keV) (integrator_for_time_invariance.pyx:318-324)
__global__ void __launch_bounds__(256) k_else_finalize(const Member* __restrict__ member, Layout Le,
                                                       const double* __restrict__ part, double cellArea,
                                                       const double* __restrict__ energies, int N_E,
                                                       double* __restrict__ out)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= N_E) return;
    double s = 0.0;
    for (int i = 0; i < member->S; i++) s += part[(size_t)i * N_E + e];
    out[e] = s * (cellArea * BB_NORM / (energies[e] * KEV));
}

// else_members[r]: the elsewhere member of request r; hot member 0 of the
// request is hot_members[r * members_per_request]
__global__ void __launch_bounds__(256) k_else_add(const Member* __restrict__ else_members, Layout Le,
                                                  const double* __restrict__ part,
                                                  const double* __restrict__ cellArea,
                                                  const double* __restrict__ energies, int N_E, int N_P,
                                                  Member* __restrict__ hot_members, int members_per_request,
                                                  double* __restrict__ flux)
{
    const int r = blockIdx.y, e = blockIdx.x * 256 + threadIdx.x;
    const Member m = else_members[r];
    if (!m.active) return;
    Member* h = hot_members + (size_t)r * members_per_request;
    if (m.status) {
        if (e == 0) atomicMax(&h->status, 2);
        return;
    }
    if (e >= N_E) return;
    const double* pr = part + (size_t)r * Le.S_max * N_E + e;
    double s = 0.0;
    for (int i = 0; i < m.S; i++) s += pr[(size_t)i * N_E];
    s *= cellArea[r] * BB_NORM / (energies[e] * KEV);
    double* fo = flux + (size_t)r * members_per_request * N_E * N_P + (size_t)e * N_P;
    for (int p = 0; p < N_P; p++) fo[p] += s;
}

}  // namespace xh
