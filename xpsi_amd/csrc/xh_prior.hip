// Prior inverse sampling on the device, for sampler proposals in batches.
//
// Reference behaviour: xpsi/Prior.py:84-160 (the base prior: flat densities
// between bounds, inverse_sample maps each hypercube coordinate u to
// lo + (hi - lo) u; __call__ is 0 inside the bounds, -inf outside) and the
// examples' compactness prior examples/examples_fast/Modules/CustomPrior.py:49-66
// (radius <= 16 km; polar radius above the photon-sphere bound; a single
// maximum of |z| on the surface), with the spacetime quantities of
// xpsi/Spacetime.py:111-190.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/xpsi_hip.h"

int xh_fail(int code, const std::string& msg);

namespace {

// one thread per point: params[q][k] = lo_k + (hi_k - lo_k) u[q][k]; the log
// prior of the point (0 or -inf)
__global__ void __launch_bounds__(256) k_prior_inverse_sample(int n, int d, const double* __restrict__ u,
                                                              const double* __restrict__ bounds,
                                                              int idx_mass, int idx_radius, double frequency,
                                                              double* __restrict__ params,
                                                              double* __restrict__ logprior)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    bool ok = true;
    double mass = 0.0, radius = 0.0;
    for (int k = 0; k < d; k++) {
        const double lo = bounds[2 * k], hi = bounds[2 * k + 1];
        const double v = lo + (hi - lo) * u[(size_t)q * d + k];
        params[(size_t)q * d + k] = v;
        if (!(lo <= v) || !(v <= hi)) ok = false;
        if (k == idx_mass) mass = v;
        if (k == idx_radius) radius = v;
    }
    if (ok && idx_mass >= 0 && idx_radius >= 0) {
        // CustomPrior.py:55-66 with the Spacetime.py derived quantities
        const double c = 2.99792458e8, G = 6.6740831e-11, GM = 0.5 * 2.95325024e3;
        if (!(radius <= 16.0)) ok = false;
        const double r_g = mass * GM, M = mass * GM * c * c / G, R = radius * 1.0e3;
        const double Omega = 2.0 * M_PI * frequency;
        const double zeta = r_g / R, eps = Omega * Omega * R * R * R / (G * M);
        const double R_r_s = R / (2.0 * r_g);
        const double R_p = 1.0 + eps * (-0.788 + 1.030 * zeta);
        if (R_p < 1.505 / R_r_s) ok = false;
        const double arg = -1.0 / (3.0 * eps * (-0.788 + 1.030 * zeta));
        if (arg >= 0.0 && sqrt(arg) < 1.0) ok = false;
    }
    logprior[q] = ok ? 0.0 : -INFINITY;
}

}  // namespace

extern "C" int xh_prior_inverse_sample(int n, int d, const double* hypercube, const double* bounds, int idx_mass,
                                       int idx_radius, double frequency, double* params, double* logprior,
                                       void* stream)
{
    if (n < 0 || d < 1) return xh_fail(XH_EINVAL, "xh_prior_inverse_sample: bad sizes");
    if (idx_mass >= d || idx_radius >= d) return xh_fail(XH_EINVAL, "xh_prior_inverse_sample: index beyond d");
    if (n == 0) return XH_OK;
    hipLaunchKernelGGL(k_prior_inverse_sample, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, d,
                       hypercube, bounds, idx_mass, idx_radius, frequency, params, logprior);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return xh_fail(XH_EHIP, hipGetErrorString(e));
    return XH_OK;
}
