// The atmosphere extension point, batched on the device.
//
// Reference behaviour: the C-level plugin interface of the hot-region
// integrators, xpsi/surface_radiation_field/hot_wrapper.pyx:72-252 -
// init_hot(threads, preloaded table, ext) -> data, eval_hot_I(thread, E, mu,
// VEC, data, beam_opt) -> intensity, eval_hot_norm() -> norm,
// free_hot(threads, data) - for the blackbody (ext 1, hot_BB.pyx:54-98) and
// the four-dimensional table (ext 2, hot_Num4D.pyx:66-460, table preloaded
// as preload.pyx:6-44 lays it out).  Here one call evaluates a whole sequence
// of queries, one thread each: for the table the reference's base-node state
// after query q-1 is reconstructed from that query alone (num4d_I), so the
// sequence gives what one reference thread gives evaluating it in order from
// a fresh init_hot.
#include <hip/hip_runtime.h>

#include <string>

#include "xh_atm.hpp"
#include "../../include/xpsi_hip.h"

int xh_fail(int code, const std::string& msg);

struct xh_atmosphere {
    int ext;
    xh::FoldArgs F;          // table axes, lengths and values (ext 2)
    double* d_axes = nullptr;
    double* d_I = nullptr;
};

namespace {

__global__ void __launch_bounds__(256) k_atm_eval(int ext, xh::FoldArgs F, int n, const double* __restrict__ E,
                                                  const double* __restrict__ mu, const double* __restrict__ vec2,
                                                  double* __restrict__ out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double kT = xh::K_B_OVER_KEV * pow(10.0, vec2[2 * q]);
    if (ext == 1) {
        const double e = E[q];
        out[q] = e * e * e / (exp(e / kT) - 1.0);
        return;
    }
    double v[4], prev[4];
    v[0] = vec2[2 * q]; v[1] = vec2[2 * q + 1]; v[2] = mu[q];
    v[3] = log10(E[q] / kT);
    if (q > 0) {
        prev[0] = vec2[2 * q - 2]; prev[1] = vec2[2 * q - 1]; prev[2] = mu[q - 1];
        prev[3] = log10(E[q - 1] / (xh::K_B_OVER_KEV * pow(10.0, vec2[2 * q - 2])));
    }
    out[q] = xh::num4d_I(F, v, prev, q > 0);
}

}  // namespace

extern "C" xh_atmosphere* xh_atm_init(int ext, const double* const* axes, const int* N, const double* I)
{
    if (ext != 1 && ext != 2) { xh_fail(XH_EINVAL, "xh_atm_init: extension must be 1 or 2"); return nullptr; }
    xh_atmosphere* a = new xh_atmosphere();
    a->ext = ext;
    if (ext == 2) {
        size_t na = 0, ni = 1;
        for (int d = 0; d < 4; d++) {
            if (!axes || !axes[d] || N[d] < 4) { xh_fail(XH_EINVAL, "xh_atm_init: table axes missing or short"); delete a; return nullptr; }
            na += N[d];
            ni *= N[d];
            a->F.atm_N[d] = N[d];
        }
        if (hipMalloc(&a->d_axes, na * 8) != hipSuccess || hipMalloc(&a->d_I, ni * 8) != hipSuccess) {
            xh_fail(XH_ENOMEM, "xh_atm_init: device memory");
            (void)hipFree(a->d_axes);
            delete a;
            return nullptr;
        }
        size_t off = 0;
        for (int d = 0; d < 4; d++) {
            if (hipMemcpy(a->d_axes + off, axes[d], (size_t)N[d] * 8, hipMemcpyHostToDevice) != hipSuccess) {
                xh_fail(XH_EHIP, "xh_atm_init: copy");
                (void)hipFree(a->d_axes); (void)hipFree(a->d_I); delete a; return nullptr;
            }
            off += N[d];
        }
        if (hipMemcpy(a->d_I, I, ni * 8, hipMemcpyHostToDevice) != hipSuccess) {
            xh_fail(XH_EHIP, "xh_atm_init: copy");
            (void)hipFree(a->d_axes); (void)hipFree(a->d_I); delete a; return nullptr;
        }
        a->F.atm_axes = a->d_axes;
        a->F.atm_I = a->d_I;
    }
    return a;
}

extern "C" int xh_atm_eval_I(xh_atmosphere* a, int n, const double* E, const double* mu, const double* VEC,
                             double* out, void* stream)
{
    if (!a || n < 0) return xh_fail(XH_EINVAL, "xh_atm_eval_I: bad arguments");
    if (n == 0) return XH_OK;
    hipLaunchKernelGGL(k_atm_eval, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, a->ext, a->F, n, E, mu,
                       VEC, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return xh_fail(XH_EHIP, hipGetErrorString(e));
    return XH_OK;
}

extern "C" double xh_atm_norm(const xh_atmosphere* a)
{
    return a && a->ext == 2 ? xh::NUM4D_NORM : xh::BB_NORM;
}

extern "C" int xh_atm_free(xh_atmosphere* a)
{
    if (!a) return XH_OK;
    (void)hipFree(a->d_axes);
    (void)hipFree(a->d_I);
    delete a;
    return XH_OK;
}
