// Allocation-free variant of the marginal-likelihood drop-in.
//
// xh_eval_marginal_likelihood takes a caller workspace for its per-channel
// scratch (log-likelihood terms and status words) and the summed lnL.  This
// variant keeps that scratch in a file-scope device array instead, so that a
// caller that cannot allocate (code under stream capture, tight serving loops,
// harnesses that forbid hipMalloc) passes only its inputs and outputs.  It
// runs the same kernels on the same inputs, so its results are the same.
//
// Limits and sharing: at most XH_NOALLOC_MAX_CH channels; every GPU the
// library runs on holds its own copy of the array (≈193 KiB) until the process
// ends, and all calls on one GPU share it, so calls that overlap on one GPU
// (different streams or threads) race.  The kernels write every scratch
// word they read (each channel's terms and status, then the sum), so the
// array needs no clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/xpsi_hip.h"

// the library's per-thread error message (xh_last_error)
int xh_fail(int code, const std::string& msg);

namespace {

constexpr size_t kScratchBytes =
    ((size_t)XH_NOALLOC_MAX_CH * (8 + 4) + 64 + 256 + 255) & ~(size_t)255;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];          // device address of `scratch`, per device

// The array's device address on the current device, looked up once per device
// with that device current.  Host code uses only this address, never the
// array's name, which in host code is a stand-in and not a device address.
int scratch_address(void** out, std::string& err)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { err = hipGetErrorString(e); return XH_EHIP; }
    if (dev < 0 || dev >= kMaxDevices) { err = "device index beyond 64"; return XH_EINVAL; }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e);
            return XH_EHIP;
        }
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return XH_OK;
}

}  // namespace

extern "C" int xh_eval_marginal_likelihood_noalloc(
    double exposure_time, int N_b, const double* phases, int N_ch, const double* counts, int n_comp,
    const double* components, int N_P, const double* comp_phases, const double* shifts,
    const double* precomp, const double* support, double epsabs, double epsrel, double epsilon,
    double sigmas, double llzero, int allow_negative, double slim, const double* background,
    double* lnl_out, double* STAR, double* B_ML, double* B_ML_supp, void* stream)
{
    if (N_ch < 1 || N_ch > XH_NOALLOC_MAX_CH)
        return xh_fail(XH_ERANGE, "xh_eval_marginal_likelihood_noalloc: N_ch must be 1.." +
                                      std::to_string(XH_NOALLOC_MAX_CH) + " (the static scratch's size)");
    void* ws = nullptr;
    std::string err;
    int rc = scratch_address(&ws, err);
    if (rc != XH_OK) return xh_fail(rc, "xh_eval_marginal_likelihood_noalloc: " + err);
    return xh_eval_marginal_likelihood(exposure_time, N_b, phases, N_ch, counts, n_comp, components, N_P,
                                       comp_phases, shifts, precomp, support, epsabs, epsrel, epsilon, sigmas,
                                       llzero, allow_negative, slim, background, lnl_out, STAR, B_ML,
                                       B_ML_supp, ws, xh_marginal_workspace_bytes(N_ch), stream);
}
