// Host launch layer and C ABI (include/xpsi_hip.h).
//
// xh_pipeline owns every device buffer of a batch slot, sized once at
// creation for the configured maxima; xh_pipeline_eval only fills the
// per-request descriptors, launches the kernel sequence on one stream and
// copies back one double per request.
#include <hip/hip_runtime.h>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "xh_common.hpp"
#include "xh_kernels.hpp"
#include "xh_types.hpp"
#include "../../include/xpsi_hip.h"

using namespace xh;

static thread_local std::string g_err;
static int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}
#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) return fail(XH_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" const char* xh_last_error(void) { return g_err.c_str(); }
int xh_fail(int code, const std::string& msg) { return fail(code, msg); }

// physical constants of xpsi/global_imports.py:67-88
// Seeds of the llzero-class draws: a process-wide counter mixed with the
// start time, so every launch and every process draws afresh (the reference
// draws from numpy's global generator, xpsi/Likelihood.py:268-272).
static std::atomic<unsigned long long> g_seed_counter{0};
static unsigned long long next_seed()
{
    static const unsigned long long base =
        (unsigned long long)std::chrono::steady_clock::now().time_since_epoch().count() * 0x9E3779B97F4A7C15ULL;
    return base + g_seed_counter.fetch_add(1) + 1;
}

static const double kC = 2.99792458e8, kKPC = 3.08567758e19, kGM = 0.5 * 2.95325024e3,
                    kG = 6.6740831e-11, kKM = 1.0e3;

static int g_fold_ec = 0;   // 0: choose from the leaf count
static bool eq12_host(double x, double y) { return std::fabs(x - y) < 1.0e-12; }

static int pick_ec(int N_L)
{
    if (g_fold_ec) {
        int ec = g_fold_ec;
        while (ec > 8 && (size_t)5 * N_L * ec * 8 > 150 * 1024) ec >>= 1;
        return ec;
    }
    // 8 energies per block measured fastest at 64 leaves (more phase slots per
    // wave, 20 KiB of LDS); keep the staged arrays within 40 KiB otherwise
    int ec = 8;
    while (ec < 64 && (size_t)5 * N_L * (ec * 2) * 8 <= 40960 && N_L < 16) ec <<= 1;
    return ec;
}

// tuning hook: energies per fold block (8, 16, 32 or 64; 0 = automatic)
extern "C" int xh_set_fold_energy_chunk(int ec)
{
    if (ec != 0 && ec != 8 && ec != 16 && ec != 32 && ec != 64) return fail(XH_EINVAL, "chunk must be 0, 8, 16, 32 or 64");
    g_fold_ec = ec;
    return XH_OK;
}

static size_t fold_lds(int N_L, int EC, int S_max)
{
    return ((size_t)N_L * EC + 4 * (size_t)(N_L - 1) * EC + 5 * (size_t)N_L + 2 * (size_t)S_max) * 8;
}

static size_t geo_lds(int N_R, int N_L) { return (5 * (size_t)N_R + 8 * (size_t)N_L) * 8 + (N_L / 2 + 2) * 4; }

template <int EC, int K, int ATM>
static hipError_t launch_fold_t(dim3 g, size_t lds, hipStream_t s, const Member* m, Layout L, FoldArgs F)
{
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void*)&k_profile_fold<EC, K, ATM>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
        if (e != hipSuccess) return e;
        attr = true;
    }
    if (g.x == 0) return hipSuccess;     // configuration only
    hipLaunchKernelGGL((k_profile_fold<EC, K, ATM>), g, dim3(256), lds, s, m, L, F);
    return hipGetLastError();
}

// phases per lane: the smallest of 4, 8, 16 covering N_P at 256/EC phase slots
static int fold_k(int EC, int N_P)
{
    int per = (N_P + (256 / EC) - 1) / (256 / EC);
    return per <= 4 ? 4 : (per <= 8 ? 8 : 16);
}

template <int EC, int ATM>
static hipError_t launch_fold_e(int K, dim3 g, size_t lds, hipStream_t s, const Member* m, Layout L, FoldArgs F)
{
    if (K == 4) return launch_fold_t<EC, 4, ATM>(g, lds, s, m, L, F);
    if (K == 8) return launch_fold_t<EC, 8, ATM>(g, lds, s, m, L, F);
    return launch_fold_t<EC, 16, ATM>(g, lds, s, m, L, F);
}

template <int ATM>
static hipError_t launch_fold_a(int EC, dim3 g, size_t lds, hipStream_t s, const Member* m, Layout L, FoldArgs F)
{
    int K = fold_k(EC, L.N_P);
    switch (EC) {
    case 64: return launch_fold_e<64, ATM>(K, g, lds, s, m, L, F);
    case 32: return launch_fold_e<32, ATM>(K, g, lds, s, m, L, F);
    case 16: return launch_fold_e<16, ATM>(K, g, lds, s, m, L, F);
    default: return launch_fold_e<8, ATM>(K, g, lds, s, m, L, F);
    }
}

static hipError_t launch_fold(int EC, dim3 g, size_t lds, hipStream_t s, const Member* m, Layout L, FoldArgs F)
{
    if (lds > 150 * 1024) return hipErrorInvalidValue;
    return L.atm_ext == 2 ? launch_fold_a<2>(EC, g, lds, s, m, L, F) : launch_fold_a<1>(EC, g, lds, s, m, L, F);
}

// ------------------------------------------------- energy-lane fold
// dynamic LDS of k_lane_fold (mirrors its carve-up)
static size_t lane_lds(int N_L, int S_max, size_t rb)
{
    const size_t head = ((5 * (size_t)N_L + 2 * (size_t)S_max + 8 + 3) & ~(size_t)3) * 8;
    return head + 64 * rb + ((size_t)N_L - 1) * 64 * 4 * rb;
}

// phases per wave of the energy-lane fold (8 or 16); a tuning choice
static int g_lane_nq = 8;

extern "C" int xh_set_lane_phases(int nq)
{
    if (nq != 8 && nq != 16) return fail(XH_EINVAL, "phases per wave must be 8 or 16");
    g_lane_nq = nq;
    return XH_OK;
}

static dim3 lane_grid(int N_E, int N_P, int nq, int ring_blocks, int members)
{
    const int pb = 4 * nq;
    return dim3((unsigned)(((N_E + 63) / 64) * ((N_P + pb - 1) / pb)), (unsigned)ring_blocks, (unsigned)members);
}

template <typename R, int ATM, int NQ>
static hipError_t launch_lane_t(dim3 g, size_t lds, hipStream_t s, const Member* m, Layout L, FoldArgs F)
{
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void*)&k_lane_fold<R, ATM, NQ>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr = true;
    }
    if (g.x == 0) return hipSuccess;     // configuration only
    hipLaunchKernelGGL((k_lane_fold<R, ATM, NQ>), g, dim3(256), lds, s, m, L, F);
    return hipGetLastError();
}

template <typename R, int ATM>
static hipError_t launch_lane_a(int nq, dim3 g, size_t lds, hipStream_t s, const Member* m, Layout L, FoldArgs F)
{
    if (nq == 16) return launch_lane_t<R, ATM, 16>(g, lds, s, m, L, F);
    return launch_lane_t<R, ATM, 8>(g, lds, s, m, L, F);
}

static hipError_t launch_lane(bool fp32, int nq, dim3 g, size_t lds, hipStream_t s, const Member* m, Layout L,
                              FoldArgs F)
{
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (fp32) return L.atm_ext == 2 ? launch_lane_a<float, 2>(nq, g, lds, s, m, L, F)
                                    : launch_lane_a<float, 1>(nq, g, lds, s, m, L, F);
    return L.atm_ext == 2 ? launch_lane_a<double, 2>(nq, g, lds, s, m, L, F)
                          : launch_lane_a<double, 1>(nq, g, lds, s, m, L, F);
}

// blocks of the wave-per-ray QAG kernel (grid-stride over its work list)
constexpr int kRayWaveBlocks = 2048;

// ----------------------------------------------- moment form (matrix cores)
static int g_fold_kernel = 0;   // xh_integrate: 0 automatic, 1 per-point, 2 moment form, 3 energy lanes
static const size_t kMomentLdsMax = 160 * 1024;

extern "C" int xh_set_fold_kernel(int kind)
{
    if (kind < 0 || kind > 3)
        return fail(XH_EINVAL, "fold kernel must be 0 (automatic), 1 (per-point), 2 (moment) or 3 (energy lanes)");
    g_fold_kernel = kind;
    return XH_OK;
}

// moment-form block shape: T 16-phase tiles x G resident 16-energy chunks
// (T = G = 1, 2 or 4) per block of eight waves, moment_groups(T) groups of G
// chunks per block.
// 0 = automatic: for batches of 32 members or more the largest shape whose
// LDS fits, else 1 (more, smaller blocks).
static int g_moment_shape = 0;

extern "C" int xh_set_moment_shape(int tiles)
{
    if (tiles != 0 && tiles != 1 && tiles != 2 && tiles != 4)
        return fail(XH_EINVAL, "moment-form block shape must be 0 (automatic), 1, 2 or 4");
    g_moment_shape = tiles;
    return XH_OK;
}

// dynamic LDS of k_moment_fold<R, ATM, T, T> (mirrors its carve-up; the
// coefficient area also holds the 8 x 256 FP64 partials summed at the end)
static size_t moment_lds(int N_L, int S_max, size_t rb, int T, int atm)
{
    if (N_L > 256) return (size_t)-1;    // one list entry per thread of four waves
    const size_t NL = (size_t)N_L, NI = NL - 1, G = (size_t)T;
    size_t off = ((atm == 2 ? 4 : 3) * NL + (size_t)S_max + 8) * 8 + (size_t)S_max * rb + G * NL * 4 +
                 ((NL + 7) / 8) * 4 + 64 * 4 + NL * 2 + (3 + (size_t)T) * NL * 2;
    off = (off + 15) & ~(size_t)15;
    size_t body = G * NL * 64 * rb + (size_t)T * NI * 64 * rb + std::max(G * NI * 16, G * (size_t)T * 256) * rb;
    if (T * T < 8) body = std::max(body, (size_t)8 * 4 * 64 * 8);
    return off + body;
}

// parallels per fold block for batches of 32 members or more (the moment
// form and the energy lanes; partial sums meet by atomics); 0 = automatic
// (every parallel of a member in one moment-form block, 8 per lane block)
static int g_fold_rings = 0;

extern "C" int xh_set_fold_rings(int rings)
{
    if (rings < 0) return fail(XH_EINVAL, "parallels per fold block must be >= 0");
    g_fold_rings = rings;
    return XH_OK;
}

// The phase interpolant (xpsi/tools/core.pyx:21-116), a process-wide choice
// as in the reference: 0 periodic Akima (default), 1 Steffen, 2 periodic
// cubic spline.  The drop-ins read it at each call; the batched pipeline's
// folds are periodic Akima and its creation refuses another choice.
static int g_phase_interp = 0;

extern "C" int xh_set_phase_interpolant(int kind)
{
    if (kind < 0 || kind > 2)
        return fail(XH_EINVAL, "phase interpolant must be 0 (Akima, periodic), 1 (Steffen) or 2 (cubic, periodic)");
    g_phase_interp = kind;
    return XH_OK;
}

extern "C" int xh_get_phase_interpolant(void) { return g_phase_interp; }

// dynamic LDS of the periodic cubic spline's tables for a marginal-family
// launch: 3 x (the components' phase counts) + one component's work
static int cspline_tab_lds(const MarginalArgs& A, hipStream_t s, size_t* out)
{
    *out = 0;
    if (A.interp != 2) return XH_OK;
    int P_tot = A.n_comp * A.N_P, max_np = A.N_P;
    if (A.comp_np) {
        int h[32];
        HIPCHK(hipMemcpyAsync(h, A.comp_np, (size_t)A.n_comp * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        P_tot = 0;
        max_np = 0;
        for (int c = 0; c < A.n_comp; c++) {
            if (h[c] > 1 && h[c] < 4) return fail(XH_EINVAL, "the periodic cubic spline needs 4 or more phases");
            P_tot += h[c];
            max_np = std::max(max_np, h[c]);
        }
    } else if (A.N_P > 1 && A.N_P < 4) {
        return fail(XH_EINVAL, "the periodic cubic spline needs 4 or more phases");
    }
    *out = (3 * (size_t)P_tot + (size_t)max_np) * 8;
    static bool attr = false;
    if (!attr) {
        HIPCHK(hipFuncSetAttribute((const void*)&k_marginal, hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024));
        HIPCHK(hipFuncSetAttribute((const void*)&k_poisson_bg, hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024));
        HIPCHK(hipFuncSetAttribute((const void*)&k_star_rate, hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024));
        HIPCHK(hipFuncSetAttribute((const void*)&k_gauss_bg, hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024));
        HIPCHK(hipFuncSetAttribute((const void*)&k_phase_integrate, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   159 * 1024));
        attr = true;
    }
    if (*out + 2 * (size_t)A.N_b * 8 > 159 * 1024)
        return fail(XH_EINVAL, "the periodic cubic spline's tables do not fit one block's LDS");
    return XH_OK;
}

// intervals per FP32 accumulation chain of the wave fold before the chain's
// sums are added to the FP64 flux (0 = one chain per parallel).  Longer chains
// round more; their error is shared by neighbouring bins, so it shows in the
// likelihood as a small scale error.
static int g_wave_chain = 8;

extern "C" int xh_set_wave_fold_chain(int intervals)
{
    if (intervals < 0) return fail(XH_EINVAL, "wave-fold chain length must be >= 0");
    g_wave_chain = intervals;
    return XH_OK;
}

// parallels per wavefront of the wave fold (0 = automatic: 1)
static int g_wave_rings = 0;

// the wave fold with one block per parallel whose energy chunks share the cell
// walk (k_wave_fold_shared), where the energies need 2 to 4 chunks
static int g_wave_shared = 0;

extern "C" int xh_set_wave_fold_shared(int on)
{
    g_wave_shared = on ? 1 : 0;
    return XH_OK;
}

extern "C" int xh_set_wave_fold_rings(int rings)
{
    if (rings < 0) return fail(XH_EINVAL, "parallels per wave-fold wavefront must be >= 0");
    g_wave_rings = rings;
    return XH_OK;
}

// the block shape of a launch over `members` members
static int moment_shape(int N_L, int S_max, size_t rb, int atm, int members)
{
    if (g_moment_shape > 0) {
        int T = g_moment_shape;
        while (T > 1 && moment_lds(N_L, S_max, rb, T, atm) > kMomentLdsMax) T >>= 1;
        return T;
    }
    if (members < 32) return 1;
    for (int T = 4; T > 1; T >>= 1)
        if (moment_lds(N_L, S_max, rb, T, atm) <= kMomentLdsMax) return T;
    return 1;
}

static dim3 moment_grid(int N_E, int N_P, int T, int ring_blocks, int members)
{
    const int epb = 16 * T * moment_groups(T), ppb = 16 * T;
    return dim3((unsigned)(((N_E + epb - 1) / epb) * ((N_P + ppb - 1) / ppb)), (unsigned)ring_blocks,
                (unsigned)members);
}

template <typename R, int ATM, int T, int OCC>
static hipError_t launch_moment_t(dim3 g, size_t lds, hipStream_t s, const Member* m, Layout L, FoldArgs F)
{
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void*)&k_moment_fold<R, ATM, T, T, OCC>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMomentLdsMax);
        if (e != hipSuccess) return e;
        attr = true;
    }
    if (g.x == 0) return hipSuccess;     // configuration only
    hipLaunchKernelGGL((k_moment_fold<R, ATM, T, T, OCC>), g, dim3(512), lds, s, m, L, F);
    return hipGetLastError();
}

// shapes T < 4 with two blocks per CU in LDS use the variant whose register
// budget is sized for four waves per SIMD
template <typename R, int ATM, int T>
static hipError_t launch_moment_o(dim3 g, size_t lds, hipStream_t s, const Member* m, Layout L, FoldArgs F)
{
    if constexpr (T < 4) {
        if (g.x == 0) {
            hipError_t e = launch_moment_t<R, ATM, T, 2>(g, 0, s, m, L, F);
            return e != hipSuccess ? e : launch_moment_t<R, ATM, T, 4>(g, 0, s, m, L, F);
        }
        if (2 * lds <= kMomentLdsMax) return launch_moment_t<R, ATM, T, 4>(g, lds, s, m, L, F);
    }
    return launch_moment_t<R, ATM, T, 2>(g, lds, s, m, L, F);
}

template <typename R, int ATM>
static hipError_t launch_moment_a(int T, dim3 g, size_t lds, hipStream_t s, const Member* m, Layout L, FoldArgs F)
{
    if (T == 4) return launch_moment_o<R, ATM, 4>(g, lds, s, m, L, F);
    if (T == 2) return launch_moment_o<R, ATM, 2>(g, lds, s, m, L, F);
    return launch_moment_o<R, ATM, 1>(g, lds, s, m, L, F);
}

// T = 0: configure every shape (no launch)
static hipError_t launch_moment(bool fp32, int T, dim3 g, size_t lds, hipStream_t s, const Member* m, Layout L,
                                FoldArgs F)
{
    if (lds > kMomentLdsMax) return hipErrorInvalidValue;
    if (T == 0) {
        for (int t = 1; t <= 4; t <<= 1) {
            hipError_t e = launch_moment(fp32, t, dim3(0, 1, 1), 0, s, m, L, F);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    if (fp32) return L.atm_ext == 2 ? launch_moment_a<float, 2>(T, g, lds, s, m, L, F)
                                    : launch_moment_a<float, 1>(T, g, lds, s, m, L, F);
    return L.atm_ext == 2 ? launch_moment_a<double, 2>(T, g, lds, s, m, L, F)
                          : launch_moment_a<double, 1>(T, g, lds, s, m, L, F);
}


// dynamic LDS of k_else_integrate: ray tables, alternative tables, slots
static size_t else_lds(const Layout& Le)
{
    return (4 * (size_t)Le.N_R + 2 * (size_t)Le.S_max * 8 + (size_t)Le.S_max) * 8;
}

static hipError_t fill(double* p, size_t n, double v, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
static hipError_t fill_int(int* p, size_t n, int v, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// ============================================================= drop-ins

extern "C" size_t xh_integrate_workspace_bytes(int S, int N_R, int N_L, int N_E, int N_P)
{
    (void)N_R; (void)N_E; (void)N_P;
    const int IO = 8;
    size_t b = 256;                                  // Member
    b += 3 * (size_t)S * 8 + (size_t)S * 4;          // theta, logT, logg, ring_ok
    b += (size_t)S * IO * 5 * N_L * 8;               // geo
    b += 2 * (size_t)S * IO * 4;                     // valid, lastq
    return b + 1024;
}

extern "C" int xh_integrate(double omega, double r_s, double inclination, int S, const double* cellArea,
                            const double* radial, const double* r_s_over_r, const double* theta,
                            const double* phi, const double* cellParams, const int* radiates, int N_R,
                            const double* deflection, const double* cos_alpha, const double* lag,
                            const double* maxDeflection, const double* cos_gamma, int N_E,
                            const double* energies, int N_L, const double* leaves, int N_P,
                            const double* phases, int hot_atm_ext, const double* atm_axes, const int* atm_N,
                            const double* atm_I, int image_order_limit, int correct, double else_logT,
                            double* flux, void* workspace, size_t workspace_bytes, int* status_out,
                            void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (hot_atm_ext != 1 && hot_atm_ext != 2)
        return fail(XH_EINVAL, "xh_integrate: atmosphere extension must be 1 (blackbody) or 2 (4-D table)");
    if (hot_atm_ext == 2 && (!atm_axes || !atm_N || !atm_I))
        return fail(XH_EINVAL, "xh_integrate: the 4-D table atmosphere needs its axes and values");
    if (S < 2 || N_R < 3 || N_L < 5 || N_E < 1 || N_P < 1) return fail(XH_EINVAL, "xh_integrate: bad sizes");
    if (g_phase_interp != 0)
        return fail(XH_EINVAL, "xh_integrate folds with the periodic Akima interpolant; xh_integrate_stokes takes the "
                               "other phase interpolants (blackbody atmospheres)");
    const int EC = pick_ec(N_L);
    const bool lane_fits = lane_lds(N_L, S, 8) <= 160 * 1024;
    const bool moment_fits = moment_lds(N_L, S, 8, 1, hot_atm_ext) <= kMomentLdsMax;
    const bool lane = g_fold_kernel == 3 || (g_fold_kernel == 0 && !moment_fits && lane_fits);
    if (g_fold_kernel == 3 && !lane_fits)
        return fail(XH_EINVAL, "xh_integrate: the energy-lane fold does not fit LDS at this leaf count");
    const bool moment = !lane && (g_fold_kernel == 2 || (g_fold_kernel == 0 && moment_fits));
    if (g_fold_kernel == 2 && !moment_fits)
        return fail(XH_EINVAL, "xh_integrate: the moment-form fold does not fit LDS at this leaf count");
    if (!lane && !moment && N_P > (256 / EC) * 16)
        return fail(XH_EINVAL, "xh_integrate: too many phases for the leaf count");
    if (workspace_bytes < xh_integrate_workspace_bytes(S, N_R, N_L, N_E, N_P))
        return fail(XH_EINVAL, "xh_integrate: workspace too small");
    int hN[4] = {0, 0, 0, 0};
    if (hot_atm_ext == 2) {
        // axis lengths first, in stream order, before anything is enqueued
        HIPCHK(hipMemcpyAsync(hN, atm_N, sizeof(hN), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int d = 0; d < 4; d++)
            if (hN[d] < 4) return fail(XH_EINVAL, "xh_integrate: table axes need at least 4 nodes");
    }
    const int IO = 8;
    char* w = (char*)workspace;
    Member* d_m = (Member*)w; w += 256;
    double* d_theta = (double*)w; w += (size_t)S * 8;
    double* d_logT = (double*)w; w += (size_t)S * 8;
    double* d_logg = (double*)w; w += (size_t)S * 8;
    int* d_ok = (int*)w; w += (size_t)S * 4;
    w = (char*)(((uintptr_t)w + 255) & ~(uintptr_t)255);
    double* d_geo = (double*)w; w += (size_t)S * IO * 5 * N_L * 8;
    int* d_valid = (int*)w; w += (size_t)S * IO * 4;
    int* d_lastq = (int*)w;
    Member hm{};
    hm.Omega = omega; hm.r_s = r_s; hm.inclination = inclination; hm.active = 1; hm.S = S;
    hm.correct = correct ? 1 : 0; hm.else_logT = else_logT;
    HIPCHK(hipMemcpyAsync(d_m, &hm, sizeof(Member), hipMemcpyHostToDevice, s));
    Layout L{S, N_R, N_L, N_P, N_E, IO, image_order_limit > 0 ? image_order_limit : 0, hot_atm_ext};
    hipLaunchKernelGGL(k_ring_prep, dim3(1), dim3(256), 0, s, S, theta, cellParams, radiates, d_theta, d_ok, d_logT,
                       d_logg);
    HIPCHK(hipGetLastError());
    RingArgs RA{d_theta, radial, r_s_over_r, cos_gamma, maxDeflection, deflection, cos_alpha, lag, leaves, d_ok,
                d_geo, d_valid, d_lastq};
    hipLaunchKernelGGL(k_geometry<0>, dim3(S, 1, 1), dim3(128), geo_lds(N_R, N_L), s, d_m, L, RA);
    HIPCHK(hipGetLastError());
    FoldArgs FA{d_geo, d_valid, cellArea, radiates, phi, S, d_logT, d_logg, d_lastq, atm_axes, atm_I,
                {0, 0, 0, 0}, energies, phases, flux, S};
    if (hot_atm_ext == 2)
        for (int d = 0; d < 4; d++) FA.atm_N[d] = hN[d];
    if (lane) {
        const int nrb = std::min(S, 8), rb = (S + nrb - 1) / nrb;
        FA.ring_block = rb;
        HIPCHK(fill(flux, (size_t)N_E * N_P, 0.0, s));
        const int nq = g_lane_nq;
        HIPCHK(launch_lane(false, nq, dim3(0, 1, 1), 0, s, d_m, L, FA));
        HIPCHK(launch_lane(false, nq, lane_grid(N_E, N_P, nq, (S + rb - 1) / rb, 1), lane_lds(N_L, S, 8), s, d_m, L,
                           FA));
    } else if (moment) {
        // parallels split over up to 8 blocks per tile; partial sums meet by atomics
        const int nrb = std::min(S, 8), rb = (S + nrb - 1) / nrb;
        FA.ring_block = rb;
        HIPCHK(fill(flux, (size_t)N_E * N_P, 0.0, s));
        const int T = moment_shape(N_L, S, 8, hot_atm_ext, 1);
        HIPCHK(launch_moment(false, T, dim3(0, 1, 1), 0, s, d_m, L, FA));
        HIPCHK(launch_moment(false, T, moment_grid(N_E, N_P, T, (S + rb - 1) / rb, 1), moment_lds(N_L, S, 8, T, hot_atm_ext), s,
                             d_m, L, FA));
    } else {
        HIPCHK(launch_fold(EC, dim3((N_E + EC - 1) / EC, 1, 1), fold_lds(N_L, EC, S), s, d_m, L, FA));
    }
    Member back{};
    HIPCHK(hipMemcpyAsync(&back, d_m, sizeof(Member), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (back.status == 4) return fail(XH_ERANGE, "xh_integrate: more image orders than supported (8)");
    *status_out = back.status ? 1 : 0;
    return XH_OK;
}

extern "C" size_t xh_integrate_stokes_workspace_bytes(int S, int N_R, int N_L)
{
    (void)N_R;
    const int IO = 8;
    size_t b = 256;                                  // Member
    b += 3 * (size_t)S * 8 + (size_t)S * 4;          // theta, logT, logg, ring_ok
    b = (b + 255) & ~(size_t)255;
    b += (size_t)S * IO * 5 * N_L * 8;               // geo
    b += (size_t)S * IO * 2 * N_L * 8;               // cos 2chi, sin 2chi
    b += 2 * (size_t)S * IO * 4;                     // valid, lastq
    return b + 1024;
}

extern "C" int xh_integrate_stokes(double omega, double inclination, int S, const double* cellArea,
                                   const double* radial, const double* r_s_over_r, const double* theta,
                                   const double* phi, const double* cellParams, int n_params, const int* radiates,
                                   int N_R, const double* deflection, const double* cos_alpha, const double* lag,
                                   const double* maxDeflection, const double* cos_gamma, int N_E,
                                   const double* energies, int N_L, const double* leaves, int N_P,
                                   const double* phases, int atm_ext, int beam_opt, int image_order_limit,
                                   double R_in, const double* correction, double* flux_I, double* flux_Q,
                                   double* flux_U, void* workspace, size_t workspace_bytes, int* status_out,
                                   void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (S < 2 || N_R < 3 || N_L < 5 || N_E < 1 || N_P < 1)
        return fail(XH_EINVAL, "xh_integrate_stokes: bad sizes");
    if (atm_ext != 1 && atm_ext != 3)
        return fail(XH_EINVAL, "xh_integrate_stokes: atmosphere extension must be 1 (blackbody) or 3 (burst blackbody)");
    if (beam_opt < 0 || beam_opt > 3 || (beam_opt == 3 && atm_ext != 1))
        return fail(XH_EINVAL, "xh_integrate_stokes: beaming option must be 0-3 (3 with the blackbody only)");
    if (n_params < 2 || (beam_opt != 0 && n_params < 7))
        return fail(XH_EINVAL, "xh_integrate_stokes: beaming needs 7 parameters per cell (log T, log g, a, b, c, d, n_mu)");
    if ((flux_Q == nullptr) != (flux_U == nullptr))
        return fail(XH_EINVAL, "xh_integrate_stokes: Q and U are requested together");
    const int lds_max = 159 * 1024;
    if (stokes_fold_lds(N_L) > (size_t)lds_max || geo_lds(N_R, N_L) > (size_t)lds_max)
        return fail(XH_EINVAL, "xh_integrate_stokes: too many leaves or rays for one block's LDS");
    const int nt = N_P <= 64 ? 64 : (N_P <= 128 ? 128 : 256);
    const long tiles = (long)N_E * ((N_P + nt - 1) / nt);
    if (tiles > 65535) return fail(XH_EINVAL, "xh_integrate_stokes: too many energies x phase tiles");
    if (workspace_bytes < xh_integrate_stokes_workspace_bytes(S, N_R, N_L))
        return fail(XH_EINVAL, "xh_integrate_stokes: workspace too small");
    static bool attr = false;
    if (!attr) {
        HIPCHK(hipFuncSetAttribute((const void*)&k_geometry<1>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
        HIPCHK(hipFuncSetAttribute((const void*)&k_stokes_fold, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
        attr = true;
    }
    const int IO = 8;
    char* w = (char*)workspace;
    Member* d_m = (Member*)w; w += 256;
    double* d_theta = (double*)w; w += (size_t)S * 8;
    double* d_logT = (double*)w; w += (size_t)S * 8;
    double* d_logg = (double*)w; w += (size_t)S * 8;
    int* d_ok = (int*)w; w += (size_t)S * 4;
    w = (char*)(((uintptr_t)w + 255) & ~(uintptr_t)255);
    double* d_geo = (double*)w; w += (size_t)S * IO * 5 * N_L * 8;
    double* d_pol = (double*)w; w += (size_t)S * IO * 2 * N_L * 8;
    int* d_valid = (int*)w; w += (size_t)S * IO * 4;
    int* d_lastq = (int*)w;
    Member hm{};
    hm.Omega = omega; hm.inclination = inclination; hm.active = 1; hm.S = S;
    HIPCHK(hipMemcpyAsync(d_m, &hm, sizeof(Member), hipMemcpyHostToDevice, s));
    Layout L{S, N_R, N_L, N_P, N_E, IO, image_order_limit > 0 ? image_order_limit : 0, atm_ext};
    hipLaunchKernelGGL(k_ring_prep, dim3(1), dim3(256), 0, s, S, theta, cellParams, radiates, d_theta, d_ok, d_logT,
                       d_logg);
    HIPCHK(hipGetLastError());
    RingArgs RA{d_theta, radial, r_s_over_r, cos_gamma, maxDeflection, deflection, cos_alpha, lag, leaves, d_ok,
                d_geo, d_valid, d_lastq};
    RA.disk = R_in < 1.0e6 ? 1 : 0;
    RA.R_in = R_in;
    RA.pol = flux_Q ? d_pol : nullptr;
    hipLaunchKernelGGL(k_geometry<1>, dim3(S, 1, 1), dim3(128), geo_lds(N_R, N_L), s, d_m, L, RA);
    HIPCHK(hipGetLastError());
    HIPCHK(fill(flux_I, (size_t)N_E * N_P, 0.0, s));
    if (flux_Q) {
        HIPCHK(fill(flux_Q, (size_t)N_E * N_P, 0.0, s));
        HIPCHK(fill(flux_U, (size_t)N_E * N_P, 0.0, s));
    }
    StokesArgs SA{d_geo, RA.pol, d_valid, cellArea, radiates, phi, cellParams, n_params, correction, atm_ext,
                  beam_opt, energies, phases, flux_I, flux_Q, flux_U};
    SA.interp = g_phase_interp;
    hipLaunchKernelGGL(k_stokes_fold, dim3(S, IO, (unsigned)tiles), dim3(nt), stokes_fold_lds(N_L), s, d_m, L, SA);
    HIPCHK(hipGetLastError());
    Member back{};
    HIPCHK(hipMemcpyAsync(&back, d_m, sizeof(Member), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (back.status == 4) return fail(XH_ERANGE, "xh_integrate_stokes: more image orders than supported (8)");
    *status_out = back.status ? 1 : 0;
    return XH_OK;
}

extern "C" size_t xh_integrate_general_workspace_bytes(int S, int N_L)
{
    const int IO = 8;
    size_t b = 256;                                  // Member
    b += 3 * (size_t)S * 8 + (size_t)S * 4;          // theta, logT, logg, ring_ok
    b = (b + 255) & ~(size_t)255;
    b += (size_t)S * IO * 4 * N_L * 8;               // geo
    b += (size_t)S * IO * 2 * 8;                     // polarisation angle per parallel and image
    b += (size_t)S * IO * 4;                         // valid
    return b + 1024;
}

extern "C" int xh_integrate_general(double omega, double inclination, int S, const double* cellArea,
                                    const double* radial, const double* r_s_over_r, const double* theta,
                                    const double* phi, const double* cellParams, int n_params, const int* radiates,
                                    int N_R, const double* deflection, const double* cos_alpha, const double* lag,
                                    const double* maxDeflection, const double* cos_gamma, int N_E,
                                    const double* energies, int N_L, const double* leaves, int N_P,
                                    const double* phases, int image_order_limit, int atm_ext, int beam_opt,
                                    const double* correction, double* flux, double* flux_Q, double* flux_U,
                                    void* workspace, size_t workspace_bytes, int* status_out, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (S < 2 || N_R < 3 || N_L < 5 || N_E < 1 || N_P < 1)
        return fail(XH_EINVAL, "xh_integrate_general: bad sizes");
    if (atm_ext != 1 && atm_ext != 3)
        return fail(XH_EINVAL, "xh_integrate_general: atmosphere extension must be 1 (blackbody) or 3 (burst blackbody)");
    if (beam_opt < 0 || beam_opt > 3 || (beam_opt == 3 && atm_ext != 1))
        return fail(XH_EINVAL, "xh_integrate_general: beaming option must be 0-3 (3 with the blackbody only)");
    if (n_params < 2 || (beam_opt != 0 && n_params < 7))
        return fail(XH_EINVAL, "xh_integrate_general: beaming needs 7 parameters per cell (log T, log g, a, b, c, d, n_mu)");
    if ((flux_Q == nullptr) != (flux_U == nullptr))
        return fail(XH_EINVAL, "xh_integrate_general: Q and U are requested together");
    // the geometry kernel keeps a few static LDS words beside its dynamic tables
    const int gen_lds_max = 159 * 1024;
    if (general_fold_lds(N_L) > (size_t)gen_lds_max || general_geo_lds(N_R, N_L) > (size_t)gen_lds_max)
        return fail(XH_EINVAL, "xh_integrate_general: too many leaves or rays for one block's LDS");
    if ((size_t)N_E > 65535u * 32u || (size_t)((N_E + 31) / 32) * ((N_P + 63) / 64) > 65535u)
        return fail(XH_EINVAL, "xh_integrate_general: too many energies x phases");
    if (workspace_bytes < xh_integrate_general_workspace_bytes(S, N_L))
        return fail(XH_EINVAL, "xh_integrate_general: workspace too small");
    static bool attr = false;
    if (!attr) {
        HIPCHK(hipFuncSetAttribute((const void*)&k_general_geometry, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   gen_lds_max));
        HIPCHK(hipFuncSetAttribute((const void*)&k_general_fold, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   gen_lds_max));
        attr = true;
    }
    const int IO = 8;
    char* w = (char*)workspace;
    Member* d_m = (Member*)w; w += 256;
    double* d_theta = (double*)w; w += (size_t)S * 8;
    double* d_logT = (double*)w; w += (size_t)S * 8;
    double* d_logg = (double*)w; w += (size_t)S * 8;
    int* d_ok = (int*)w; w += (size_t)S * 4;
    w = (char*)(((uintptr_t)w + 255) & ~(uintptr_t)255);
    double* d_geo = (double*)w; w += (size_t)S * IO * 4 * N_L * 8;
    double* d_pol = (double*)w; w += (size_t)S * IO * 2 * 8;
    int* d_valid = (int*)w;
    Member hm{};
    hm.Omega = omega; hm.inclination = inclination; hm.active = 1; hm.S = S;
    HIPCHK(hipMemcpyAsync(d_m, &hm, sizeof(Member), hipMemcpyHostToDevice, s));
    Layout L{S, N_R, N_L, N_P, N_E, IO, image_order_limit > 0 ? image_order_limit : 0, atm_ext};
    hipLaunchKernelGGL(k_ring_prep, dim3(1), dim3(256), 0, s, S, theta, cellParams, radiates, d_theta, d_ok, d_logT,
                       d_logg);
    HIPCHK(hipGetLastError());
    RingArgs RA{d_theta, radial, r_s_over_r, cos_gamma, maxDeflection, deflection, cos_alpha, lag, leaves, d_ok,
                d_geo, d_valid, nullptr};
    RA.pol = flux_Q ? d_pol : nullptr;     // [ring][image][2] here
    hipLaunchKernelGGL(k_general_geometry, dim3(S), dim3(128), general_geo_lds(N_R, N_L), s, d_m, L, RA);
    HIPCHK(hipGetLastError());
    HIPCHK(fill(flux, (size_t)N_E * N_P, 0.0, s));
    if (flux_Q) {
        HIPCHK(fill(flux_Q, (size_t)N_E * N_P, 0.0, s));
        HIPCHK(fill(flux_U, (size_t)N_E * N_P, 0.0, s));
    }
    GeneralArgs GA{d_geo, d_valid, cellArea, radiates, phi, cellParams, correction, energies, phases, flux,
                   n_params, atm_ext, beam_opt, RA.pol, flux_Q, flux_U};
    GA.interp = g_phase_interp;
    const int tiles = ((N_E + 31) / 32) * ((N_P + 63) / 64);
    hipLaunchKernelGGL(k_general_fold, dim3(S, IO, tiles), dim3(256), general_fold_lds(N_L), s, d_m, L, GA);
    HIPCHK(hipGetLastError());
    Member back{};
    HIPCHK(hipMemcpyAsync(&back, d_m, sizeof(Member), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (back.status == 4) return fail(XH_ERANGE, "xh_integrate_general: more image orders than supported (8)");
    *status_out = back.status ? 1 : 0;
    return XH_OK;
}

extern "C" size_t xh_integrate_time_invariant_workspace_bytes(int S, int N_E)
{
    return 256 + (size_t)S * N_E * 8 + 256;
}

extern "C" int xh_integrate_time_invariant(double omega, double r_s, double inclination, int S, double cellArea,
                                           const double* radial, const double* colat, const double* phi,
                                           const double* cellParams, int param_stride, int N_R,
                                           const double* deflection, const double* cos_alpha,
                                           const double* maxDeflection, const double* cos_gamma, int N_E,
                                           const double* energies, int image_order_limit, double* flux,
                                           void* workspace, size_t workspace_bytes, int* status_out, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (S < 2 || N_R < 3 || N_E < 1 || param_stride < 1)
        return fail(XH_EINVAL, "xh_integrate_time_invariant: bad sizes");
    if (workspace_bytes < xh_integrate_time_invariant_workspace_bytes(S, N_E))
        return fail(XH_EINVAL, "xh_integrate_time_invariant: workspace too small");
    Layout Le{S, N_R, 0, 0, N_E, 8, image_order_limit > 0 ? image_order_limit : 0, 1};
    if (else_lds(Le) > 64 * 1024)
        return fail(XH_EINVAL, "xh_integrate_time_invariant: too many rays or cells for one block's LDS");
    if (image_order_limit > 8) return fail(XH_EINVAL, "xh_integrate_time_invariant: at most 8 image orders");
    char* w = (char*)workspace;
    Member* d_m = (Member*)w; w += 256;
    double* d_part = (double*)w;
    Member hm{};
    hm.Omega = omega; hm.r_s = r_s; hm.inclination = inclination; hm.active = 1; hm.S = S;
    HIPCHK(hipMemcpyAsync(d_m, &hm, sizeof(Member), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_else_integrate, dim3(S, 1), dim3(256), else_lds(Le), s, d_m, Le, colat, radial, cos_gamma,
                       maxDeflection, deflection, cos_alpha, energies, N_E, phi, S, cellParams, param_stride, d_part);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_else_finalize, dim3((N_E + 255) / 256), dim3(256), 0, s, (const Member*)d_m, Le, d_part,
                       cellArea, energies, N_E, flux);
    HIPCHK(hipGetLastError());
    Member back{};
    HIPCHK(hipMemcpyAsync(&back, d_m, sizeof(Member), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (back.status == 4) return fail(XH_ERANGE, "xh_integrate_time_invariant: more than 8 image orders");
    *status_out = back.status ? 1 : 0;
    return XH_OK;
}

extern "C" size_t xh_compute_rays_workspace_bytes(int N_u, int N_R)
{
    return 256 + (size_t)N_u * 8 + (size_t)N_u * N_R * 4 + 64 + 256;
}

extern "C" int xh_compute_rays(int N_u, double r_s, const double* r_s_over_R, double* maxAlpha, int N_R,
                               double* deflection, double* cos_alpha, double* lag, double* maxDeflection,
                               void* workspace, size_t workspace_bytes, int* terminate_out, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (N_u < 1 || N_R < 3) return fail(XH_EINVAL, "xh_compute_rays: bad sizes");
    if (workspace_bytes < xh_compute_rays_workspace_bytes(N_u, N_R))
        return fail(XH_EINVAL, "xh_compute_rays: workspace too small");
    char* w = (char*)workspace;
    Member* d_m = (Member*)w; w += 256;
    double* d_rr = (double*)w; w += (size_t)N_u * 8;
    int* d_work = (int*)w; w += (size_t)N_u * N_R * 4;
    w = (char*)(((uintptr_t)w + 63) & ~(uintptr_t)63);
    int* d_count = (int*)w;
    Member hm{};
    hm.r_s = r_s; hm.active = 1; hm.S = N_u;
    HIPCHK(hipMemcpyAsync(d_m, &hm, sizeof(Member), hipMemcpyHostToDevice, s));
    // radial coordinates r = r_s / u, as the kernels take them
    hipLaunchKernelGGL(k_over, dim3((N_u + 255) / 256), dim3(256), 0, s, d_rr, r_s_over_R, (size_t)N_u, r_s);
    HIPCHK(hipGetLastError());
    HIPCHK(fill_int(d_count, 2, 0, s));
    Layout L{N_u, N_R, 0, 0, 0, 8, 0, 1};
    hipLaunchKernelGGL(k_rays_lane, dim3((unsigned)(((size_t)N_u * (N_R - 1) + 63) / 64), 1, 1), dim3(64), 0, s,
                       (const Member*)d_m, L,
                       (const double*)d_rr, (const double*)maxAlpha, cos_alpha, deflection, lag, d_work, d_count,
                       r_s_over_R);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_rays, dim3(kRayWaveBlocks), dim3(256), 0, s, (const Member*)d_m, L, (const double*)d_rr,
                       (const double*)maxAlpha, deflection, lag, (const int*)d_work, (const int*)d_count,
                       r_s_over_R);
    HIPCHK(hipGetLastError());
    // bad rays repaired in the reference's order, maxAlpha clamped in place
    // (rays.pyx:310-369); the lag is returned uncalibrated
    hipLaunchKernelGGL(k_rays_fix, dim3(1), dim3(256), 0, s, d_m, L, (const double*)d_rr, maxAlpha, deflection, lag,
                       maxDeflection, 0, r_s_over_R);
    HIPCHK(hipGetLastError());
    Member back{};
    HIPCHK(hipMemcpyAsync(&back, d_m, sizeof(Member), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *terminate_out = back.status == 1 ? 1 : 0;
    return XH_OK;
}

extern "C" size_t xh_allocate_cells_workspace_bytes(void) { return 256 + 2 * 256 + 256; }

extern "C" int xh_allocate_cells(int num_cells, int min_sqrt_num_cells, int max_sqrt_num_cells, double mass,
                                 double r_s, double R_eq, double Omega, double zeta, double epsilon,
                                 int star_shape_ind, double superRadius, double cedeRadius,
                                 double superColatitude, double cedeColatitude, double superAzimuth,
                                 double holeColatitude, double holeRadius, double holeAzimuth, double* out4,
                                 void* workspace, size_t workspace_bytes, void* stream)
{
    (void)mass; (void)r_s;
    hipStream_t s = (hipStream_t)stream;
    if (num_cells < 1 || min_sqrt_num_cells < 1 || max_sqrt_num_cells < min_sqrt_num_cells)
        return fail(XH_EINVAL, "xh_allocate_cells: bad cell counts");
    if (workspace_bytes < xh_allocate_cells_workspace_bytes())
        return fail(XH_EINVAL, "xh_allocate_cells: workspace too small");
    char* w = (char*)workspace;
    RegionAlloc* d_r = (RegionAlloc*)w; w += 256;
    Member* d_m = (Member*)w;
    RegionAlloc hr{};
    hr.superR = superRadius; hr.cedeR = cedeRadius; hr.superColat = superColatitude; hr.cedeColat = cedeColatitude;
    hr.superAzi = superAzimuth; hr.holeColat = holeColatitude; hr.holeR = holeRadius; hr.holeAzi = holeAzimuth;
    hr.R_eq = R_eq; hr.Omega = Omega; hr.zeta = zeta; hr.epsilon = epsilon; hr.shape = star_shape_ind;
    hr.num_cells = num_cells; hr.min_sqrt = min_sqrt_num_cells; hr.max_sqrt = max_sqrt_num_cells;
    hr.member0 = 0; hr.member1 = eq12_host(cedeRadius, 0.0) ? -1 : 1;
    Member hm[2] = {};
    hm[0].active = 1; hm[1].active = 1;
    HIPCHK(hipMemcpyAsync(d_r, &hr, sizeof(RegionAlloc), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_m, hm, 2 * sizeof(Member), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_allocate, dim3(1), dim3(256), 0, s, (const RegionAlloc*)d_r, d_m, max_sqrt_num_cells);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hm, d_m, 2 * sizeof(Member), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    out4[0] = (double)hm[0].S;
    out4[1] = hm[0].numCell;
    out4[2] = hr.member1 >= 0 ? (double)hm[1].S : 0.0;
    out4[3] = hr.member1 >= 0 ? hm[1].numCell : 0.0;
    return XH_OK;
}

// The cell mesh after the area table: parallels' limits, cell kinds and the
// boundary-cell list, then the boundary cells (k_ring_bounds, k_mesh_classify,
// k_mesh_cells); blist holds members x S x S entries, every cell's worst case.
static hipError_t launch_mesh_cells(const Member* m, const double* tab, const double* tot, int table_n, Layout L,
                                    int nm, int S, double* theta, double* phi, double* r, double* cellArea,
                                    double* maxAlpha, double* cos_gamma, double* effGrav, double* ringlim,
                                    int2* blist, int* bcount, hipStream_t s)
{
    hipError_t e = fill_int(bcount, 1, 0, s);
    if (e != hipSuccess) return e;
    const dim3 g((unsigned)((S + 3) / 4), (unsigned)nm);
    hipLaunchKernelGGL(k_ring_bounds, g, dim3(256), 0, s, m, tab, tot, table_n, L, theta, r, maxAlpha, cos_gamma,
                       effGrav, ringlim);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_mesh_classify, dim3((unsigned)((S + 7) / 8), (unsigned)nm), dim3(512), 0, s, m, tot, L,
                       (const double*)ringlim, phi, cellArea, blist, bcount);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const unsigned nb = (unsigned)std::max(1, std::min(3072, nm * S / 8));   // three waves per SIMD
    hipLaunchKernelGGL(k_mesh_cells, dim3(nb), dim3(256), 0, s, m, tot, L, (const double*)ringlim, cellArea,
                       (const int2*)blist, (const int*)bcount);
    return hipGetLastError();
}

extern "C" size_t xh_cell_mesh_workspace_bytes(int S)
{
    // member, area table, totals, colatitudes, azimuths, parallel limits,
    // boundary-cell list and its count
    return 256 + (1000 + 64) * 8 + 2 * (size_t)S * 8 + (size_t)S * 3 * 8 + (size_t)S * S * 8 + 256 + 256;
}

extern "C" int xh_construct_cell_mesh(int polar, double numCell, int S, double mass, double r_s, double R_eq,
                                      double zeta, double epsilon, int star_shape_ind, double cedeRadius,
                                      double cedeColatitude, double superRadius, double superColatitude,
                                      double superAzimuth, double* theta, double* phi, double* r, double* cellArea,
                                      double* maxEmissionAngle, double* cos_gamma, double* effGrav, void* workspace,
                                      size_t workspace_bytes, int* status_out, void* stream)
{
    (void)mass;
    hipStream_t s = (hipStream_t)stream;
    if (S < 2 || S > 32767) return fail(XH_EINVAL, "xh_construct_cell_mesh: bad size");
    if (workspace_bytes < xh_cell_mesh_workspace_bytes(S))
        return fail(XH_EINVAL, "xh_construct_cell_mesh: workspace too small");
    const int table_n = 1000;
    char* w = (char*)workspace;
    Member* d_m = (Member*)w; w += 256;
    double* d_tab = (double*)w; w += (size_t)table_n * 8;
    double* d_total = (double*)w; w += 64 * 8;
    double* d_th = (double*)w; w += (size_t)S * 8;
    double* d_ph = (double*)w; w += (size_t)S * 8;
    double* d_rl = (double*)w; w += (size_t)S * 3 * 8;
    int* d_bc = (int*)w; w += 256;
    int2* d_bl = (int2*)w;
    Member hm{};
    hm.R_eq = R_eq; hm.epsilon = epsilon; hm.zeta = zeta; hm.r_s = r_s; hm.shape = star_shape_ind; hm.active = 1;
    hm.cedeR = cedeRadius; hm.cedeColat = cedeColatitude; hm.superR = superRadius; hm.superColat = superColatitude;
    hm.superAzi = superAzimuth; hm.polar = polar ? 1 : 0; hm.S = S; hm.numCell = numCell;
    HIPCHK(hipMemcpyAsync(d_m, &hm, sizeof(Member), hipMemcpyHostToDevice, s));
    HIPCHK(fill(cellArea, (size_t)S * S, 0.0, s));
    Layout L{S, 3, 0, 0, 0, 8, 0, 1};
    hipLaunchKernelGGL(k_mesh_table, dim3(1), dim3(1024), 0, s, (const Member*)d_m, d_tab,
                       d_total, table_n);
    HIPCHK(hipGetLastError());
    HIPCHK(launch_mesh_cells(d_m, d_tab, d_total, table_n, L, 1, S, d_th, d_ph, r, cellArea, maxEmissionAngle,
                             cos_gamma, effGrav, d_rl, d_bl, d_bc, s));
    // the reference's [S, S] colatitude and azimuth arrays
    const unsigned nb = (unsigned)(((size_t)S * S + 255) / 256);
    hipLaunchKernelGGL(k_spread, dim3(nb), dim3(256), 0, s, theta, (const double*)d_th, S, S, 1);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_spread, dim3(nb), dim3(256), 0, s, phi, (const double*)d_ph, S, S, 0);
    HIPCHK(hipGetLastError());
    Member back{};
    HIPCHK(hipMemcpyAsync(&back, d_m, sizeof(Member), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *status_out = back.status;
    return XH_OK;
}

extern "C" int xh_energy_integrator(int N_E, int N_P, const double* signal, const double* logE, int N_edges,
                                    const double* logEdges, double* out, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (N_E < 5 || N_P < 1 || N_edges < 2) return fail(XH_EINVAL, "xh_energy_integrator: bad sizes");
    if (N_E > 8192) return fail(XH_EINVAL, "xh_energy_integrator: at most 8192 energies (one LDS row)");
    // no member table: the kernel integrates the signal itself (d^2 = 1)
    EnergyArgs A{signal, logE, logEdges, nullptr, nullptr, out, N_E, N_P, N_edges, 1};
    A.interp = g_phase_interp;
    if (A.interp == 2) {
        // the periodic cubic spline: one block per phase column solves it in LDS
        const size_t lds = (size_t)5 * N_E * 8;
        if (lds > 159 * 1024) return fail(XH_EINVAL, "xh_energy_integrator: at most 4070 energies with the cubic spline");
        static bool attr = false;
        if (!attr) {
            HIPCHK(hipFuncSetAttribute((const void*)&k_energy_integrate_cs, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       159 * 1024));
            attr = true;
        }
        hipLaunchKernelGGL(k_energy_integrate_cs, dim3(N_P), dim3(256), lds, s, A);
        HIPCHK(hipGetLastError());
        return XH_OK;
    }
    int n = (N_edges - 1) * N_P;
    hipLaunchKernelGGL(k_energy_integrate, dim3((n + 255) / 256, 1), dim3(256), (size_t)N_E * 8, s,
                       (const Member*)nullptr, A);
    HIPCHK(hipGetLastError());
    return XH_OK;
}

extern "C" int xh_fold(int N_ch, int N_in, int N_P, const double* M, int ld_M, const double* signal,
                       double* out, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    FoldGemmArgs A{M, signal, out, N_ch, N_in, N_P, ld_M};
    hipLaunchKernelGGL(k_fold_f64, dim3((N_P + 31) / 32, (N_ch + 31) / 32, 1), dim3(256), 0, s, A);
    HIPCHK(hipGetLastError());
    return XH_OK;
}

extern "C" int xh_precomputation(int N_ch, int N_b, const int* counts, double* out)
{
    for (int i = 0; i < N_ch; i++) {
        double acc = 0.0;
        for (int j = 0; j < N_b; j++) acc += std::lgamma((double)(unsigned int)counts[(size_t)i * N_b + j] + 1.0);
        out[i] = -acc;
    }
    return XH_OK;
}

extern "C" size_t xh_marginal_workspace_bytes(int N_ch) { return (size_t)N_ch * (8 + 4) + 64 + 256; }

// shared body of the two marginal drop-ins: comp_np (device, per component)
// or nullptr for one shared grid of N_P phases; allow: bit c lets component c
// keep negative bin integrals
static int marginal_impl(double T, int N_b, const double* phases, int N_ch, const double* counts, int n_comp,
                         const double* comps, int N_P, const int* comp_np, const double* comp_phases,
                         const double* shifts, const double* precomp, const double* support, double epsabs,
                         double epsrel, double epsilon, double sigmas, double llzero, unsigned allow, double slim,
                         const double* background, double* lnl_out, double* STAR, double* B_ML, double* B_ML_supp,
                         void* workspace, size_t workspace_bytes, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (n_comp < 1 || n_comp > 32) return fail(XH_EINVAL, "marginal likelihood: 1 to 32 components");
    if (workspace_bytes < xh_marginal_workspace_bytes(N_ch)) return fail(XH_EINVAL, "workspace too small");
    char* w = (char*)workspace;
    double* d_ll = (double*)w; w += (size_t)N_ch * 8;
    double* d_lnl = (double*)w; w += 64;
    int* d_st = (int*)w;
    MarginalArgs A{};
    A.signals = comps; A.comp_phases = comp_phases; A.comp_np = comp_np; A.shifts = shifts; A.phases = phases;
    A.counts = counts;
    A.precomp = precomp; A.support = support; A.background = background; A.request_status = nullptr;
    A.ll = d_ll; A.status = d_st; A.B_ML = B_ML; A.B_ML_supp = B_ML_supp; A.STAR = STAR;
    A.exposure = T; A.epsabs = epsabs; A.epsrel = epsrel; A.epsilon = epsilon; A.sigmas = sigmas;
    A.llzero = llzero; A.slim = slim; A.N_ch = N_ch; A.N_b = N_b; A.N_P = N_P; A.n_comp = n_comp;
    A.allow_negative = allow; A.seed = next_seed(); A.seed_ptr = nullptr;
    A.interp = g_phase_interp;
    size_t tab = 0;
    {
        const int rc = cspline_tab_lds(A, s, &tab);
        if (rc != XH_OK) return rc;
    }
    HIPCHK(fill(STAR, (size_t)N_ch * N_b, 0.0, s));
    HIPCHK(fill(B_ML, N_ch, 0.0, s));
    HIPCHK(fill(B_ML_supp, N_ch, 0.0, s));
    hipLaunchKernelGGL(k_marginal, dim3(N_ch, 1), dim3(64), 2 * N_b * 8 + tab, s, A);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_lnl_reduce, dim3(1), dim3(64), 0, s, A, (const Member*)nullptr, 0, d_lnl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(lnl_out, d_lnl, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return XH_OK;
}

static unsigned allow_mask(int n_comp, const int* allow_negative)
{
    unsigned m = 0u;
    for (int c = 0; allow_negative && c < n_comp && c < 32; c++)
        if (allow_negative[c]) m |= 1u << c;
    return m;
}

extern "C" int xh_eval_marginal_likelihood(double T, int N_b, const double* phases, int N_ch, const double* counts,
                                           int n_comp, const double* comps, int N_P, const double* comp_phases,
                                           const double* shifts, const double* precomp, const double* support,
                                           double epsabs, double epsrel, double epsilon, double sigmas,
                                           double llzero, int allow_negative, double slim,
                                           const double* background, double* lnl_out, double* STAR,
                                           double* B_ML, double* B_ML_supp, void* workspace,
                                           size_t workspace_bytes, void* stream)
{
    return marginal_impl(T, N_b, phases, N_ch, counts, n_comp, comps, N_P, nullptr, comp_phases, shifts, precomp,
                         support, epsabs, epsrel, epsilon, sigmas, llzero, allow_negative ? ~0u : 0u, slim,
                         background, lnl_out, STAR, B_ML, B_ML_supp, workspace, workspace_bytes, stream);
}

extern "C" int xh_eval_marginal_likelihood_ex(double T, int N_b, const double* phases, int N_ch,
                                              const double* counts, int n_comp, const double* comps,
                                              const int* comp_num_phases, const double* comp_phases,
                                              const double* shifts, const double* precomp, const double* support,
                                              double epsabs, double epsrel, double epsilon, double sigmas,
                                              double llzero, const int* allow_negative, double slim,
                                              const double* background, double* lnl_out, double* STAR,
                                              double* B_ML, double* B_ML_supp, void* workspace,
                                              size_t workspace_bytes, void* stream)
{
    if (!comp_num_phases) return fail(XH_EINVAL, "xh_eval_marginal_likelihood_ex: phases per component required");
    return marginal_impl(T, N_b, phases, N_ch, counts, n_comp, comps, 0, comp_num_phases, comp_phases, shifts,
                         precomp, support, epsabs, epsrel, epsilon, sigmas, llzero, allow_mask(n_comp, allow_negative),
                         slim, background, lnl_out, STAR, B_ML, B_ML_supp, workspace, workspace_bytes, stream);
}

extern "C" size_t xh_poisson_workspace_bytes(int N_ch) { return (size_t)N_ch * (8 + 4) + 64 + 256; }

static int poisson_impl(double T, int N_b, const double* phases, int N_ch, const double* counts, int n_comp,
                        const double* comps, int N_P, const int* comp_np, const double* comp_phases,
                        const double* shifts, const double* background, const double* precomp, unsigned allow,
                        double* lnl_out, double* EXPEC, void* workspace, size_t workspace_bytes, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (N_ch < 1 || N_b < 1 || n_comp < 1 || n_comp > 32 || (!comp_np && N_P < 1))
        return fail(XH_EINVAL, "xh_poisson_likelihood_given_background: bad sizes");
    if (!background) return fail(XH_EINVAL, "xh_poisson_likelihood_given_background: the background is required");
    if (workspace_bytes < xh_poisson_workspace_bytes(N_ch))
        return fail(XH_EINVAL, "xh_poisson_likelihood_given_background: workspace too small");
    char* w = (char*)workspace;
    double* d_ll = (double*)w; w += (size_t)N_ch * 8;
    double* d_lnl = (double*)w; w += 64;
    int* d_st = (int*)w;
    MarginalArgs A{};
    A.signals = comps; A.comp_phases = comp_phases; A.comp_np = comp_np; A.shifts = shifts; A.phases = phases;
    A.counts = counts;
    A.precomp = precomp; A.background = background; A.ll = d_ll; A.status = d_st; A.STAR = EXPEC;
    A.exposure = T; A.llzero = -1.0e90; A.N_ch = N_ch; A.N_b = N_b; A.N_P = N_P; A.n_comp = n_comp;
    A.allow_negative = allow; A.seed = next_seed(); A.seed_ptr = nullptr;
    A.interp = g_phase_interp;
    size_t tab = 0;
    {
        const int rc = cspline_tab_lds(A, s, &tab);
        if (rc != XH_OK) return rc;
    }
    hipLaunchKernelGGL(k_poisson_bg, dim3(N_ch, 1), dim3(64), 2 * (size_t)N_b * 8 + tab, s, A);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_poisson_reduce, dim3(1), dim3(64), 0, s, A, d_lnl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(lnl_out, d_lnl, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return XH_OK;
}

extern "C" int xh_poisson_likelihood_given_background(double T, int N_b, const double* phases, int N_ch,
                                                      const double* counts, int n_comp, const double* comps,
                                                      int N_P, const double* comp_phases, const double* shifts,
                                                      const double* background, const double* precomp,
                                                      int allow_negative, double* lnl_out, double* EXPEC,
                                                      void* workspace, size_t workspace_bytes, void* stream)
{
    return poisson_impl(T, N_b, phases, N_ch, counts, n_comp, comps, N_P, nullptr, comp_phases, shifts, background,
                        precomp, allow_negative ? ~0u : 0u, lnl_out, EXPEC, workspace, workspace_bytes, stream);
}

extern "C" int xh_poisson_likelihood_given_background_ex(double T, int N_b, const double* phases, int N_ch,
                                                         const double* counts, int n_comp, const double* comps,
                                                         const int* comp_num_phases, const double* comp_phases,
                                                         const double* shifts, const double* background,
                                                         const double* precomp, const int* allow_negative,
                                                         double* lnl_out, double* EXPEC, void* workspace,
                                                         size_t workspace_bytes, void* stream)
{
    if (!comp_num_phases)
        return fail(XH_EINVAL, "xh_poisson_likelihood_given_background_ex: phases per component required");
    return poisson_impl(T, N_b, phases, N_ch, counts, n_comp, comps, 0, comp_num_phases, comp_phases, shifts,
                        background, precomp, allow_mask(n_comp, allow_negative), lnl_out, EXPEC, workspace,
                        workspace_bytes, stream);
}

extern "C" size_t xh_gaussian_workspace_bytes(int N_ch, int N_b) { return (size_t)N_ch * N_b * 8 + 64 + 256; }

static int gaussian_impl(double T, int N_b, const double* phases, int N_ch, const double* counts,
                         const double* errors, int n_comp, const double* comps, int N_P, const int* comp_np,
                         const double* comp_phases, const double* shifts, const double* background, unsigned allow,
                         double* lnl_out, double* EXPEC, void* workspace, size_t workspace_bytes, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (N_ch < 1 || N_b < 1 || n_comp < 1 || n_comp > 32 || (!comp_np && N_P < 1))
        return fail(XH_EINVAL, "xh_gaussian_likelihood_given_background: bad sizes");
    if (!background || !errors)
        return fail(XH_EINVAL, "xh_gaussian_likelihood_given_background: the background and errors are required");
    if (workspace_bytes < xh_gaussian_workspace_bytes(N_ch, N_b))
        return fail(XH_EINVAL, "xh_gaussian_likelihood_given_background: workspace too small");
    double* tm = (double*)workspace;
    double* d_lnl = tm + (size_t)N_ch * N_b;
    MarginalArgs A{};
    A.signals = comps; A.comp_phases = comp_phases; A.comp_np = comp_np; A.shifts = shifts; A.phases = phases;
    A.counts = counts; A.background = background; A.STAR = EXPEC;
    A.exposure = T; A.N_ch = N_ch; A.N_b = N_b; A.N_P = N_P; A.n_comp = n_comp; A.allow_negative = allow;
    A.interp = g_phase_interp;
    size_t tab = 0;
    {
        const int rc = cspline_tab_lds(A, s, &tab);
        if (rc != XH_OK) return rc;
    }
    hipLaunchKernelGGL(k_gauss_bg, dim3(N_ch), dim3(64), tab, s, A, errors, tm);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_gauss_reduce, dim3(1), dim3(64), 0, s, (const double*)tm, (size_t)N_ch * N_b, d_lnl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(lnl_out, d_lnl, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return XH_OK;
}

extern "C" int xh_gaussian_likelihood_given_background(double T, int N_b, const double* phases, int N_ch,
                                                       const double* counts, const double* errors, int n_comp,
                                                       const double* comps, int N_P, const double* comp_phases,
                                                       const double* shifts, const double* background,
                                                       int allow_negative, double* lnl_out, double* EXPEC,
                                                       void* workspace, size_t workspace_bytes, void* stream)
{
    return gaussian_impl(T, N_b, phases, N_ch, counts, errors, n_comp, comps, N_P, nullptr, comp_phases, shifts,
                         background, allow_negative ? ~0u : 0u, lnl_out, EXPEC, workspace, workspace_bytes, stream);
}

extern "C" int xh_gaussian_likelihood_given_background_ex(double T, int N_b, const double* phases, int N_ch,
                                                          const double* counts, const double* errors, int n_comp,
                                                          const double* comps, const int* comp_num_phases,
                                                          const double* comp_phases, const double* shifts,
                                                          const double* background, const int* allow_negative,
                                                          double* lnl_out, double* EXPEC, void* workspace,
                                                          size_t workspace_bytes, void* stream)
{
    if (!comp_num_phases)
        return fail(XH_EINVAL, "xh_gaussian_likelihood_given_background_ex: phases per component required");
    return gaussian_impl(T, N_b, phases, N_ch, counts, errors, n_comp, comps, 0, comp_num_phases, comp_phases, shifts,
                         background, allow_mask(n_comp, allow_negative), lnl_out, EXPEC, workspace, workspace_bytes,
                         stream);
}

extern "C" int xh_gaussian_likelihood_QnUn(int N_rows, int N_b, const double* counts, const double* errors,
                                           const double* star, double* lnl_out, void* workspace,
                                           size_t workspace_bytes, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (N_rows < 1 || N_b < 1) return fail(XH_EINVAL, "xh_gaussian_likelihood_QnUn: bad sizes");
    if (!counts || !errors || !star) return fail(XH_EINVAL, "xh_gaussian_likelihood_QnUn: null input");
    if (workspace_bytes < xh_gaussian_workspace_bytes(N_rows, N_b))
        return fail(XH_EINVAL, "xh_gaussian_likelihood_QnUn: workspace too small");
    double* tm = (double*)workspace;
    double* d_lnl = tm + (size_t)N_rows * N_b;
    const size_t n = (size_t)N_rows * N_b;
    hipLaunchKernelGGL(k_gauss_qnun, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N_rows, N_b, counts, errors,
                       star, tm);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_gauss_reduce, dim3(1), dim3(64), 0, s, (const double*)tm, n, d_lnl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(lnl_out, d_lnl, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return XH_OK;
}

extern "C" int xh_phase_integrator(double exposure_time, int N_b, const double* phases, int N_rows,
                                   const double* signal, int N_P, const double* signal_phases, double phase_shift,
                                   int allow_negative, double* out, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (N_rows < 1 || N_b < 1 || N_P < 2) return fail(XH_EINVAL, "xh_phase_integrator: bad sizes (two or more phases)");
    if (!phases || !signal || !signal_phases || !out) return fail(XH_EINVAL, "xh_phase_integrator: null input");
    MarginalArgs A{};
    A.signals = signal; A.comp_phases = signal_phases; A.phases = phases;
    A.exposure = exposure_time; A.N_ch = N_rows; A.N_b = N_b; A.N_P = N_P; A.n_comp = 1;
    A.allow_negative = allow_negative ? 1u : 0u;
    A.interp = g_phase_interp;
    size_t tab = 0;
    {
        const int rc = cspline_tab_lds(A, s, &tab);
        if (rc != XH_OK) return rc;
    }
    hipLaunchKernelGGL(k_phase_integrate, dim3(N_rows), dim3(64), tab, s, A, phase_shift, out);
    HIPCHK(hipGetLastError());
    return XH_OK;
}

extern "C" size_t xh_synthesise_workspace_bytes(int N_ch, int N_b) { return (size_t)N_ch * N_b * 8 + 256; }

// shared body of the two synthesis entry points (mode 0: given the exposure,
// mode 1: given the total expected star counts)
static int synthesise(int mode, double T, double n_star, int N_b, const double* phases, int N_ch, int n_comp,
                      const double* comps, int N_P, const double* comp_phases, const double* shifts, double n_bg,
                      const double* background, int allow_negative, unsigned long long seed, double* EXPEC,
                      double* SYNTH, double* scales_out, void* workspace, size_t workspace_bytes, hipStream_t s)
{
    if (N_ch < 1 || N_b < 1 || n_comp < 1 || N_P < 1) return fail(XH_EINVAL, "synthesise: bad sizes");
    if (!background) return fail(XH_EINVAL, "synthesise: the background is required");
    if (workspace_bytes < xh_synthesise_workspace_bytes(N_ch, N_b))
        return fail(XH_EINVAL, "synthesise: workspace too small");
    double* rate = (double*)workspace;
    double* scal = rate + (size_t)N_ch * N_b;
    MarginalArgs A{};
    A.signals = comps; A.comp_phases = comp_phases; A.shifts = shifts; A.phases = phases;
    A.N_ch = N_ch; A.N_b = N_b; A.N_P = N_P; A.n_comp = n_comp; A.allow_negative = allow_negative ? ~0u : 0u;
    A.interp = g_phase_interp;
    size_t tab = 0;
    {
        const int rc = cspline_tab_lds(A, s, &tab);
        if (rc != XH_OK) return rc;
    }
    hipLaunchKernelGGL(k_star_rate, dim3(N_ch), dim3(64), tab, s, A, rate);
    HIPCHK(hipGetLastError());
    const int n = N_ch * N_b;
    hipLaunchKernelGGL(k_synth_totals, dim3(1), dim3(64), 0, s, (const double*)rate, background, n,
                       mode == 1 ? 1.0e5 : 1.0, scal);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_synth_draw, dim3((n + 255) / 256), dim3(256), 0, s, mode, (const double*)rate, background, n,
                       T, n_star, n_bg, scal, seed, EXPEC, SYNTH);
    HIPCHK(hipGetLastError());
    double h[4];
    HIPCHK(hipMemcpyAsync(h, scal, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    scales_out[0] = h[2];
    scales_out[1] = h[3];
    return XH_OK;
}

extern "C" int xh_synthesise_exposure(double exposure_time, int N_b, const double* phases, int N_ch, int n_comp,
                                      const double* comps, int N_P, const double* comp_phases,
                                      const double* shifts, double expected_background_counts,
                                      const double* background, int allow_negative, unsigned long long seed,
                                      double* EXPEC, double* SYNTHETIC, double* scale_background,
                                      void* workspace, size_t workspace_bytes, void* stream)
{
    double sc[2];
    int rc = synthesise(0, exposure_time, 0.0, N_b, phases, N_ch, n_comp, comps, N_P, comp_phases, shifts,
                        expected_background_counts, background, allow_negative, seed, EXPEC, SYNTHETIC, sc,
                        workspace, workspace_bytes, (hipStream_t)stream);
    if (rc == XH_OK) *scale_background = sc[0];
    return rc;
}

extern "C" int xh_synthesise_given_total_count_number(int N_b, const double* phases, double expected_star_counts,
                                                      int N_ch, int n_comp, const double* comps, int N_P,
                                                      const double* comp_phases, const double* shifts,
                                                      double expected_background_counts, const double* background,
                                                      int allow_negative, unsigned long long seed, double* EXPEC,
                                                      double* SYNTHETIC, double* scales_out, void* workspace,
                                                      size_t workspace_bytes, void* stream)
{
    double sc[2];
    int rc = synthesise(1, 0.0, expected_star_counts, N_b, phases, N_ch, n_comp, comps, N_P, comp_phases, shifts,
                        expected_background_counts, background, allow_negative, seed, EXPEC, SYNTHETIC, sc,
                        workspace, workspace_bytes, (hipStream_t)stream);
    if (rc == XH_OK) {
        scales_out[0] = sc[1];               // SCALE_STAR: the exposure is 1e5 s times it
        scales_out[1] = sc[0] / sc[1];       // SCALE_BACKGROUND / SCALE_STAR
    }
    return rc;
}

// ============================================================= pipeline

struct xh_pipeline {
    xh_model_config cfg;
    Layout L;
    int nreg, mpr, maxb, N_in, N_ch, N_b, EC, table_n;
    int fold_kind;               // 1 per-point fold, 2 moment form, 3 energy lanes, 4 wave fold
    int fold_rings;              // parallels per fold block (0 = automatic), fixed at creation
    int wave_rings;              // parallels per wave-fold wavefront (0 = automatic)
    int wave_chain;              // intervals per wave-fold FP32 chain (0 = per parallel)
    int wave_shared;             // 1: k_wave_fold_shared where it applies
    int shape;                   // moment form: block shape T (0 = by batch size)
    int lane_nq;                 // energy-lane fold: phases per wave
    bool fp32;                   // precision form
    int fold_terms;              // precision form's response fold: 0 FP32 MFMA, 3 or 6 bf16 products
    float* d_Mf;                 // FP32-rounded response (precision form, fold_terms 0)
    __bf16 *d_Mh, *d_Mm, *d_Ml;  // three-term bf16 split of the response (fold_terms 3 or 6)
    int moff[3];                 // first member of each region within a request
    bool timing;
    hipEvent_t ev[16];
    double stage_ms[8];
    double* d_work;
    std::vector<void*> allocs;
    Member* d_members;
    RegionAlloc* d_regions;
    double *d_area_tab, *d_total, *d_theta, *d_phi, *d_r, *d_cellArea, *d_maxAlpha, *d_cos_gamma, *d_effGrav;
    double* d_ringlim;           // per parallel (lower, upper, centroid) colatitude
    int2* d_blist;               // boundary cells of the batch (member, ring << 16 | column)
    int* d_bcount;
    double *d_u, *d_logT, *d_maxDefl, *d_defl, *d_cosa, *d_lag, *d_geo, *d_flux;
    int *d_ring_ok, *d_valid, *d_member_index, *d_status, *d_lastq;
    int *d_raywork, *d_raycount, *d_else_raywork;   // rays left to the full QAG, and their counts
    double *d_atm_axes, *d_atm_I;
    double* d_slab;              // table contracted over (log T, log g) per parallel (moment form)
    double* d_atm_S;             // Lagrange denominators per axis and base node
    int atm_N[4];
    double *d_energies, *d_logE, *d_logEdges, *d_atten, *d_leaves, *d_phases, *d_comp_phases;
    double *d_integrated, *d_M, *d_folded, *d_data_phases, *d_counts, *d_precomp, *d_support, *d_shifts;
    double *d_ll, *d_bml, *d_bmls, *d_star, *d_lnl;
    Member* h_members;
    RegionAlloc* h_regions;
    double *h_shifts, *h_lnl;
    unsigned long long *h_seed, *d_seed;   // llzero-class draw of each launch (pinned host -> device)
    bool inflight;                         // a launch not yet waited for
    hipStream_t stream;
    int last_n;
    bool use_graphs;
    std::vector<hipGraphExec_t> graphs;   // per batch size n (index n), null until captured
    // elsewhere component: one member per request on its own closed mesh
    bool has_else;
    Layout Le;
    Member *d_else, *h_else;
    double *d_else_tab, *d_else_total, *d_else_colat, *d_else_r, *d_else_maxA, *d_else_cg, *d_else_eg;
    double *d_else_area, *d_else_defl, *d_else_cosa, *d_else_lag, *d_else_maxD, *d_else_part;
};

template <class T>
static int dalloc(xh_pipeline* p, T** ptr, size_t n)
{
    void* v = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&v, n * sizeof(T));
    if (e != hipSuccess) return fail(XH_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    p->allocs.push_back(v);
    *ptr = (T*)v;
    return XH_OK;
}

extern "C" void xh_pipeline_destroy(xh_pipeline* p)
{
    if (!p) return;
    for (hipGraphExec_t g : p->graphs) if (g) (void)hipGraphExecDestroy(g);
    for (void* v : p->allocs) (void)hipFree(v);
    for (int e = 0; e < 16; e++) if (p->ev[e]) (void)hipEventDestroy(p->ev[e]);
    if (p->h_members) (void)hipHostFree(p->h_members);
    if (p->h_regions) (void)hipHostFree(p->h_regions);
    if (p->h_shifts) (void)hipHostFree(p->h_shifts);
    if (p->h_lnl) (void)hipHostFree(p->h_lnl);
    if (p->h_seed) (void)hipHostFree(p->h_seed);
    if (p->h_else) (void)hipHostFree(p->h_else);
    delete p;
}

#define PCHK(x) do { if ((x) != XH_OK) { xh_pipeline_destroy(p); return nullptr; } } while (0)
#define PHIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { fail(XH_EHIP, hipGetErrorString(e_)); xh_pipeline_destroy(p); return nullptr; } } while (0)

extern "C" xh_pipeline* xh_pipeline_create(const xh_model_config* c)
{
    if (!c || c->n_regions < 1 || c->n_regions > 2 || c->max_batch < 1 || c->N_E < 5 || c->N_E > 8192 || c->num_leaves < 5 ||
        c->num_rays < 3 || c->N_in < 1 || c->N_ch < 1 || c->N_b < 1) {
        fail(XH_EINVAL, "xh_pipeline_create: bad configuration");
        return nullptr;
    }
    if (g_phase_interp != 0) {
        fail(XH_EINVAL, "xh_pipeline_create: the batched pipeline uses the periodic Akima phase interpolant");
        return nullptr;
    }
    xh_pipeline* p = new xh_pipeline();   // value-initialised: every field zero
    p->cfg = *c;
    p->nreg = c->n_regions;
    p->moff[0] = 0;
    for (int g = 0; g < c->n_regions; g++) p->moff[g + 1] = p->moff[g] + 1 + (c->region[g].cede ? 1 : 0);
    p->mpr = p->moff[c->n_regions];
    p->maxb = c->max_batch;
    p->N_in = c->N_in; p->N_ch = c->N_ch; p->N_b = c->N_b;
    p->table_n = 1000;
    int smax = c->max_sqrt_cells;
    for (int r = 0; r < c->n_regions; r++) smax = std::max(smax, c->region[r].max_sqrt_num_cells);
    smax += smax & 1;
    const int N_P = c->num_phases > 0 ? c->num_phases : c->num_leaves;
    const int IO = c->image_order_limit > 0 ? c->image_order_limit : 8;
    p->L = Layout{smax, c->num_rays, c->num_leaves, N_P, c->N_E, IO, c->image_order_limit > 0 ? c->image_order_limit : 0,
                  c->atm_ext};
    p->EC = pick_ec(c->num_leaves);
    if (c->precision < 0 || c->precision > 1 || c->fold_kernel < 0 || c->fold_kernel > 4) {
        fail(XH_EINVAL, "pipeline: precision must be 0 or 1 and fold_kernel 0 .. 4"); delete p; return nullptr;
    }
    if (c->fold_kernel == 4 && (c->precision != 1 || c->atm_ext != 1 || c->num_leaves < 5)) {
        fail(XH_EINVAL, "pipeline: the wave fold runs the precision form with the blackbody and at least 5 leaves");
        delete p; return nullptr;
    }
    p->fp32 = c->precision == 1;
    p->fold_terms = c->fold_terms;
    if (c->fold_terms != 0 && c->fold_terms != 3 && c->fold_terms != 6) {
        fail(XH_EINVAL, "pipeline: fold_terms must be 0, 3 or 6"); delete p; return nullptr;
    }
    {
        // a tuned shape larger than this model's LDS allows falls back to smaller ones
        p->shape = g_moment_shape;
        p->fold_rings = g_fold_rings;
        p->wave_rings = g_wave_rings;
        p->wave_chain = g_wave_chain;
        p->wave_shared = g_wave_shared;
        while (p->shape > 1 && moment_lds(c->num_leaves, smax, p->fp32 ? 4 : 8, p->shape, c->atm_ext) > kMomentLdsMax)
            p->shape >>= 1;
        const bool fits = moment_lds(c->num_leaves, smax, p->fp32 ? 4 : 8, p->shape > 0 ? p->shape : 1, c->atm_ext) <= kMomentLdsMax;
        if ((c->fold_kernel == 2 || p->fp32) && !fits) {
            fail(XH_EINVAL, "pipeline: the moment-form fold does not fit LDS at this leaf count and mesh bound");
            delete p; return nullptr;
        }
        if (p->fp32 && c->fold_kernel == 1) {
            fail(XH_EINVAL, "pipeline: the precision form uses the moment-form fold"); delete p; return nullptr;
        }
        const bool lane_fits = lane_lds(c->num_leaves, smax, p->fp32 ? 4 : 8) <= 160 * 1024;
        if (c->fold_kernel == 3 && !lane_fits) {
            fail(XH_EINVAL, "pipeline: the energy-lane fold does not fit LDS at this leaf count and mesh bound");
            delete p; return nullptr;
        }
        // automatic: the moment form, else the energy lanes, else per-point
        // automatic: the wave fold for the precision form with the blackbody
        // (measured fastest on every energy group of the batched workload),
        // else the block moment form, else energy lanes, else per-point
        const bool wave_ok = p->fp32 && c->atm_ext == 1 && c->num_leaves >= 5 &&
                             wave_fold_lds(c->num_leaves, smax) <= 64 * 1024;
        if (c->fold_kernel == 4) p->fold_kind = 4;
        else if (c->fold_kernel == 3) p->fold_kind = 3;
        else if (c->fold_kernel == 1) p->fold_kind = 1;
        else if (c->fold_kernel == 2) p->fold_kind = 2;
        else if (wave_ok) p->fold_kind = 4;
        else if (fits) p->fold_kind = 2;
        else p->fold_kind = lane_fits ? 3 : 1;
    }
    {
        // configure the fold kernel's LDS limit now, never inside a capture
        Layout L0{};
        L0.N_P = c->num_phases > 0 ? c->num_phases : c->num_leaves;
        L0.atm_ext = c->atm_ext;
        FoldArgs F0{};
        p->lane_nq = g_lane_nq;
        if (p->fold_kind == 4) {
            if (wave_fold_lds(c->num_leaves, smax) > 64 * 1024) {
                fail(XH_EINVAL, "pipeline: the wave fold's LDS exceeds 64 KiB at this leaf count and mesh bound");
                delete p; return nullptr;
            }
            // the shared walk for 2 to 4 energy chunks, where its block fits 64 KiB
            const int ne = c->N_E;
            if (p->wave_shared && !(ne > 64 && ne <= 256 && wave_fold_shared_lds(c->num_leaves, smax) <= 64 * 1024))
                p->wave_shared = 0;
        } else if (p->fold_kind == 3) PHIP(launch_lane(p->fp32, p->lane_nq, dim3(0, 1, 1), 0, 0, nullptr, L0, F0));
        else if (p->fold_kind == 2) PHIP(launch_moment(p->fp32, 0, dim3(0, 1, 1), 0, 0, nullptr, L0, F0));
        else PHIP(launch_fold(p->EC, dim3(0, 1, 1), 0, 0, nullptr, L0, F0));
    }
    if (p->fold_kind == 1 && N_P > (256 / p->EC) * 16) { fail(XH_EINVAL, "too many phases for the leaf count"); delete p; return nullptr; }
    if (c->atm_ext != 1 && c->atm_ext != 2) { fail(XH_EINVAL, "pipeline: atmosphere must be 1 (blackbody) or 2 (4-D table)"); delete p; return nullptr; }
    if (c->atm_ext == 2) {
        for (int d = 0; d < 4; d++)
            if (!c->atm_axes[d] || c->atm_N[d] < 4) { fail(XH_EINVAL, "pipeline: 4-D table axes missing or too short"); delete p; return nullptr; }
        if (!c->atm_I) { fail(XH_EINVAL, "pipeline: 4-D table values missing"); delete p; return nullptr; }
    }
    const size_t nm = (size_t)p->maxb * p->mpr, S = smax, NR = c->num_rays, NE = c->N_E;
    const size_t ng = (size_t)p->maxb * p->nreg;
    PCHK(dalloc(p, &p->d_members, nm));
    PCHK(dalloc(p, &p->d_regions, ng));
    PCHK(dalloc(p, &p->d_area_tab, nm * p->table_n));
    PCHK(dalloc(p, &p->d_total, nm));
    PCHK(dalloc(p, &p->d_theta, nm * S));
    PCHK(dalloc(p, &p->d_phi, nm * S));
    PCHK(dalloc(p, &p->d_r, nm * S));
    PCHK(dalloc(p, &p->d_cellArea, nm * S * S));
    PCHK(dalloc(p, &p->d_ringlim, nm * S * 3));
    PCHK(dalloc(p, &p->d_blist, nm * S * S));
    PCHK(dalloc(p, &p->d_bcount, 1));
    PCHK(dalloc(p, &p->d_maxAlpha, nm * S));
    PCHK(dalloc(p, &p->d_cos_gamma, nm * S));
    PCHK(dalloc(p, &p->d_effGrav, nm * S));
    PCHK(dalloc(p, &p->d_u, nm * S));
    PCHK(dalloc(p, &p->d_logT, nm * S));
    PCHK(dalloc(p, &p->d_maxDefl, nm * S));
    PCHK(dalloc(p, &p->d_ring_ok, nm * S));
    PCHK(dalloc(p, &p->d_raywork, nm * S * NR));
    PCHK(dalloc(p, &p->d_raycount, 2));
    PCHK(dalloc(p, &p->d_defl, nm * S * NR));
    PCHK(dalloc(p, &p->d_cosa, nm * S * NR));
    PCHK(dalloc(p, &p->d_lag, nm * S * NR));
    PCHK(dalloc(p, &p->d_geo, nm * S * IO * 5 * c->num_leaves));
    PCHK(dalloc(p, &p->d_valid, nm * S * IO));
    PCHK(dalloc(p, &p->d_lastq, nm * S * IO));
    if (c->atm_ext == 2) {
        size_t na = 0, ni = 1;
        for (int d = 0; d < 4; d++) { p->atm_N[d] = c->atm_N[d]; na += c->atm_N[d]; ni *= c->atm_N[d]; }
        PCHK(dalloc(p, &p->d_atm_axes, na));
        PCHK(dalloc(p, &p->d_atm_I, ni));
        size_t off = 0;
        for (int d = 0; d < 4; d++) {
            PHIP(hipMemcpy(p->d_atm_axes + off, c->atm_axes[d], (size_t)c->atm_N[d] * 8, hipMemcpyHostToDevice));
            off += c->atm_N[d];
        }
        PHIP(hipMemcpy(p->d_atm_I, c->atm_I, ni * 8, hipMemcpyHostToDevice));
        if (p->fold_kind == 2) PCHK(dalloc(p, &p->d_slab, nm * S * (size_t)c->atm_N[2] * c->atm_N[3]));
        // 1 / prod_{m != k} (x_{b+k} - x_{b+m}) per axis and base node, in the
        // order hot_Num4D.pyx:385-402 divides
        std::vector<double> Sd;
        for (int d = 0; d < 4; d++) {
            const double* x = c->atm_axes[d];
            for (int b = 0; b + 3 < c->atm_N[d]; b++)
                for (int k = 0; k < 4; k++) {
                    double v = 1.0;
                    bool first = true;
                    for (int mm = 0; mm < 4; mm++) {
                        if (mm == k) continue;
                        if (first) { v = 1.0 / (x[b + k] - x[b + mm]); first = false; }
                        else v /= x[b + k] - x[b + mm];
                    }
                    Sd.push_back(v);
                }
        }
        PCHK(dalloc(p, &p->d_atm_S, Sd.size()));
        PHIP(hipMemcpy(p->d_atm_S, Sd.data(), Sd.size() * 8, hipMemcpyHostToDevice));
    }
    PCHK(dalloc(p, &p->d_flux, nm * NE * N_P));
    PCHK(dalloc(p, &p->d_member_index, ng * 2));
    PCHK(dalloc(p, &p->d_work, nm));
    PCHK(dalloc(p, &p->d_energies, NE));
    PCHK(dalloc(p, &p->d_logE, NE));
    PCHK(dalloc(p, &p->d_logEdges, (size_t)c->N_in + 1));
    PCHK(dalloc(p, &p->d_atten, (size_t)c->N_in));
    PCHK(dalloc(p, &p->d_leaves, (size_t)c->num_leaves));
    PCHK(dalloc(p, &p->d_phases, (size_t)N_P));
    PCHK(dalloc(p, &p->d_comp_phases, (size_t)N_P));
    PCHK(dalloc(p, &p->d_integrated, ng * c->N_in * N_P));
    PCHK(dalloc(p, &p->d_M, (size_t)c->N_ch * c->N_in));
    PCHK(dalloc(p, &p->d_folded, ng * c->N_ch * N_P));
    PCHK(dalloc(p, &p->d_data_phases, (size_t)c->N_b + 1));
    PCHK(dalloc(p, &p->d_counts, (size_t)c->N_ch * c->N_b));
    PCHK(dalloc(p, &p->d_precomp, (size_t)c->N_ch));
    PCHK(dalloc(p, &p->d_support, (size_t)c->N_ch * 2));
    PCHK(dalloc(p, &p->d_shifts, ng));
    PCHK(dalloc(p, &p->d_ll, (size_t)p->maxb * c->N_ch));
    PCHK(dalloc(p, &p->d_status, (size_t)p->maxb * c->N_ch));
    PCHK(dalloc(p, &p->d_bml, (size_t)p->maxb * c->N_ch));
    PCHK(dalloc(p, &p->d_bmls, (size_t)p->maxb * c->N_ch));
    PCHK(dalloc(p, &p->d_star, (size_t)p->maxb * c->N_ch * c->N_b));
    PCHK(dalloc(p, &p->d_lnl, (size_t)p->maxb));
    if (c->elsewhere) {
        if (c->else_sqrt_num_cells < 10 || c->else_sqrt_num_cells % 2 || c->else_num_rays < 3) {
            fail(XH_EINVAL, "pipeline: elsewhere needs an even sqrt_num_cells >= 10 and >= 3 rays");
            xh_pipeline_destroy(p); return nullptr;
        }
        p->has_else = true;
        const int Se = c->else_sqrt_num_cells, NRe = c->else_num_rays;
        p->Le = Layout{Se, NRe, c->num_leaves, N_P, c->N_E, 8,
                       c->else_image_order_limit > 0 ? c->else_image_order_limit : 0, 1};
        const size_t nb = (size_t)p->maxb;
        PCHK(dalloc(p, &p->d_else, nb));
        PCHK(dalloc(p, &p->d_else_tab, nb * p->table_n));
        PCHK(dalloc(p, &p->d_else_total, nb));
        PCHK(dalloc(p, &p->d_else_colat, nb * Se));
        PCHK(dalloc(p, &p->d_else_r, nb * Se));
        PCHK(dalloc(p, &p->d_else_maxA, nb * Se));
        PCHK(dalloc(p, &p->d_else_cg, nb * Se));
        PCHK(dalloc(p, &p->d_else_eg, nb * Se));
        PCHK(dalloc(p, &p->d_else_area, nb));
        PCHK(dalloc(p, &p->d_else_defl, nb * Se * NRe));
        PCHK(dalloc(p, &p->d_else_cosa, nb * Se * NRe));
        PCHK(dalloc(p, &p->d_else_lag, nb * Se * NRe));
        PCHK(dalloc(p, &p->d_else_maxD, nb * Se));
        PCHK(dalloc(p, &p->d_else_part, nb * Se * NE));
        PCHK(dalloc(p, &p->d_else_raywork, nb * Se * NRe));
        PHIP(hipHostMalloc((void**)&p->h_else, nb * sizeof(Member), hipHostMallocDefault));
    }
    PHIP(hipHostMalloc((void**)&p->h_members, nm * sizeof(Member), hipHostMallocDefault));
    PHIP(hipHostMalloc((void**)&p->h_regions, ng * sizeof(RegionAlloc), hipHostMallocDefault));
    PHIP(hipHostMalloc((void**)&p->h_shifts, ng * sizeof(double), hipHostMallocDefault));
    PHIP(hipHostMalloc((void**)&p->h_lnl, (size_t)p->maxb * sizeof(double), hipHostMallocDefault));
    PHIP(hipHostMalloc((void**)&p->h_seed, sizeof(unsigned long long), hipHostMallocDefault));
    PCHK(dalloc(p, &p->d_seed, 1));
    p->h_seed[0] = next_seed();
    // constant inputs
    std::vector<double> v(std::max<size_t>({NE, (size_t)c->N_in + 1, (size_t)c->num_leaves, (size_t)N_P}));
    PHIP(hipMemcpy(p->d_energies, c->energies, NE * 8, hipMemcpyHostToDevice));
    for (size_t i = 0; i < NE; i++) v[i] = std::log10(c->energies[i]);
    PHIP(hipMemcpy(p->d_logE, v.data(), NE * 8, hipMemcpyHostToDevice));
    for (int i = 0; i <= c->N_in; i++) v[i] = std::log10(c->energy_edges[i]);
    PHIP(hipMemcpy(p->d_logEdges, v.data(), ((size_t)c->N_in + 1) * 8, hipMemcpyHostToDevice));
    if (c->attenuation) PHIP(hipMemcpy(p->d_atten, c->attenuation, (size_t)c->N_in * 8, hipMemcpyHostToDevice));
    // numpy.linspace(0, 2 pi, N_L) and 2 pi * numpy.linspace(0, 1, N_P)
    const double two_pi = 2.0 * M_PI;
    for (int i = 0; i < c->num_leaves; i++)
        v[i] = (i == c->num_leaves - 1) ? two_pi : (double)i * (two_pi / (double)(c->num_leaves - 1)) + 0.0;
    PHIP(hipMemcpy(p->d_leaves, v.data(), (size_t)c->num_leaves * 8, hipMemcpyHostToDevice));
    std::vector<double> cyc(N_P);
    for (int i = 0; i < N_P; i++) cyc[i] = (i == N_P - 1) ? 1.0 : (double)i * (1.0 / (double)(N_P - 1)) + 0.0;
    PHIP(hipMemcpy(p->d_comp_phases, cyc.data(), (size_t)N_P * 8, hipMemcpyHostToDevice));
    for (int i = 0; i < N_P; i++) v[i] = two_pi * cyc[i];
    PHIP(hipMemcpy(p->d_phases, v.data(), (size_t)N_P * 8, hipMemcpyHostToDevice));
    PHIP(hipMemcpy(p->d_M, c->matrix, (size_t)c->N_ch * c->N_in * 8, hipMemcpyHostToDevice));
    if (p->fp32 && p->fold_terms == 0) {
        const size_t nmat = (size_t)c->N_ch * c->N_in;
        std::vector<float> f(nmat);
        for (size_t q = 0; q < nmat; q++) f[q] = (float)c->matrix[q];
        PCHK(dalloc(p, &p->d_Mf, nmat));
        PHIP(hipMemcpy(p->d_Mf, f.data(), nmat * sizeof(float), hipMemcpyHostToDevice));
    } else if (p->fp32) {
        // three-term bf16 split of the FP32-rounded response, once
        const size_t nmat = (size_t)c->N_ch * c->N_in;
        std::vector<__bf16> h(nmat), m(nmat), l(nmat);
        for (size_t q = 0; q < nmat; q++) {
            const float x = (float)c->matrix[q];
            h[q] = (__bf16)x;
            const float r = x - (float)h[q];
            m[q] = (__bf16)r;
            l[q] = (__bf16)(r - (float)m[q]);
        }
        PCHK(dalloc(p, &p->d_Mh, nmat));
        PCHK(dalloc(p, &p->d_Mm, nmat));
        PCHK(dalloc(p, &p->d_Ml, nmat));
        PHIP(hipMemcpy(p->d_Mh, h.data(), nmat * sizeof(__bf16), hipMemcpyHostToDevice));
        PHIP(hipMemcpy(p->d_Mm, m.data(), nmat * sizeof(__bf16), hipMemcpyHostToDevice));
        PHIP(hipMemcpy(p->d_Ml, l.data(), nmat * sizeof(__bf16), hipMemcpyHostToDevice));
    }
    PHIP(hipMemcpy(p->d_data_phases, c->data_phases, ((size_t)c->N_b + 1) * 8, hipMemcpyHostToDevice));
    PHIP(hipMemcpy(p->d_counts, c->counts, (size_t)c->N_ch * c->N_b * 8, hipMemcpyHostToDevice));
    std::vector<int> ic((size_t)c->N_ch * c->N_b);
    for (size_t i = 0; i < ic.size(); i++) ic[i] = (int)c->counts[i];
    std::vector<double> pre(c->N_ch);
    xh_precomputation(c->N_ch, c->N_b, ic.data(), pre.data());
    PHIP(hipMemcpy(p->d_precomp, pre.data(), (size_t)c->N_ch * 8, hipMemcpyHostToDevice));
    std::vector<double> sup((size_t)c->N_ch * 2);
    for (int i = 0; i < c->N_ch; i++) {
        sup[2 * i] = c->support ? c->support[2 * i] : 0.0;
        sup[2 * i + 1] = c->support ? c->support[2 * i + 1] : -1.0;
    }
    PHIP(hipMemcpy(p->d_support, sup.data(), sup.size() * 8, hipMemcpyHostToDevice));
    // energy integration groups: the members of each (request, region)
    std::vector<int> midx(ng * 2);
    for (int rq = 0; rq < p->maxb; rq++)
        for (int g = 0; g < p->nreg; g++) {
            int gi = rq * p->nreg + g, m0 = rq * p->mpr + p->moff[g];
            midx[2 * gi] = m0;
            midx[2 * gi + 1] = (p->moff[g + 1] - p->moff[g] > 1) ? m0 + 1 : -1;
        }
    PHIP(hipMemcpy(p->d_member_index, midx.data(), ng * 2 * sizeof(int), hipMemcpyHostToDevice));
    for (int e = 0; e < 16; e++) PHIP(hipEventCreate(&p->ev[e]));
    return p;
}

// host-side spacetime and descriptors for one request (Spacetime.py:111-190,
// HotRegion.py:774-871)
static void describe(const xh_pipeline* p, int rq, const xh_request& r)
{
    const xh_model_config& c = p->cfg;
    const double r_g = r.mass * kGM;
    const double Msi = r.mass * kGM * (kC * kC) / kG;
    const double R = r.radius * kKM;
    const double Omega = 2.0 * M_PI * c.frequency;
    const double d = r.distance * kKPC;
    const double zeta = r_g / R;
    const double eps = std::pow(Omega, 2.0) * std::pow(R, 3.0) / (kG * Msi);
    const double incl = std::acos(r.cos_inclination);
    for (int g = 0; g < p->nreg; g++) {
        const double* q = r.region[g];
        const double shift = q[0], sc = q[1], sr = q[2], sT = q[3];
        const bool cede = c.region[g].cede != 0;
        const double cc = cede ? q[4] : sc, cr = cede ? q[5] : 0.0, ca = cede ? q[6] : 0.0, cT = q[7];
        // an omitted region cuts the superseding member; with a ceding ring
        // the reference ignores the omission (HotRegion.py:44)
        const bool omit = c.region[g].omit != 0 && !cede;
        const double oc = omit ? r.omit[g][0] : sc, orad = omit ? r.omit[g][1] : 0.0,
                     oa = omit ? r.omit[g][2] : 0.0;
        const int gi = rq * p->nreg + g;
        const int m0 = rq * p->mpr + p->moff[g], m1 = m0 + 1;
        const int nmem = p->moff[g + 1] - p->moff[g];
        for (int s = 0; s < nmem; s++) {
            Member& m = p->h_members[m0 + s];
            std::memset((void*)&m, 0, sizeof(Member));
            m.R_eq = R; m.epsilon = eps; m.zeta = zeta; m.Omega = Omega; m.r_s = 2.0 * r_g;
            m.inclination = incl; m.mode_freq = c.mode_frequency; m.d_sq = d * d; m.shape = c.star_shape;
            m.antiphased = c.region[g].is_antiphased;
            if (s == 0) {
                m.active = 1;
                m.cedeR = sr; m.cedeColat = sc; m.superR = orad; m.superColat = oc; m.superAzi = oa;
                m.phi_offset = -oa;
                m.polar = (sc - sr < 0.0) || (sc + sr > M_PI);
                m.logT = sT;
            } else {
                m.active = cede && cr > 0.0;
                m.cedeR = cr; m.cedeColat = cc; m.superR = sr; m.superColat = sc; m.superAzi = -ca;
                m.phi_offset = ca;
                m.polar = (cc - cr < 0.0) || (cc + cr > M_PI);
                m.logT = cT;
            }
            if (p->has_else) { m.correct = 1; m.else_logT = r.elsewhere_temperature; }
        }
        RegionAlloc& A = p->h_regions[gi];
        int snc = r.sqrt_num_cells[g] > 0 ? r.sqrt_num_cells[g] : c.region[g].sqrt_num_cells;
        A.superR = sr; A.cedeR = cr; A.superColat = sc; A.cedeColat = cc; A.superAzi = -ca;
        A.holeColat = oc; A.holeR = orad; A.holeAzi = oa;
        A.R_eq = R; A.Omega = Omega; A.zeta = zeta; A.epsilon = eps; A.shape = c.star_shape;
        A.num_cells = snc * snc; A.min_sqrt = c.region[g].min_sqrt_num_cells;
        A.max_sqrt = c.region[g].max_sqrt_num_cells;
        A.member0 = m0; A.member1 = (nmem > 1 && cr > 0.0) ? m1 : -1;
        p->h_shifts[gi] = shift;
    }
    if (p->has_else) {
        // Elsewhere.embed (xpsi/Elsewhere.py:254-303)
        Member& e = p->h_else[rq];
        std::memset((void*)&e, 0, sizeof(Member));
        e.R_eq = R; e.epsilon = eps; e.zeta = zeta; e.Omega = Omega; e.r_s = 2.0 * r_g;
        e.inclination = incl; e.mode_freq = c.mode_frequency; e.d_sq = d * d; e.shape = c.star_shape;
        e.active = 1; e.S = c.else_sqrt_num_cells; e.numCell = (double)e.S * e.S;
        e.logT = r.elsewhere_temperature;
    }
}

static int enqueue(xh_pipeline* p, int n, hipStream_t s);
static int capture_graph(xh_pipeline* p, int n, hipStream_t s);

extern "C" int xh_pipeline_launch(xh_pipeline* p, int n, const xh_request* reqs, void* stream)
{
    if (!p || n < 1 || n > p->maxb) return fail(XH_EINVAL, "xh_pipeline_launch: batch size out of range");
    hipStream_t s = (hipStream_t)stream;
    // the pinned descriptors and results are shared with the previous launch:
    // let it finish before they are rewritten
    if (p->inflight) {
        HIPCHK(hipStreamSynchronize(p->stream));
        p->inflight = false;
    }
    p->stream = s;
    p->last_n = n;
    for (int rq = 0; rq < n; rq++) describe(p, rq, reqs[rq]);
    p->h_seed[0] = next_seed();
    p->inflight = true;
    if (!p->use_graphs || p->timing || s == nullptr) return enqueue(p, n, s);
    int rc = capture_graph(p, n, s);
    if (rc != XH_OK) return rc;
    HIPCHK(hipGraphLaunch(p->graphs[n], s));
    return XH_OK;
}

// Capture the whole device sequence once for batch size n; the host
// descriptors it copies from are refreshed before each replay.
static int capture_graph(xh_pipeline* p, int n, hipStream_t s)
{
    if ((int)p->graphs.size() <= n) p->graphs.resize(n + 1, nullptr);
    if (p->graphs[n]) return XH_OK;
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(p, n, s);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != XH_OK) return rc;
    HIPCHK(e);
    hipGraphExec_t ge;
    HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    p->graphs[n] = ge;
    return XH_OK;
}

extern "C" int xh_pipeline_capture(xh_pipeline* p, int n, void* stream)
{
    if (!p || n < 1 || n > p->maxb) return fail(XH_EINVAL, "xh_pipeline_capture: batch size out of range");
    if (stream == nullptr) return fail(XH_EINVAL, "xh_pipeline_capture: needs a non-default stream");
    return capture_graph(p, n, (hipStream_t)stream);
}

static int enqueue(xh_pipeline* p, int n, hipStream_t s)
{
    const Layout L = p->L;
    const int nm = n * p->mpr, ng = n * p->nreg;
    const int S = L.S_max, N_P = L.N_P, N_E = L.N_E;
    HIPCHK(hipMemcpyAsync(p->d_members, p->h_members, (size_t)nm * sizeof(Member), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(p->d_regions, p->h_regions, (size_t)ng * sizeof(RegionAlloc), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(p->d_shifts, p->h_shifts, (size_t)ng * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(p->d_seed, p->h_seed, sizeof(unsigned long long), hipMemcpyHostToDevice, s));
    if (p->has_else)
        HIPCHK(hipMemcpyAsync(p->d_else, p->h_else, (size_t)n * sizeof(Member), hipMemcpyHostToDevice, s));
    HIPCHK(fill(p->d_cellArea, (size_t)nm * S * S, 0.0, s));
    HIPCHK(fill(p->d_flux, (size_t)nm * N_E * N_P, 0.0, s));
    HIPCHK(fill_int(p->d_ring_ok, (size_t)nm * S, 0, s));
    HIPCHK(fill_int(p->d_valid, (size_t)nm * S * L.IO_max, 0, s));
    if (p->timing) HIPCHK(hipEventRecord(p->ev[0], s));
    hipLaunchKernelGGL(k_allocate, dim3(ng), dim3(256), 0, s, p->d_regions, p->d_members, S);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_mesh_table, dim3(nm), dim3(1024), 0, s, p->d_members,
                       p->d_area_tab, p->d_total, p->table_n);
    HIPCHK(hipGetLastError());
    HIPCHK(launch_mesh_cells(p->d_members, p->d_area_tab, p->d_total, p->table_n, L, (int)nm, S, p->d_theta,
                             p->d_phi, p->d_r, p->d_cellArea, p->d_maxAlpha, p->d_cos_gamma, p->d_effGrav,
                             p->d_ringlim, p->d_blist, p->d_bcount, s));
    hipLaunchKernelGGL(k_ring_finalize, dim3(nm), dim3(256), 0, s, p->d_members, L, p->d_r, p->d_cellArea,
                       p->d_u, p->d_ring_ok, p->d_logT);
    HIPCHK(hipGetLastError());
    const Layout Le = p->Le;
    if (p->has_else) {
        hipLaunchKernelGGL(k_closed_table, dim3((p->table_n + 1 + 3) / 4, n), dim3(256), 0, s, p->d_else,
                           p->d_else_tab, p->d_else_total, p->table_n);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_closed_rings, dim3(Le.S_max / 2, n), dim3(64), 0, s, p->d_else, p->d_else_tab,
                           p->d_else_total, p->table_n, Le.S_max, p->d_else_colat, p->d_else_r, p->d_else_maxA,
                           p->d_else_cg, p->d_else_eg, p->d_else_area);
        HIPCHK(hipGetLastError());
    }
    if (p->timing) HIPCHK(hipEventRecord(p->ev[1], s));
    // rays: lane-per-ray first pass, then the listed rays with full QAG
    HIPCHK(fill_int(p->d_raycount, 2, 0, s));
    hipLaunchKernelGGL(k_rays_lane, dim3((unsigned)(((size_t)S * (L.N_R - 1) + 63) / 64), 1, nm), dim3(64), 0, s,
                       p->d_members, L, p->d_r,
                       p->d_maxAlpha, p->d_cosa, p->d_defl, p->d_lag, p->d_raywork, p->d_raycount,
                       (const double*)nullptr);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_rays, dim3(kRayWaveBlocks), dim3(256), 0, s, p->d_members, L, p->d_r, p->d_maxAlpha,
                       p->d_defl, p->d_lag, p->d_raywork, p->d_raycount, (const double*)nullptr);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_rays_fix, dim3(nm), dim3(256), 0, s, p->d_members, L, p->d_r, p->d_maxAlpha,
                       p->d_defl, p->d_lag, p->d_maxDefl, 1, (const double*)nullptr);
    HIPCHK(hipGetLastError());
    if (p->has_else) {
        hipLaunchKernelGGL(k_rays_lane, dim3((unsigned)(((size_t)Le.S_max * (Le.N_R - 1) + 63) / 64), 1, n), dim3(64), 0,
                           s, p->d_else, Le,
                           p->d_else_r, p->d_else_maxA, p->d_else_cosa, p->d_else_defl, p->d_else_lag,
                           p->d_else_raywork, p->d_raycount + 1, (const double*)nullptr);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_rays, dim3(kRayWaveBlocks), dim3(256), 0, s, p->d_else, Le, p->d_else_r,
                           p->d_else_maxA, p->d_else_defl, p->d_else_lag, p->d_else_raywork, p->d_raycount + 1,
                           (const double*)nullptr);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_rays_fix, dim3(n), dim3(256), 0, s, p->d_else, Le, p->d_else_r, p->d_else_maxA,
                           p->d_else_defl, p->d_else_lag, p->d_else_maxD, 0, (const double*)nullptr);
        HIPCHK(hipGetLastError());
    }
    RingArgs RA{p->d_theta, p->d_r, p->d_u, p->d_cos_gamma, p->d_maxDefl, p->d_defl, p->d_cosa, p->d_lag,
                p->d_leaves, p->d_ring_ok, p->d_geo, p->d_valid, p->d_lastq};
    if (p->timing) HIPCHK(hipEventRecord(p->ev[2], s));
    hipLaunchKernelGGL(k_geometry<0>, dim3(S, 1, nm), dim3(64), geo_lds(L.N_R, L.N_L), s, p->d_members, L, RA);
    HIPCHK(hipGetLastError());
    if (p->has_else) {
        hipLaunchKernelGGL(k_else_integrate, dim3(Le.S_max, n), dim3(256), else_lds(Le), s, p->d_else, Le,
                           p->d_else_colat, p->d_else_r, p->d_else_cg, p->d_else_maxD, p->d_else_defl,
                           p->d_else_cosa, p->d_energies, N_E, nullptr, 0, nullptr, 0, p->d_else_part);
        HIPCHK(hipGetLastError());
    }
    // parallels per block: all of them when the batch fills the GPU
    // (the energy-lane fold: 8 parallels per block, so that members of very
    // different sizes finish in similar times; partial sums meet by atomics)
    int rb = (n >= 32) ? S : 1;
    if (p->fold_kind == 4) rb = std::min(S, p->wave_rings > 0 ? p->wave_rings : 1);
    else if (n >= 32 && p->fold_rings > 0 && p->fold_kind != 1) rb = std::min(S, p->fold_rings);
    else if (n >= 32 && p->fold_kind == 3) rb = std::min(S, 8);
    FoldArgs FA{p->d_geo, p->d_valid, p->d_cellArea, nullptr, p->d_phi, 0, p->d_logT, p->d_effGrav, p->d_lastq,
                p->d_atm_axes, p->d_atm_I, {p->atm_N[0], p->atm_N[1], p->atm_N[2], p->atm_N[3]}, p->d_energies,
                p->d_phases, p->d_flux, rb, nullptr};
    FA.chain = p->wave_chain;
    if (p->timing) HIPCHK(hipEventRecord(p->ev[3], s));
    if (p->d_slab && p->fold_kind == 2) {
        hipLaunchKernelGGL(k_table_slab, dim3(8, S, nm), dim3(256), 0, s, (const Member*)p->d_members, L, FA,
                           p->d_slab);
        HIPCHK(hipGetLastError());
        FA.slab = p->d_slab;
        FA.atm_S = p->d_atm_S;
    }
    if (p->fold_kind == 4) {
        const unsigned chunks = (unsigned)(((N_E + 63) / 64) * ((N_P + 63) / 64));
        if (p->wave_shared)
            hipLaunchKernelGGL(k_wave_fold_shared, dim3((unsigned)S, (unsigned)((N_P + 63) / 64), nm),
                               dim3(64 * ((N_E + 63) / 64)), wave_fold_shared_lds(L.N_L, L.S_max), s, p->d_members,
                               L, FA);
        else
            hipLaunchKernelGGL(k_wave_fold, dim3((unsigned)((S + rb - 1) / rb), chunks, nm), dim3(64),
                               wave_fold_lds(L.N_L, L.S_max), s, p->d_members, L, FA);
        HIPCHK(hipGetLastError());
    }
    else if (p->fold_kind == 3)
        HIPCHK(launch_lane(p->fp32, p->lane_nq, lane_grid(N_E, N_P, p->lane_nq, (S + rb - 1) / rb, nm),
                           lane_lds(L.N_L, L.S_max, p->fp32 ? 4 : 8), s, p->d_members, L, FA));
    else if (p->fold_kind == 2)
    {
        const size_t rbytes = p->fp32 ? 4 : 8;
        const int T = p->shape > 0 ? p->shape : moment_shape(L.N_L, L.S_max, rbytes, L.atm_ext, n);
        HIPCHK(launch_moment(p->fp32, T, moment_grid(N_E, N_P, T, (S + rb - 1) / rb, nm),
                             moment_lds(L.N_L, L.S_max, rbytes, T, L.atm_ext), s, p->d_members, L, FA));
    }
    else
        HIPCHK(launch_fold(p->EC, dim3((N_E + p->EC - 1) / p->EC, (S + rb - 1) / rb, nm), fold_lds(L.N_L, p->EC, S),
                           s, p->d_members, L, FA));
    EnergyArgs EA{p->d_flux, p->d_logE, p->d_logEdges, p->cfg.attenuation ? p->d_atten : nullptr,
                  p->d_member_index, p->d_integrated, N_E, N_P, p->N_in + 1, 2};
    if (p->timing) HIPCHK(hipEventRecord(p->ev[4], s));
    if (p->has_else) {
        hipLaunchKernelGGL(k_else_add, dim3((N_E + 255) / 256, n), dim3(256), 0, s, p->d_else, Le, p->d_else_part,
                           p->d_else_area, p->d_energies, N_E, N_P, p->d_members, p->mpr, p->d_flux);
        HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_energy_integrate, dim3((p->N_in * N_P + 255) / 256, ng), dim3(256), (size_t)N_E * 8, s,
                       p->d_members, EA);
    HIPCHK(hipGetLastError());
    FoldGemmArgs GA{p->d_M, p->d_integrated, p->d_folded, p->N_ch, p->N_in, N_P, p->N_in};
    if (p->fp32 && p->fold_terms == 0)
        hipLaunchKernelGGL(k_fold_f32, dim3((N_P + 31) / 32, (p->N_ch + 31) / 32, ng), dim3(256), 0, s, GA, p->d_Mf);
    else if (p->fp32 && p->fold_terms == 3)
        hipLaunchKernelGGL(k_fold_bf16<3>, dim3((N_P + 31) / 32, (p->N_ch + 31) / 32, ng), dim3(256), 0, s, GA,
                           p->d_Mh, p->d_Mm, p->d_Ml);
    else if (p->fp32)
        hipLaunchKernelGGL(k_fold_bf16<6>, dim3((N_P + 31) / 32, (p->N_ch + 31) / 32, ng), dim3(256), 0, s, GA,
                           p->d_Mh, p->d_Mm, p->d_Ml);
    else
        hipLaunchKernelGGL(k_fold_f64, dim3((N_P + 31) / 32, (p->N_ch + 31) / 32, ng), dim3(256), 0, s, GA);
    HIPCHK(hipGetLastError());
    MarginalArgs MA{};
    MA.signals = p->d_folded; MA.comp_phases = p->d_comp_phases; MA.shifts = p->d_shifts;
    MA.phases = p->d_data_phases; MA.counts = p->d_counts; MA.precomp = p->d_precomp; MA.support = p->d_support;
    MA.background = nullptr; MA.request_status = nullptr; MA.ll = p->d_ll; MA.status = p->d_status;
    MA.B_ML = p->d_bml; MA.B_ML_supp = p->d_bmls; MA.STAR = p->d_star;
    MA.exposure = p->cfg.exposure_time; MA.epsabs = p->cfg.epsabs; MA.epsrel = p->cfg.epsrel;
    MA.epsilon = p->cfg.epsilon; MA.sigmas = p->cfg.sigmas; MA.llzero = p->cfg.llzero; MA.slim = p->cfg.slim;
    MA.N_ch = p->N_ch; MA.N_b = p->N_b; MA.N_P = N_P; MA.n_comp = p->nreg; MA.allow_negative = 0u;
    MA.seed = 0; MA.seed_ptr = p->d_seed;
    if (p->timing) HIPCHK(hipEventRecord(p->ev[5], s));
    hipLaunchKernelGGL(k_marginal, dim3(p->N_ch, n), dim3(64), 2 * p->N_b * 8, s, MA);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_lnl_reduce, dim3(n), dim3(64), 0, s, MA, (const Member*)p->d_members, p->mpr, p->d_lnl);
    HIPCHK(hipGetLastError());
    if (p->timing) HIPCHK(hipEventRecord(p->ev[6], s));
    HIPCHK(hipMemcpyAsync(p->h_lnl, p->d_lnl, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    return XH_OK;
}

extern "C" int xh_pipeline_use_graphs(xh_pipeline* p, int on)
{
    p->use_graphs = on != 0;
    return XH_OK;
}

extern "C" int xh_pipeline_sync(xh_pipeline* p)
{
    HIPCHK(hipStreamSynchronize(p->stream));
    p->inflight = false;
    if (p->timing) {
        for (int k = 0; k < 6; k++) {
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, p->ev[k], p->ev[k + 1]));
            p->stage_ms[k] += ms;
        }
    }
    return XH_OK;
}

extern "C" int xh_pipeline_set_timing(xh_pipeline* p, int on)
{
    p->timing = on != 0;
    for (int k = 0; k < 8; k++) p->stage_ms[k] = 0.0;
    return XH_OK;
}

extern "C" int xh_pipeline_stage_ms(xh_pipeline* p, double* out, int n)
{
    for (int k = 0; k < n && k < 8; k++) out[k] = p->stage_ms[k];
    return XH_OK;
}

extern "C" int xh_pipeline_members_per_request(xh_pipeline* p) { return p->mpr; }

extern "C" int xh_pipeline_work(xh_pipeline* p, int n, double* out)
{
    const int nm = n * p->mpr;
    hipLaunchKernelGGL(k_work_count, dim3(nm), dim3(256), 0, p->stream, (const Member*)p->d_members, p->L,
                       (const double*)p->d_cellArea, (const int*)p->d_valid, p->d_work);
    HIPCHK(hipGetLastError());
    // on the pipeline's stream: a blocking copy on the null stream would not
    // wait for a kernel on a non-blocking stream
    HIPCHK(hipMemcpyAsync(out, p->d_work, (size_t)nm * 8, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return XH_OK;
}

extern "C" const double* xh_pipeline_results(xh_pipeline* p) { return p->h_lnl; }

extern "C" int xh_pipeline_eval(xh_pipeline* p, int n, const xh_request* reqs, double* lnl_out, void* stream)
{
    int rc = xh_pipeline_launch(p, n, reqs, stream);
    if (rc != XH_OK) return rc;
    rc = xh_pipeline_sync(p);
    if (rc != XH_OK) return rc;
    std::memcpy(lnl_out, p->h_lnl, (size_t)n * 8);
    return XH_OK;
}

extern "C" int xh_pipeline_fetch(xh_pipeline* p, int what, int req, int index, double* out, size_t n)
{
    const Layout L = p->L;
    const size_t S = L.S_max;
    if (req < 0 || req >= p->maxb) return fail(XH_EINVAL, "xh_pipeline_fetch: request out of range");
    const bool by_region = (what == 1 || what == 8);
    if (index < 0 || index >= (by_region ? p->nreg : p->mpr))
        return fail(XH_EINVAL, "xh_pipeline_fetch: index out of range");
    const int mi = req * p->mpr + (by_region ? 0 : index);
    const double* src = nullptr;
    size_t cnt = 0;
    Member m{};
    // the last launch may sit on a non-blocking stream, which the blocking
    // copies below (null stream) would not wait for
    HIPCHK(hipStreamSynchronize(p->stream));
    HIPCHK(hipMemcpy(&m, p->d_members + mi, sizeof(Member), hipMemcpyDeviceToHost));
    switch (what) {
    case 0: src = p->d_flux + (size_t)mi * L.N_E * L.N_P; cnt = (size_t)L.N_E * L.N_P; break;
    case 1: src = p->d_folded + ((size_t)req * p->nreg + index) * p->N_ch * L.N_P; cnt = (size_t)p->N_ch * L.N_P; break;
    case 2: out[0] = m.S; if (n > 1) out[1] = m.status; if (n > 2) out[2] = m.numCell; return XH_OK;
    case 3: src = p->d_cellArea + (size_t)mi * S * S; cnt = (size_t)m.S * m.S; break;
    case 4: src = p->d_defl + (size_t)mi * S * L.N_R; cnt = (size_t)m.S * L.N_R; break;
    case 5: src = p->d_lag + (size_t)mi * S * L.N_R; cnt = (size_t)m.S * L.N_R; break;
    case 6: src = p->d_theta + (size_t)mi * S; cnt = m.S; break;
    case 7: src = p->d_phi + (size_t)mi * S; cnt = m.S; break;
    case 8: src = p->d_star + (size_t)req * p->N_ch * p->N_b; cnt = (size_t)p->N_ch * p->N_b; break;
    case 9: {
        // elsewhere spectrum [N_E]: ring partial sums x cell area x norm / (E keV)
        if (!p->has_else) return fail(XH_EINVAL, "xh_pipeline_fetch: the model has no elsewhere component");
        const Layout& Le = p->Le;
        std::vector<double> part((size_t)Le.S_max * L.N_E), E(L.N_E);
        double area = 0.0;
        Member em{};
        HIPCHK(hipMemcpy(&em, p->d_else + req, sizeof(Member), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(part.data(), p->d_else_part + (size_t)req * Le.S_max * L.N_E, part.size() * 8,
                         hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&area, p->d_else_area + req, 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(E.data(), p->d_energies, (size_t)L.N_E * 8, hipMemcpyDeviceToHost));
        const double bb_norm = 1.0e-7 * 5.040366110812353e22, kev = 1.60217662e-16;
        for (size_t e = 0; e < (size_t)L.N_E && e < n; e++) {
            double s = 0.0;
            for (int i = 0; i < em.S; i++) s += part[(size_t)i * L.N_E + e];
            out[e] = s * (area * bb_norm / (E[e] * kev));
        }
        return XH_OK;
    }
    default: return fail(XH_EINVAL, "xh_pipeline_fetch: unknown array");
    }
    if (cnt > n) cnt = n;
    HIPCHK(hipMemcpy(out, src, cnt * 8, hipMemcpyDeviceToHost));
    return XH_OK;
}
