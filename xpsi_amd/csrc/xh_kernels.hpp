// Kernel declarations and argument blocks shared by the launch code.
#pragma once
#include "xh_types.hpp"

namespace xh {

struct EnergyArgs {
    const double* flux;         // [member][N_E][N_P]
    const double* logE;         // N_E
    const double* logEdges;     // N_edges
    const double* atten;        // N_edges-1 or nullptr
    const int* member_index;    // [region][members_per_region], -1 = none
    double* out;                // [region][N_in][N_P]
    int N_E, N_P, N_edges, members_per_region;
    int interp;                 // phase interpolant (0 periodic Akima, 1 Steffen, 2 periodic cubic spline)
};

struct FoldGemmArgs {
    const double* M;            // [N_ch][ld_M]
    const double* in;           // [region][N_in][N_P]
    double* out;                // [region][N_ch][N_P]
    int N_ch, N_in, N_P, ld_M;
};

struct MarginalArgs {
    const double* signals;      // [request][comp][N_ch][N_P] (comp_np: component c's block is N_ch x comp_np[c])
    const double* comp_phases;  // N_P (cycles), shared; with comp_np the components' grids one after another
    const int* comp_np;         // [comp] phases per component, or nullptr (every component on the shared grid)
    const double* shifts;       // [request][comp]
    const double* phases;       // N_b + 1
    const double* counts;       // [N_ch][N_b]
    const double* precomp;      // N_ch
    const double* support;      // [N_ch][2]
    const double* background;   // [N_ch][N_b] or nullptr
    const int* request_status;  // [request] or nullptr
    double* ll;                 // [request][N_ch]
    int* status;                // [request][N_ch]
    double* B_ML;               // [request][N_ch]
    double* B_ML_supp;          // [request][N_ch]
    double* STAR;               // [request][N_ch][N_b] or nullptr
    double exposure, epsabs, epsrel, epsilon, sigmas, llzero, slim;
    int N_ch, N_b, N_P, n_comp;
    unsigned allow_negative;    // bit c: component c keeps negative bin integrals
    int interp;                 // phase interpolant: 0 periodic Akima, 1 Steffen, 2 periodic cubic spline
    // draw of the llzero-class values: *seed_ptr when set (device memory the
    // host refreshes before each launch, so replayed graphs draw afresh), else seed
    unsigned long long seed;
    const unsigned long long* seed_ptr;
};

__global__ void k_allocate(const RegionAlloc* regions, Member* members, int S_max);
__global__ void k_mesh_table(const Member* members, double* area_tab, double* total, int table_n);
__global__ void k_ring_bounds(const Member* members, const double* area_tab, const double* total, int table_n,
                              Layout L, double* theta, double* rr, double* maxAlpha, double* cos_gamma,
                              double* effGrav, double* ringlim);
__global__ void k_mesh_classify(const Member* members, const double* total, Layout L, const double* ringlim,
                                double* phi, double* cellArea, int2* blist, int* bcount);
__global__ void k_mesh_cells(const Member* members, const double* total, Layout L, const double* ringlim,
                             double* cellArea, const int2* blist, const int* bcount);
__global__ void k_rays_lane(const Member* members, Layout L, const double* rr, const double* maxAlpha_in,
                            double* cos_alpha, double* defl_raw, double* lag_raw, int* work, int* nwork,
                            const double* uu);
__global__ void k_rays(const Member* members, Layout L, const double* rr, const double* maxAlpha_in,
                       double* defl_raw, double* lag_raw, const int* work, const int* nwork,
                       const double* uu);
__global__ void k_rays_fix(Member* members, Layout L, const double* rr, double* maxAlpha,
                           double* defl, double* lag, double* maxDefl, int calibrate, const double* uu);
__global__ void k_energy_integrate(const Member* members, EnergyArgs A);
// the periodic cubic spline's energy integral (drop-in): block per phase column
__global__ void k_energy_integrate_cs(EnergyArgs A);
__global__ void k_fold_f64(FoldGemmArgs A);
__global__ void k_fold_f32(FoldGemmArgs A, const float* Mf);
template <int NT>
__global__ void k_fold_bf16(FoldGemmArgs A, const __bf16* Mh, const __bf16* Mm, const __bf16* Ml);
__global__ void k_marginal(MarginalArgs A);
__global__ void k_lnl_reduce(MarginalArgs A, const Member* members, int members_per_request,
                             double* lnl);

}  // namespace xh

namespace xh {
__global__ void k_ring_finalize(const Member* members, Layout L, const double* rr,
                                const double* cellArea, double* u, int* ring_ok, double* logT);
__global__ void k_ring_prep(int S, const double* theta2d, const double* params, const int* radiates,
                            double* theta, int* ring_ok, double* logT, double* logg);
// per-member, per-ring inputs
struct RingArgs {
    const double *theta, *r, *u, *cos_gamma, *maxDefl, *defl, *cos_alpha, *lag, *leaves;
    const int* ring_ok;
    double* geo;     // [member][ring][image][5][N_L]: PHASE, Z, ABB(mu eta), GEOM, PREV
    int* valid;      // [member][ring][image]
    int* lastq;      // [member][ring][image]: leaf index of the last intensity query
    int disk;        // a disk blocks rays (R_in below)
    double R_in;     // the disk's inner edge (m)
    double* pol;     // [member][ring][image][2][N_L]: cos 2chi, sin 2chi, or nullptr
};

struct FoldArgs {
    const double* geo;
    const int* valid;
    const double* cellArea;   // [member][S_max*S_max] (row stride S)
    const int* radiates;      // same layout or nullptr (then area > 0)
    const double* phi;        // [member][...], see phi_row_stride
    int phi_row_stride;       // 0: one azimuth row per member, else S
    const double* logT;       // per member per ring [member][S_max]
    const double* logg;       // per member per ring [member][S_max] (table atmosphere)
    const int* lastq;         // [member][ring][image]
    const double* atm_axes;   // table axes: logT, logg, mu, log10(E/kT) concatenated
    const double* atm_I;      // table values, C order (T, g, mu, E)
    int atm_N[4];
    const double* energies;   // N_E
    const double* phases;     // N_P (radians)
    double* flux;             // [member][N_E][N_P]
    int ring_block;
    // table atmosphere contracted over log T and log g per parallel,
    // [member][S_max][N_mu][N_E'] (k_table_slab), or nullptr
    const double* slab;
    // per axis d and base node b (0 .. N_d - 4) the four Lagrange
    // denominators 1 / prod_{m != k} (x_{b+k} - x_{b+m}), axes one after
    // another; or nullptr (then formed at each query)
    const double* atm_S;
    // wave fold: intervals per FP32 accumulation chain before its sums go to
    // the FP64 flux (0 = one chain per parallel)
    int chain;
};

// Stokes integrator with azimuthal invariance and the hot-region options
struct StokesArgs {
    const double* geo;         // [ring][image][5][N_L] from k_geometry
    const double* pol;         // [ring][image][2][N_L] cos 2chi, sin 2chi, or nullptr (I only)
    const int* valid;          // [ring][image]
    const double* cellArea;    // [S][S]
    const int* radiates;       // [S][S]
    const double* phi;         // [S][S]
    const double* params;      // [S][S][n_params], log10 T first
    int n_params;
    const double* correction;  // [S][S][2] elsewhere log10 T first, or nullptr
    int atm;                   // 1 blackbody, 3 burst blackbody
    int beam;                  // beaming option 0-3
    const double* energies;    // N_E
    const double* phases;      // N_P (radians)
    double* flux_I;            // [N_E][N_P], zeroed beforehand
    double* flux_Q;            // same, or nullptr
    double* flux_U;
    int interp;                // phase interpolant of the profiles
};
__global__ void k_stokes_fold(Member* members, Layout L, StokesArgs F);
inline size_t stokes_fold_lds(int N_L) { return (size_t)16 * N_L * 8; }

// integrator without azimuthal invariance (blackbody, per-cell temperatures)
struct GeneralArgs {
    const double* geo;         // [ring][image][4][N_L]: PHASE, Z, ABB, GEOM
    const int* valid;          // [ring][image]
    const double* cellArea;    // [S][S]
    const int* radiates;       // [S][S]
    const double* phi;         // [S][S]
    const double* params;      // [S][S][n_params], log10 T first
    const double* correction;  // [S][S][2] elsewhere log10 T first, or nullptr
    const double* energies;    // N_E
    const double* phases;      // N_P (radians)
    double* flux;              // [N_E][N_P], zeroed beforehand
    int n_params;
    int atm;                   // 1 blackbody, 3 burst blackbody
    int beam;                  // beaming option 0-3
    const double* pol;         // [ring][image][2] cos 2chi, sin 2chi (Stokes form) or nullptr
    double* flux_Q;            // [N_E][N_P] or nullptr
    double* flux_U;
    int interp;                // phase interpolant of GEOM
};
__global__ void k_general_geometry(Member* members, Layout L, RingArgs A);
__global__ void k_general_fold(Member* members, Layout L, GeneralArgs F);
inline size_t general_geo_lds(int N_R, int N_L) { return (size_t)(5 * N_R + 8 * N_L) * 8 + N_L + 16; }
inline size_t general_fold_lds(int N_L) { return (size_t)14 * N_L * 8; }

// wave-autonomous moment fold (precision form, blackbody): one wavefront per
// (ring block, 64 energies x 64 phases, member); flux zeroed beforehand
__global__ void k_wave_fold(Member* members, Layout L, FoldArgs F);
// the wave fold with a parallel's cell walk shared by its energy chunks:
// grid (S_max, phase chunks, members), 64 x ceil(N_E / 64) <= 256 threads
__global__ void k_wave_fold_shared(Member* members, Layout L, FoldArgs F);
__host__ __device__ inline size_t wave_fold_lds(int N_L, int S_max)
{
    return (((size_t)(3 * N_L + S_max) * 8 + (size_t)(S_max + 2 * N_L) * 4 + (size_t)N_L) + 15) & ~(size_t)15;
}

constexpr int WF_SLAB = 16;   // intervals per slab of k_wave_fold_shared's moments
inline size_t wave_fold_shared_lds(int N_L, int S_max)
{
    return wave_fold_lds(N_L, S_max) + (size_t)WF_SLAB * 4 * 64 * 4 + (size_t)WF_SLAB * 8 +
           (size_t)(WF_SLAB + 1) * 64 * 4;
}


// moment form: groups of T resident 16-energy chunks per block of shape T.
// Shape 2 takes 8 groups (256 energies), so a 256-energy member shares one
// block's per-parallel set-up per phase tile pair; the wave's 32 FP64
// accumulators spill under the 4-waves-per-SIMD register cap, which measured
// 6-7% faster than 4 groups on the batched workload.  Shape 4 keeps 2 groups
// and shape 1 keeps 4 (the accumulators of a wave stay within 16 doubles).
__host__ __device__ constexpr int moment_groups(int T) { return T == 4 ? 2 : (T == 2 ? 8 : 4); }

template <int OPT>
__global__ void k_geometry(Member* members, Layout L, RingArgs A);
__global__ void k_poisson_bg(MarginalArgs A);
__global__ void k_star_rate(MarginalArgs A, double* out);
__global__ void k_synth_totals(const double* rate, const double* bg, int n, double mul, double* scal);
__global__ void k_synth_draw(int mode, const double* rate, const double* bg, int n, double T, double n_star,
                             double n_bg, double* scal, unsigned long long seed, double* expec, double* synth);
__global__ void k_over(double* out, const double* x, size_t n, double a);
__global__ void k_spread(double* out, const double* v, int rows, int cols, int by_row);
__global__ void k_poisson_reduce(MarginalArgs A, double* lnl);
__global__ void k_gauss_bg(MarginalArgs A, const double* errors, double* tm);
__global__ void k_gauss_qnun(int n_rows, int N_b, const double* counts, const double* errors, const double* star,
                             double* tm);
__global__ void k_gauss_reduce(const double* tm, size_t n, double* lnl);
__global__ void k_phase_integrate(MarginalArgs A, double shift, double* out);
// the table atmosphere contracted over (log T, log g) for every parallel
__global__ void k_table_slab(const Member* members, Layout L, FoldArgs F, double* slab);
// closed mesh and time-invariant spectrum of the elsewhere component
__global__ void k_closed_table(const Member* members, double* area_tab, double* total, int table_n);
__global__ void k_closed_rings(const Member* members, const double* area_tab, const double* total,
                               int table_n, int S_max, double* colat, double* rr, double* maxAlpha,
                               double* cos_gamma, double* effGrav, double* cellArea);
__global__ void k_else_integrate(Member* members, Layout Le, const double* colat, const double* rr,
                                 const double* cos_gamma, const double* maxDefl, const double* defl,
                                 const double* cosa, const double* energies, int N_E, const double* phi,
                                 int phi_row_stride, const double* params, int param_stride, double* part);
__global__ void k_else_finalize(const Member* member, Layout Le, const double* part, double cellArea,
                                const double* energies, int N_E, double* out);
__global__ void k_else_add(const Member* else_members, Layout Le, const double* part,
                           const double* cellArea, const double* energies, int N_E, int N_P,
                           Member* hot_members, int members_per_request, double* flux);
template <int EC, int MAXK, int ATM>
__global__ void k_profile_fold(const Member* members, Layout L, FoldArgs F);
template <typename R, int ATM, int T, int G, int OCC>
__global__ void k_moment_fold(const Member* members, Layout L, FoldArgs F);
template <typename R, int ATM, int NQ>
__global__ void k_lane_fold(const Member* members, Layout L, FoldArgs F);
__global__ void k_fill(double* p, size_t n, double v);
__global__ void k_work_count(const Member* members, Layout L, const double* cellArea, const int* valid,
                             double* work);
__global__ void k_fill_int(int* p, size_t n, int v);
}
