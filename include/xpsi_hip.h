/*
 * C ABI of the MI355X (gfx950) X-PSI pulse-profile likelihood library.
 *
 * All array arguments are plain pointers with explicit sizes; "device"
 * pointers are HIP device memory (e.g. torch.Tensor.data_ptr() on a ROCm
 * device), "host" pointers are ordinary process memory.  `stream` is a
 * hipStream_t (NULL = the default stream).  Every entry point returns 0 on
 * success or a negative XH_E* code; xh_last_error() describes the failure.
 * Device memory is allocated only by xh_pipeline_create(),
 * xh_server_create() (the server's pipelines) and xh_atm_init() (a table's
 * device copy); the drop-ins work in the caller's buffers and workspaces.
 */
#ifndef XPSI_HIP_H
#define XPSI_HIP_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XH_OK 0
#define XH_EINVAL (-1)
#define XH_EHIP (-2)
#define XH_ENOMEM (-3)
#define XH_ERANGE (-4)

/* Message for the last failing call on this thread. */
const char* xh_last_error(void);

/* ------------------------------------------------------------------------
 * Drop-in kernels.  Each replaces one reference function; sizes follow the
 * reference's array shapes.  Device pointers, C order, float64 unless noted.
 * --------------------------------------------------------------------- */

/* Replaces xpsi/cellmesh/integrator_for_azimuthal_invariance.pyx:70-98
 * `integrate(...) -> (status, flux[N_E, N_P])` for the blackbody (hot_atm_ext
 * 1) and four-dimensional table (hot_atm_ext 2, hot_Num4D.pyx) atmospheres,
 * without a disk.  The reference's correction_srcCellParams (a blackbody
 * elsewhere temperature uniform over the cells, HotRegion.py:1056-1069) is
 * passed as correct = 1 and its log10 temperature else_logT; every cell then
 * subtracts that intensity and negative values are kept (:469-478, :593).
 * The table is passed as device
 * arrays: atm_axes = logT, logg, mu, log10(E/kT) concatenated, atm_N [4]
 * (int32, device) their lengths, atm_I the values in C order (T, g, mu, E);
 * all three may be NULL for the blackbody.
 * cellArea, theta, phi: [S, S]; radiates: int32 [S, S]; cellParams: [S, S, 2];
 * radial, r_s_over_r, maxDeflection, cos_gamma: [S];
 * deflection, cos_alpha, lag: [S, N_R]; energies [N_E] (keV);
 * leaves [N_L]; phases [N_P] (radians); flux out [N_E, N_P].
 * workspace: device scratch of xh_integrate_workspace_bytes() bytes.
 * *status_out (host) = 0 success, 1 numerical failure (the reference's ERROR). */
size_t xh_integrate_workspace_bytes(int S, int N_R, int N_L, int N_E, int N_P);
int xh_integrate(double omega, double r_s, double inclination, int S,
                 const double* cellArea, const double* radial,
                 const double* r_s_over_r, const double* theta,
                 const double* phi, const double* cellParams,
                 const int* radiates, int N_R, const double* deflection,
                 const double* cos_alpha, const double* lag,
                 const double* maxDeflection, const double* cos_gamma,
                 int N_E, const double* energies, int N_L,
                 const double* leaves, int N_P, const double* phases,
                 int hot_atm_ext, const double* atm_axes, const int* atm_N,
                 const double* atm_I, int image_order_limit, int correct,
                 double else_logT, double* flux, void* workspace,
                 size_t workspace_bytes, int* status_out, void* stream);

/* Replaces xpsi/cellmesh/integratorIQU_for_azimuthal_invariance.pyx:68-746
 * `integrate(...) -> (status, I, Q, U)` (flux_Q, flux_U non-NULL), and serves
 * integrator_for_azimuthal_invariance.pyx:70-665 for the options xh_integrate
 * does not take (flux_Q = flux_U = NULL: intensity only): the burst blackbody
 * atmosphere (atm_ext 3, hot_BB_burst.pyx: I = B (0.421 + 0.868 mu), Q = I *
 * 0.1171 (mu - 1) / (1 + 3.582 mu)), beaming (beam_opt 1-3,
 * hot_wrapper.pyx:115-199; 3 with the blackbody only), disk blocking (R_in <
 * 1e6 m, common_functions.pyx disk_block) and the polarisation angle
 * (common_functions.pyx compute_pol_ang).  cellParams [S, S, n_params]: log10 T,
 * log g, and with beaming a, b, c, d, n_mu; correction [S, S, 2] (blackbody
 * elsewhere log10 T first) or NULL.  Other arrays as xh_integrate.  13 N_L and
 * 5 N_R + 8 N_L doubles must each fit in 159 KiB of one block's LDS.
 * workspace: xh_integrate_stokes_workspace_bytes() bytes.
 * *status_out = 0 success, 1 numerical failure (the reference's ERROR). */
size_t xh_integrate_stokes_workspace_bytes(int S, int N_R, int N_L);
int xh_integrate_stokes(double omega, double inclination, int S,
                        const double* cellArea, const double* radial,
                        const double* r_s_over_r, const double* theta,
                        const double* phi, const double* cellParams, int n_params,
                        const int* radiates, int N_R, const double* deflection,
                        const double* cos_alpha, const double* lag,
                        const double* maxDeflection, const double* cos_gamma,
                        int N_E, const double* energies, int N_L,
                        const double* leaves, int N_P, const double* phases,
                        int atm_ext, int beam_opt, int image_order_limit, double R_in,
                        const double* correction, double* flux_I, double* flux_Q,
                        double* flux_U, void* workspace, size_t workspace_bytes,
                        int* status_out, void* stream);

/* Replaces xpsi/cellmesh/integrator.pyx:48-667 `integrate(...) -> (status,
 * flux[N_E, N_P])`, the integrator without azimuthal invariance
 * (HotRegion.symmetry False), and its Stokes form
 * xpsi/cellmesh/integratorIQU.pyx:51-717 `-> (status, I, Q, U)` (flux_Q,
 * flux_U non-NULL): every cell has its own parameters.  Same arrays as
 * xh_integrate; cellParams [S, S, n_params]: log10 T, log g, and with beaming
 * a, b, c, d, n_mu; atm_ext 1 (blackbody) or 3 (burst blackbody, with its Q);
 * beam_opt 0-3 (3 with the blackbody); correction: [S, S, 2] per-cell
 * blackbody elsewhere parameters (log10 T first) whose intensity every cell
 * subtracts (else_atm_ext 1), or NULL.  The Stokes form uses one polarisation
 * angle per parallel and image, formed as the reference forms it from its
 * leaf loop's last values.  13 N_L and 5 N_R + 8 N_L doubles must each fit in
 * 159 KiB of one block's LDS (N_L up to about 1500).
 * workspace: xh_integrate_general_workspace_bytes() bytes.
 * *status_out = 0 success, 1 numerical failure (the reference's ERROR). */
size_t xh_integrate_general_workspace_bytes(int S, int N_L);
int xh_integrate_general(double omega, double inclination, int S,
                         const double* cellArea, const double* radial,
                         const double* r_s_over_r, const double* theta,
                         const double* phi, const double* cellParams, int n_params,
                         const int* radiates, int N_R, const double* deflection,
                         const double* cos_alpha, const double* lag,
                         const double* maxDeflection, const double* cos_gamma,
                         int N_E, const double* energies, int N_L,
                         const double* leaves, int N_P, const double* phases,
                         int image_order_limit, int atm_ext, int beam_opt,
                         const double* correction, double* flux, double* flux_Q,
                         double* flux_U, void* workspace, size_t workspace_bytes,
                         int* status_out, void* stream);

/* Replaces xpsi/cellmesh/integrator_for_time_invariance.pyx:59-98
 * `integrate(...) -> (status, flux[N_E])` for the blackbody elsewhere
 * atmosphere (else_atm_ext 1).  S rings of S cells; cellArea: the common cell
 * area (scalar); radial, colat (= theta[:, 0]), maxDeflection, cos_gamma: [S];
 * phi: [S, S]; cellParams: [S, S, param_stride] with log10 T first;
 * deflection, cos_alpha: [S, N_R]; energies [N_E] (keV); flux out [N_E].
 * workspace: xh_integrate_time_invariant_workspace_bytes() bytes.
 * *status_out = 0 success, 1 a deflection outside the ray table (ERROR). */
size_t xh_integrate_time_invariant_workspace_bytes(int S, int N_E);
int xh_integrate_time_invariant(double omega, double r_s, double inclination,
                                int S, double cellArea, const double* radial,
                                const double* colat, const double* phi,
                                const double* cellParams, int param_stride,
                                int N_R, const double* deflection,
                                const double* cos_alpha,
                                const double* maxDeflection,
                                const double* cos_gamma, int N_E,
                                const double* energies, int image_order_limit,
                                double* flux, void* workspace,
                                size_t workspace_bytes, int* status_out,
                                void* stream);

/* Replaces the atmosphere extension point of the integrators,
 * xpsi/surface_radiation_field/hot_wrapper.pyx:72-252 (init_hot,
 * eval_hot_I, eval_hot_norm, free_hot), for the blackbody (ext 1,
 * hot_BB.pyx:54-98) and the 4-D table (ext 2, hot_Num4D.pyx:66-460; axes
 * logT, logg, mu, log10(E/kT) host arrays of lengths N[4], values I host,
 * C order (T, g, mu, E) as preload.pyx:6-44 lays them out).  xh_atm_eval_I
 * evaluates n queries at once (device arrays E [n] keV, mu [n], VEC [n, 2] =
 * (log10 T, log10 g)) into out [n], the intensity without the norm, with the
 * values one reference thread returns evaluating the queries in order from
 * a fresh init_hot (the table's base-node history included); xh_atm_norm
 * returns eval_hot_norm's factor.  xh_atm_init allocates the table's device
 * copy; xh_atm_free releases it. */
typedef struct xh_atmosphere xh_atmosphere;
xh_atmosphere* xh_atm_init(int ext, const double* const* axes, const int* N,
                           const double* I);
int xh_atm_eval_I(xh_atmosphere* a, int n, const double* E, const double* mu,
                  const double* VEC, double* out, void* stream);
double xh_atm_norm(const xh_atmosphere* a);
int xh_atm_free(xh_atmosphere* a);

/* Replaces xpsi/cellmesh/rays.pyx:249-390 `compute_rays(N_T, N_u, r_s,
 * r_s_over_R[N_u], maxAlpha[N_u], N_R) -> (terminate, deflection, cos_alpha,
 * lag, maxDeflection)`: per u = r_s/R, N_R rays with cos(alpha) uniform from 1
 * to cos(maxAlpha); deflection and (uncalibrated) lag by adaptive
 * Gauss-Kronrod quadrature to 1e-12 relative.  maxAlpha (device) is clamped
 * in place as the reference does; bad rays are repaired linearly in the
 * reference's order.  deflection, cos_alpha, lag: [N_u, N_R]; maxDeflection
 * [N_u] (device).  *terminate_out (host) = 1 when a parallel's last ray
 * fails (the reference's ERROR).  workspace of xh_compute_rays_workspace_bytes(). */
size_t xh_compute_rays_workspace_bytes(int N_u, int N_R);
int xh_compute_rays(int N_u, double r_s, const double* r_s_over_R,
                    double* maxAlpha, int N_R, double* deflection,
                    double* cos_alpha, double* lag, double* maxDeflection,
                    void* workspace, size_t workspace_bytes,
                    int* terminate_out, void* stream);

/* Replaces xpsi/cellmesh/mesh_tools.pyx:892-1099 `allocate_cells(...) ->
 * (super_sqrt_num_cells, super_num_cells, cede_sqrt_num_cells,
 * cede_num_cells)`, the four values written to out4 (host); same argument
 * order.  workspace of xh_allocate_cells_workspace_bytes(). */
size_t xh_allocate_cells_workspace_bytes(void);
int xh_allocate_cells(int num_cells, int min_sqrt_num_cells,
                      int max_sqrt_num_cells, double mass, double r_s,
                      double R_eq, double Omega, double zeta, double epsilon,
                      int star_shape_ind, double superRadius,
                      double cedeRadius, double superColatitude,
                      double cedeColatitude, double superAzimuth,
                      double holeColatitude, double holeRadius,
                      double holeAzimuth, double* out4, void* workspace,
                      size_t workspace_bytes, void* stream);

/* Replaces xpsi/cellmesh/mesh.pyx:23-417 construct_spot_cellMesh (polar = 0)
 * and xpsi/cellmesh/polar_mesh.pyx:23-430 construct_polar_cellMesh (polar =
 * 1): (theta, phi, r, cellAreas, maxEmissionAngle, cos_gamma, effGrav) for
 * the region of cedeRadius at cedeColatitude less the superseding region.
 * Outputs (device): theta, phi, cellArea [S, S]; r, maxEmissionAngle,
 * cos_gamma, effGrav [S].  *status_out (host) 0 on success.
 * workspace of xh_cell_mesh_workspace_bytes(S). */
size_t xh_cell_mesh_workspace_bytes(int S);
int xh_construct_cell_mesh(int polar, double numCell, int S, double mass,
                           double r_s, double R_eq, double zeta,
                           double epsilon, int star_shape_ind,
                           double cedeRadius, double cedeColatitude,
                           double superRadius, double superColatitude,
                           double superAzimuth, double* theta, double* phi,
                           double* r, double* cellArea,
                           double* maxEmissionAngle, double* cos_gamma,
                           double* effGrav, void* workspace,
                           size_t workspace_bytes, int* status_out,
                           void* stream);

/* Replaces xpsi/tools/energy_integrator.pyx:27-30,114
 * `energy_integrator(N_Ts, signal[N_E, N_P], log10_E[N_E], log10_edges[N_in+1])`
 * -> out [N_in, N_P]. */
int xh_energy_integrator(int N_E, int N_P, const double* signal,
                         const double* log10_energies, int N_edges,
                         const double* log10_edges, double* out, void* stream);

/* Replaces xpsi/Instrument.py:163-197 `matrix @ signal`:
 * out[N_ch, N_P] = M[N_ch, N_in] (row stride ld_M) @ signal[N_in, N_P]. */
int xh_fold(int N_ch, int N_in, int N_P, const double* M, int ld_M,
            const double* signal, double* out, void* stream);

/* Replaces xpsi/likelihoods/default_background_marginalisation.pyx:38-68
 * `precomputation(int32 counts[N_ch, N_b]) -> -sum ln n!` (host memory). */
int xh_precomputation(int N_ch, int N_b, const int* counts, double* out);

/* Replaces default_background_marginalisation.pyx:450-466,731-734
 * `eval_marginal_likelihood(...) -> (lnL, STAR, B_ML, B_ML_supp)` for
 * n_comp components sharing one phase grid comp_phases [N_P] (cycles).
 * components: [n_comp, N_ch, N_P]; phases [N_b+1]; counts [N_ch, N_b];
 * shifts [n_comp]; precomp [N_ch]; support [N_ch, 2]; background [N_ch, N_b]
 * or NULL; allow_negative nonzero lets every component keep negative bin
 * integrals.  lnl_out is a host pointer; STAR [N_ch, N_b], B_ML, B_ML_supp
 * [N_ch] device outputs; workspace of xh_marginal_workspace_bytes().  At most
 * 32 components. */
size_t xh_marginal_workspace_bytes(int N_ch);
int xh_eval_marginal_likelihood(double exposure_time, int N_b,
                                const double* phases, int N_ch,
                                const double* counts, int n_comp,
                                const double* components, int N_P,
                                const double* comp_phases,
                                const double* shifts, const double* precomp,
                                const double* support, double epsabs,
                                double epsrel, double epsilon, double sigmas,
                                double llzero, int allow_negative,
                                double slim, const double* background,
                                double* lnl_out, double* STAR, double* B_ML,
                                double* B_ML_supp, void* workspace,
                                size_t workspace_bytes, void* stream);

/* xh_eval_marginal_likelihood with the components as the reference takes them
 * (xpsi/tools/compute_expected_counts.pyx:26-197): each component on its own
 * phase grid, and allow_negative per component.  components: the components'
 * [N_ch][comp_num_phases[c]] blocks one after another (device);
 * comp_num_phases: phases per component (device, n_comp ints); comp_phases:
 * the grids one after another (device); allow_negative: n_comp flags (host,
 * nullptr = none).  At most 32 components. */
int xh_eval_marginal_likelihood_ex(double exposure_time, int N_b, const double* phases, int N_ch,
                                   const double* counts, int n_comp, const double* components,
                                   const int* comp_num_phases, const double* comp_phases,
                                   const double* shifts, const double* precomp, const double* support,
                                   double epsabs, double epsrel, double epsilon, double sigmas,
                                   double llzero, const int* allow_negative, double slim,
                                   const double* background, double* lnl_out, double* STAR,
                                   double* B_ML, double* B_ML_supp, void* workspace,
                                   size_t workspace_bytes, void* stream);

/* Replaces xpsi/likelihoods/_poisson_likelihood_given_background.pyx:14-113
 * `poisson_likelihood_given_background(...) -> (lnL, EXPEC)`: the Poisson
 * likelihood of the counts given the star's signal plus a known background.
 * components [n_comp, N_ch, N_P] on one phase grid comp_phases [N_P]
 * (cycles); phases [N_b+1]; counts [N_ch, N_b]; shifts [n_comp];
 * background [N_ch, N_b] count rates (required); precomp [N_ch] (the
 * -sum ln n! of xh_precomputation) or NULL.  EXPEC [N_ch, N_b] (device) gets
 * (star + background) x exposure_time; lnl_out is a host pointer.  A bin with
 * data but no expected counts adds -1e90 (0.1 + 0.9 U), U uniform, fresh per
 * call.  workspace of xh_poisson_workspace_bytes(). */
size_t xh_poisson_workspace_bytes(int N_ch);
int xh_poisson_likelihood_given_background(double exposure_time, int N_b,
                                           const double* phases, int N_ch,
                                           const double* counts, int n_comp,
                                           const double* components, int N_P,
                                           const double* comp_phases,
                                           const double* shifts,
                                           const double* background,
                                           const double* precomp,
                                           int allow_negative, double* lnl_out,
                                           double* EXPEC, void* workspace,
                                           size_t workspace_bytes, void* stream);

/* xh_poisson_likelihood_given_background with components on their own phase
 * grids and allow_negative per component, laid out as for
 * xh_eval_marginal_likelihood_ex. */
int xh_poisson_likelihood_given_background_ex(double exposure_time, int N_b, const double* phases, int N_ch,
                                              const double* counts, int n_comp, const double* components,
                                              const int* comp_num_phases, const double* comp_phases,
                                              const double* shifts, const double* background,
                                              const double* precomp, const int* allow_negative,
                                              double* lnl_out, double* EXPEC, void* workspace,
                                              size_t workspace_bytes, void* stream);

/* Replaces xpsi/likelihoods/_gaussian_likelihood_given_background_IQU.pyx:14-91
 * `gaussian_likelihood_given_background(exposure_time, phases, counts, errors,
 * components, component_phases, phase_shifts, background, allow_negative) ->
 * (LOGLIKE, EXPEC)`: EXPEC [N_ch, N_b] = (phase-binned star count rate +
 * background count rate) x exposure_time, as the Poisson statistic forms it
 * (compute_expected_counts.pyx:199-315), and LOGLIKE = -sum over channels and
 * bins of (EXPEC - counts)^2 / (2 s^2) + ln(2 pi s^2) / 2 with s = errors
 * [N_ch, N_b], summed in the reference's order.  Device inputs as for
 * xh_poisson_likelihood_given_background; *lnl_out to host memory; EXPEC may
 * be NULL.  workspace of xh_gaussian_workspace_bytes(N_ch, N_b). */
size_t xh_gaussian_workspace_bytes(int N_ch, int N_b);
int xh_gaussian_likelihood_given_background(double exposure_time, int N_b, const double* phases, int N_ch,
                                            const double* counts, const double* errors, int n_comp,
                                            const double* components, int N_P, const double* comp_phases,
                                            const double* shifts, const double* background, int allow_negative,
                                            double* lnl_out, double* EXPEC, void* workspace,
                                            size_t workspace_bytes, void* stream);

/* xh_gaussian_likelihood_given_background with components on their own phase
 * grids and allow_negative per component, laid out as for
 * xh_eval_marginal_likelihood_ex. */
int xh_gaussian_likelihood_given_background_ex(double exposure_time, int N_b, const double* phases, int N_ch,
                                               const double* counts, const double* errors, int n_comp,
                                               const double* components, const int* comp_num_phases,
                                               const double* comp_phases, const double* shifts,
                                               const double* background, const int* allow_negative,
                                               double* lnl_out, double* EXPEC, void* workspace,
                                               size_t workspace_bytes, void* stream);

/* Replaces xpsi/likelihoods/_gaussian_likelihood_QnUn.pyx:14-44
 * `gaussian_likelihood_QnUn(phases, counts, errors, star) -> (LOGLIKE, star)`:
 * counts, errors [N_rows, N_b] and one model row star [N_b] (device memory;
 * the phase edges are not used, as in the reference); LOGLIKE = -sum over rows
 * and bins of (star_j - counts_ij)^2 / (2 s_ij^2) + ln(2 pi s_ij^2) / 2 to
 * host memory.  workspace of xh_gaussian_workspace_bytes(N_rows, N_b). */
int xh_gaussian_likelihood_QnUn(int N_rows, int N_b, const double* counts, const double* errors,
                                const double* star, double* lnl_out, void* workspace, size_t workspace_bytes,
                                void* stream);

/* Replaces xpsi/tools/phase_integrator.pyx:23-121 `phase_integrator(
 * exposure_time, phases, signal, signal_phases, phase_shift, allow_negative)`:
 * each row of signal [N_rows, N_P] (device, on the grid signal_phases [N_P],
 * N_P >= 2) interpolated with the process-wide phase interpolant and
 * integrated over the intervals phases [N_b + 1] shifted by phase_shift, a
 * piece kept where positive or allow_negative, times exposure_time, into out
 * [N_rows, N_b] (device).  Stream-ordered; allocates nothing. */
int xh_phase_integrator(double exposure_time, int N_b, const double* phases, int N_rows, const double* signal,
                        int N_P, const double* signal_phases, double phase_shift, int allow_negative, double* out,
                        void* stream);

/* Replace xpsi/tools/synthesise.pyx:40-148 `synthesise_exposure(...) ->
 * (EXPEC, SYNTHETIC, SCALE_BACKGROUND)` and :152-276
 * `synthesise_given_total_count_number(...) -> (EXPEC, SYNTHETIC, SCALE_STAR,
 * SCALE_BACKGROUND / SCALE_STAR)`: expected counts of the phase-binned star
 * signal plus the background scaled to expected_background_counts, and one
 * Poisson realisation per (channel, bin) drawn on the device from `seed`
 * (splitmix64 streams; sequential inversion below mean 30, Hormann's PTRS
 * above; the reference's GSL generator is not reproduced, only the
 * distribution).  components [n_comp, N_ch, N_P] on one phase grid
 * comp_phases [N_P]; phases [N_b+1]; background [N_ch, N_b] expected counts;
 * EXPEC, SYNTHETIC [N_ch, N_b] device outputs; the scales to host memory
 * (scales_out: two values).  workspace of xh_synthesise_workspace_bytes(). */
size_t xh_synthesise_workspace_bytes(int N_ch, int N_b);
int xh_synthesise_exposure(double exposure_time, int N_b, const double* phases,
                           int N_ch, int n_comp, const double* components,
                           int N_P, const double* comp_phases,
                           const double* shifts,
                           double expected_background_counts,
                           const double* background, int allow_negative,
                           unsigned long long seed, double* EXPEC,
                           double* SYNTHETIC, double* scale_background,
                           void* workspace, size_t workspace_bytes,
                           void* stream);
int xh_synthesise_given_total_count_number(int N_b, const double* phases,
                                           double expected_star_counts,
                                           int N_ch, int n_comp,
                                           const double* components, int N_P,
                                           const double* comp_phases,
                                           const double* shifts,
                                           double expected_background_counts,
                                           const double* background,
                                           int allow_negative,
                                           unsigned long long seed,
                                           double* EXPEC, double* SYNTHETIC,
                                           double* scales_out, void* workspace,
                                           size_t workspace_bytes, void* stream);

/* No counterpart in the reference: xh_eval_marginal_likelihood without the
 * workspace argument, for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its scratch is a static device array of the library, one per
 * GPU, holding XH_NOALLOC_MAX_CH channels; more channels fail with
 * XH_ERANGE.  Calls that overlap on one GPU share that array and race.
 * It allocates nothing; results equal xh_eval_marginal_likelihood's. */
#define XH_NOALLOC_MAX_CH 16384
int xh_eval_marginal_likelihood_noalloc(double exposure_time, int N_b,
                                        const double* phases, int N_ch,
                                        const double* counts, int n_comp,
                                        const double* components, int N_P,
                                        const double* comp_phases,
                                        const double* shifts,
                                        const double* precomp,
                                        const double* support, double epsabs,
                                        double epsrel, double epsilon,
                                        double sigmas, double llzero,
                                        int allow_negative, double slim,
                                        const double* background,
                                        double* lnl_out, double* STAR,
                                        double* B_ML, double* B_ML_supp,
                                        void* stream);

/* ------------------------------------------------------------------------
 * Whole-likelihood pipeline: mesh, rays, integration, registration and the
 * marginal likelihood for a batch of parameter vectors, all on the device.
 * Replaces the body of xpsi/Likelihood.py:297-432 (_driver) for models made
 * of up to two circular hot regions (each optionally with a ceding ring).
 * --------------------------------------------------------------------- */
typedef struct xh_region_config {
    int cede;                 /* ceding ring present */
    int is_antiphased;
    int sqrt_num_cells, min_sqrt_num_cells, max_sqrt_num_cells;
    int omit;                 /* a circular region cut out of the superseding one
                                 (HotRegion omit=True; ignored with a ceding ring,
                                 as the reference ignores it) */
} xh_region_config;

typedef struct xh_model_config {
    int n_regions;            /* 1 or 2 */
    xh_region_config region[2];
    int num_rays, num_leaves, num_phases, image_order_limit, atm_ext;
    const double* atm_axes[4];/* atm_ext 2: table axes (host) logT, logg, mu, log10(E/kT) */
    int atm_N[4];
    const double* atm_I;      /* table values (host), C order (T, g, mu, E) */
    int star_shape;           /* 0 oblate (AGM_14), 1 sphere */
    double frequency, mode_frequency;
    int N_E;                  /* energies (host, keV) */
    const double* energies;
    int N_in;                 /* input bins: edges [N_in + 1] keV (host) */
    const double* energy_edges;
    const double* attenuation;/* [N_in] multiplier or NULL */
    int N_ch;                 /* response [N_ch, N_in] (host) */
    const double* matrix;
    int N_b;                  /* data: phases [N_b+1], counts [N_ch, N_b] */
    const double* data_phases;
    const double* counts;
    double exposure_time;
    const double* support;    /* [N_ch, 2] or NULL (unbounded) */
    double epsabs, epsrel, epsilon, sigmas, llzero, slim;
    int max_batch;            /* requests per xh_pipeline_eval call */
    int max_sqrt_cells;       /* buffer bound on mesh size (>= max_sqrt_num_cells) */
    int precision;            /* 0: FP64 throughout.  1: precision form - FP32
                                 intensity and spline fold (matrix cores), response fold
                                 on FP32 MFMA with FP64 sums per 4-deep slice (fold_terms
                                 selects split-BF16 instead); rays, phase-bin accumulation
                                 and likelihood quadrature stay FP64 */
    int fold_kernel;          /* 0 automatic (the wave fold for the precision form with
                                 the blackbody, else the moment form, else energy lanes,
                                 else per-point), 1 per-point fold (FP64 only), 2 moment form,
                                 3 energy lanes, 4 wave-autonomous moment fold (precision
                                 form, blackbody; one wavefront per 64 energies x 64 phases
                                 x a block of parallels, no block barriers);
                                 xpsi_amd.pipeline.autotune_fold measures */
    /* time-invariant blackbody emission from the rest of the surface
       (xpsi/Elsewhere.py): a closed mesh of else_sqrt_num_cells^2 cells and
       else_num_rays rays per ring; its spectrum joins the first hot region's
       first component and every hot cell subtracts it (no clipping) */
    int elsewhere;            /* 0 none, 1 present */
    int else_sqrt_num_cells;  /* even, >= 10 (reference default 64) */
    int else_num_rays;        /* reference default 1000 */
    int else_image_order_limit; /* 0 = inferred from the deflection */
    int fold_terms;           /* precision form's response fold: 0 FP32 MFMA with FP64
                                 sums per 4-deep slice; 3 or 6 bf16 products of a
                                 three-term split; see xpsi_amd/csrc/xh_signal.hip */
} xh_model_config;

/* One parameter vector.  Per region: phase_shift, super_colatitude,
 * super_radius, super_temperature, cede_colatitude, cede_radius,
 * cede_azimuth, cede_temperature (log10 K, angles in radians). */
typedef struct xh_request {
    double mass, radius, distance, cos_inclination;
    double region[2][8];
    int sqrt_num_cells[2];    /* 0 = the configured value */
    double elsewhere_temperature; /* log10 K, used when the model has elsewhere */
    double omit[2][3];        /* omit_colatitude, omit_radius, omit_azimuth (radians)
                                 of regions configured with omit */
} xh_request;

typedef struct xh_pipeline xh_pipeline;

xh_pipeline* xh_pipeline_create(const xh_model_config* cfg);
void xh_pipeline_destroy(xh_pipeline* p);
/* Evaluate n <= max_batch requests; lnl_out host [n].  Blocks until done.
 * Every launch draws its own llzero-class values (requests that fail or fall
 * below llzero get llzero * U(0.1, 1), fresh per launch, as the reference's
 * random_near_llzero).  A launch while the previous one is unwaited first
 * waits for it (the pinned descriptors and results are reused). */
int xh_pipeline_eval(xh_pipeline* p, int n, const xh_request* reqs,
                     double* lnl_out, void* stream);
/* Same without waiting: results land in the pipeline's host buffer, read
 * with xh_pipeline_results() after xh_pipeline_sync(). */
int xh_pipeline_launch(xh_pipeline* p, int n, const xh_request* reqs, void* stream);
int xh_pipeline_sync(xh_pipeline* p);
const double* xh_pipeline_results(xh_pipeline* p);

/* Per-stage device time: record HIP events around the kernel stages of every
 * launch (on the launch stream) when on != 0; xh_pipeline_stage_ms() returns
 * the accumulated milliseconds of mesh, rays, geometry, profile-and-fold,
 * energy-and-response and marginal stages (6 values) since the last reset. */
int xh_pipeline_set_timing(xh_pipeline* p, int on);
int xh_pipeline_stage_ms(xh_pipeline* p, double* out, int n);
/* Members (super + ceding) per request. */
int xh_pipeline_members_per_request(xh_pipeline* p);
/* Work accounting of the last launch of n requests: per member, the number
 * of (radiating cell, parallel, image order) triples folded; the phase-fold
 * evaluation count is that times N_E * N_P.  out: host [n * members]. */
int xh_pipeline_work(xh_pipeline* p, int n, double* out);

/* Tuning: energies per block of the profile-and-fold kernel (8, 16, 32, 64;
 * 0 = chosen from the leaf count).  Applies to pipelines created afterwards
 * and to xh_integrate calls. */
int xh_set_fold_energy_chunk(int ec);

/* Fold kernel of xh_integrate: 0 automatic (the moment form when its LDS
 * fits, else energy lanes, else per-point), 1 per-point evaluation,
 * 2 moment form on the matrix cores, 3 energy lanes.  All give the
 * reference's sums (the moment form evaluates sign-changing spline pieces
 * point by point, so the >0 clipping is kept). */
int xh_set_fold_kernel(int kind);

/* Tuning: block shape of the moment-form fold: T 16-phase tiles x T resident
 * 16-energy chunks per block of eight waves (T = 1, 2 or 4; larger shapes
 * share each parallel's moments and coefficients over more of the output but
 * need more LDS and give fewer blocks).  0 = automatic: the largest shape
 * whose LDS fits for batches of 32 members or more, 1 below.  Applies to
 * pipelines created afterwards and to xh_integrate;
 * xpsi_amd.pipeline.autotune_fold measures the choices. */
int xh_set_moment_shape(int tiles);

/* Tuning: parallels per block of the moment-form and energy-lane folds for
 * batches of 32 members or more (blocks of one member meet by atomic adds; 0 =
 * automatic: all of a member's parallels in one moment-form block, 8 per
 * energy-lane block).  Applies to pipelines created afterwards (and to later
 * xh_integrate calls). */
int xh_set_fold_rings(int rings);

/* Tuning: parallels per wavefront of the wave fold (fold_kernel 4), at any
 * batch size (0 = automatic: 1, measured fastest on the batched workload).
 * Applies to pipelines created afterwards. */
int xh_set_wave_fold_rings(int rings);

/* Tuning: intervals per FP32 accumulation chain of the wave fold before the
 * chain's sums are added to the FP64 flux (default 8; 0 = one chain per
 * parallel).  Longer chains issue fewer atomic additions but round more, and
 * that rounding is shared by neighbouring bins, so it acts on the likelihood
 * as a scale error: on the 256-leaf two-region model with 300 channels x 64
 * bins one chain per parallel moved |dlnL| to 2.1e-3, chains of 8 to 4.5e-4.
 * Applies to pipelines created afterwards. */
int xh_set_wave_fold_chain(int intervals);

/* Tuning: the wave fold with one block per parallel whose 64-energy chunks
 * (two to four of them) share the phase lanes' cell walk: the walk's moments
 * go through LDS in slabs of 16 intervals, and each wavefront folds its own
 * energies against them (1 = on where it applies, 0 = off, the default).
 * Applies to pipelines created afterwards.  No counterpart in the reference
 * (a launch form of this fold). */
int xh_set_wave_fold_shared(int on);

/* Replaces xpsi/tools/core.pyx:21-116 set_phase_interpolant /
 * get_phase_interpolant: the process-wide phase interpolant of the
 * integrators' profiles, the energy integral and the likelihoods' phase-bin
 * integrals: 0 periodic Akima (default), 1 Steffen, 2 periodic cubic spline.
 * Read at each drop-in call; xh_integrate (whose folds are periodic Akima)
 * and xh_pipeline_create refuse a choice other than 0. */
int xh_set_phase_interpolant(int kind);
int xh_get_phase_interpolant(void);

/* Tuning: phases per wave of the energy-lane fold (8 or 16).  Applies to
 * pipelines created afterwards and to xh_integrate. */
int xh_set_lane_phases(int nq);

/* Replay each launch as a hipGraph captured once per batch size (requires a
 * non-default stream; timing mode launches eagerly). */
int xh_pipeline_use_graphs(xh_pipeline* p, int on);
/* Capture (without running) the graph of batch size n on stream, so that a
 * later launch of that size replays at once. */
int xh_pipeline_capture(xh_pipeline* p, int n, void* stream);

/* Prior inverse sampling for a batch of proposals (xpsi/Prior.py:84-160 with
 * the examples' compactness prior, examples/examples_fast/Modules/CustomPrior.py:49-66):
 * params[q][k] = lo_k + (hi_k - lo_k) hypercube[q][k] for n points of d
 * coordinates (device, [n, d]); bounds [d, 2] (device); logprior[q] = 0 when
 * every coordinate lies within its bounds and, when idx_mass and idx_radius
 * are >= 0, the star of that mass (Msun) and equatorial radius (km) spinning
 * at `frequency` Hz passes the compactness cuts, else -inf. */
int xh_prior_inverse_sample(int n, int d, const double* hypercube,
                            const double* bounds, int idx_mass, int idx_radius,
                            double frequency, double* params, double* logprior,
                            void* stream);

/* ------------------------------------------------------------------------
 * Continuous-batching server: n_slots pipelines of slot_batch requests, each
 * on its own stream; a dispatcher thread admits queued requests into free
 * slots while the in-flight device footprint stays within hbm_budget_bytes
 * (<= 0: unlimited).  Additive to the reference API (its samplers call the
 * likelihood one vector at a time, xpsi/Sample.py:32-130).
 * --------------------------------------------------------------------- */
typedef struct xh_server xh_server;
xh_server* xh_server_create(const xh_model_config* cfg, int n_slots, int slot_batch,
                            double hbm_budget_bytes, int use_graphs);
int xh_server_submit(xh_server* s, const xh_request* req, long long* id_out);
/* Finished requests: ids, lnL and submit-to-result latency in seconds. */
int xh_server_poll(xh_server* s, long long* ids, double* lnl, double* latency_s, int max_n);
/* queued, in-flight bytes, launched batches, completed, mean batch fill,
 * peak in-flight bytes, largest batch, a failed batch seen (0/1): out[<= 8] */
int xh_server_stats(xh_server* s, double* out, int n);
/* Batch formation: a free slot takes a partial batch only once the oldest
 * queued request has waited max_wait_s seconds, or when no slot is busy
 * (default 0: dispatch at once).  A full batch is always dispatched. */
int xh_server_set_policy(xh_server* s, double max_wait_s);
/* Capture every slot's graphs for every batch size (servers created with
 * use_graphs), so that no batch pays for a capture while serving. */
int xh_server_warm(xh_server* s);
/* Device footprint (bytes) the admission control charges for a request. */
double xh_server_footprint(xh_server* s, const xh_request* req);
void xh_server_destroy(xh_server* s);

/* Copy an intermediate array of request `req` to host memory for checks.
 * what: 0 flux[member], 1 folded signal[region] [N_ch, N_P],
 *       2 member sizes (S, status) as doubles, 3 cellArea[member] [S,S],
 *       4 deflection[member] [S,N_R], 5 lag[member] [S,N_R],
 *       6 theta[member] [S], 7 phi[member] [S], 8 STAR [N_ch,N_b]. */
int xh_pipeline_fetch(xh_pipeline* p, int what, int req, int index,
                      double* host_out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
