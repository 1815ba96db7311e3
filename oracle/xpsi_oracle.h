/*
 * CPU oracle for the X-PSI pulse-profile likelihood path.
 * TEST INFRASTRUCTURE ONLY: only tests/, the smoke check and the benchmark's
 * CPU-baseline leg load this library, and only as the checker.
 */
#ifndef XPSI_ORACLE_H
#define XPSI_ORACLE_H
#include <stddef.h>

#define OR_SUCCESS 0
#define OR_EFAILED 5
#define OR_EMAXITER 11
#define OR_EROUND 18
#define OR_ESING 21
#define OR_QAG_LIMIT 128

/* relative tolerance used at the reference's cquad call sites */
#define OR_MESH_EPSREL 1.0e-10
#define OR_MARG_EPSREL 1.0e-11

typedef double (*or_fn)(double, void *);

/* numerics core */
size_t or_interval(const double *xa, size_t n, double x);
void or_steffen_init(const double *x, const double *y, size_t n,
                     double *a, double *b, double *c, double *d);
double or_steffen_eval(const double *x, size_t n, const double *a,
                       const double *b, const double *c, const double *d,
                       double xq);
double or_steffen_deriv(const double *x, size_t n, const double *a,
                        const double *b, const double *c, double xq);
void or_akima_init(const double *x, const double *y, size_t n, int periodic,
                   double *b, double *c, double *d, double *mbuf);
double or_akima_eval(const double *x, const double *y, size_t n,
                     const double *b, const double *c, const double *d,
                     double xq);
void or_cspline_periodic_init(const double *x, const double *y, size_t n, double *b, double *c, double *d,
                              double *work);
/* the process-wide phase interpolant (xpsi/tools/core.pyx:21-116): 0 periodic
 * Akima, 1 Steffen, 2 periodic cubic spline; or_phase_init fills the shared
 * y_i + t (b + t (c + t d)) form (work: 3 n + 3 doubles) */
int or_set_phase_interpolant(int kind);
int or_get_phase_interpolant(void);
void or_phase_init(const double *x, const double *y, size_t n, double *b, double *c, double *d, double *work);
double or_akima_integ(const double *x, const double *y, size_t n,
                      const double *b, const double *c, const double *d,
                      double lo, double hi);
void or_gk61(or_fn f, void *p, double a, double b, double *res,
             double *err, double *rabs, double *rasc);
int or_qag61(or_fn f, void *p, double a, double b, double epsabs,
             double epsrel, int limit, double *result, double *abserr);

/* threads of the OpenMP baseline build (returns the count in use; 1 in the
   checker build) */
int or_set_threads(int n);

/* cell mesh (mesh_tools.pyx, mesh.pyx, polar_mesh.pyx) */
double or_effective_gravity(double mu, double R_eq, double zeta,
                            double epsilon, int shape);
void or_allocate_cells(size_t num_cells, size_t min_sqrt, size_t max_sqrt,
                       double mass, double r_s, double R_eq, double Omega,
                       double zeta, double epsilon, int shape,
                       double superRadius, double cedeRadius,
                       double superColatitude, double cedeColatitude,
                       double superAzimuth, double holeColatitude,
                       double holeRadius, double holeAzimuth, double *out4);
/* outputs: theta[S*S], phi[S*S], r[S], cellArea[S*S], maxAlpha[S],
   cos_gamma[S], effGrav[S] */
int or_construct_cell_mesh(int polar, size_t numCell, size_t S, double mass,
                           double r_s, double R_eq, double zeta,
                           double epsilon, int shape, double cedeRadius,
                           double cedeColatitude, double superRadius,
                           double superColatitude, double superAzimuth,
                           double *theta, double *phi, double *r,
                           double *cellArea, double *maxAlpha,
                           double *cos_gamma, double *effGrav);

/* rays.pyx compute_rays: maxAlpha is updated in place */
int or_compute_rays(size_t N_u, double r_s, const double *r_s_over_R,
                    double *maxAlpha, size_t N_R, double *deflection,
                    double *cos_alpha, double *lag, double *maxDeflection);

/* integrator_for_azimuthal_invariance.pyx integrate (one thread) */
typedef struct {
    size_t ndims;          /* 4 */
    const double *params[4];
    size_t N[4];
    const double *intensity;
} or_atmosphere;

/* hot_Num4D.pyx:248-438 over a query sequence from a fresh state (no norm) */
void or_num4d_sequence(const or_atmosphere *a, size_t n, const double *E,
                       const double *mu, const double *vec2, double *out);

int or_integrate(double R, double omega, double r_s, double inclination,
                 size_t S_rows, size_t S_cols, const double *cellArea,
                 const double *radial, const double *r_s_over_r,
                 const double *theta, const double *phi,
                 const double *srcCellParams, const int *radiates,
                 size_t N_R, const double *deflection,
                 const double *cos_alpha, const double *lag,
                 const double *maxDeflection, const double *cos_gamma,
                 size_t N_E, const double *energies, size_t N_L,
                 const double *leaves, size_t N_P, const double *phases,
                 int atm_ext, const or_atmosphere *atm, int image_order_limit,
                 const double *correction, double *flux);

/* xpsi/cellmesh/integratorIQU_for_azimuthal_invariance.pyx:68-746 with the
 * options shared by the plain invariant integrator: disk blocking (R_in),
 * atmospheres 1 and 3, beaming, the per-cell blackbody correction.  flux_Q
 * and flux_U may be NULL (intensity only).  Returns 0, 1 (the reference's
 * ERROR) or 2 (an option outside the restatement). */
int or_integrate_stokes(double omega, double inclination, size_t S_rows, size_t S_cols, const double *cellArea,
                        const double *radial, const double *r_s_over_r, const double *theta, const double *phi,
                        const double *cellParams, size_t n_params, const int *radiates, size_t N_R,
                        const double *deflection, const double *cos_alpha, const double *lag,
                        const double *maxDeflection, const double *cos_gammaArray, size_t N_E,
                        const double *energies, size_t N_L, const double *leaves, size_t N_P,
                        const double *phases, int atm_ext, int beam_opt, int image_order_limit, double R_in,
                        const double *correction, double *flux_I, double *flux_Q, double *flux_U);
double or_hot_I(int atm, int beam_opt, double E, double mu, const double *VEC);
double or_hot_Q(int atm, double E, double mu, const double *VEC);
double or_pol_angle(double leaf, double sin_psi, double cos_psi, double sin_alpha, double cos_alpha,
                    double sin_theta, double cos_theta, double sin_i, double cos_i, double sin_gamma,
                    double cos_gamma, double Grav_z, double mu, double eta, double beta, double Lorentz,
                    double cos_xi);
int or_disk_block(double R_in, double cos_i, double cos_psi, double cos_theta, double r_s_over_r, double radius,
                  double sin_alpha, double theta_over_pi);

/* xpsi/cellmesh/integrator.pyx:48-667 and integratorIQU.pyx:51-717: per-cell
 * parameters [S][S][n_params], atmospheres 1 and 3, beaming, the per-cell
 * blackbody correction; flux_Q and flux_U may be NULL.  Returns 0, 1 (ERROR)
 * or 2 (an option outside the restatement). */
int or_integrate_general_ext(double omega, double inclination, size_t S_rows, size_t S_cols,
                             const double *cellArea, const double *radial, const double *r_s_over_r,
                             const double *theta, const double *phi, const double *cellParams,
                             size_t n_params, const int *radiates, size_t N_R, const double *deflection,
                             const double *cos_alpha, const double *lag, const double *maxDeflection,
                             const double *cos_gammaArray, size_t N_E, const double *energies, size_t N_L,
                             const double *leaves, size_t N_P, const double *phases, int image_order_limit,
                             int atm_ext, int beam_opt, const double *correction, double *flux,
                             double *flux_Q, double *flux_U);
/* xpsi/cellmesh/integrator.pyx:48-667 (blackbody): per-cell parameters */
int or_integrate_general(double omega, double inclination, size_t S_rows, size_t S_cols,
                         const double *cellArea, const double *radial, const double *r_s_over_r,
                         const double *theta, const double *phi, const double *cellParams,
                         const int *radiates, size_t N_R, const double *deflection,
                         const double *cos_alpha, const double *lag, const double *maxDeflection,
                         const double *cos_gammaArray, size_t N_E, const double *energies, size_t N_L,
                         const double *leaves, size_t N_P, const double *phases, int image_order_limit,
                         const double *correction, double *flux);

/* global_mesh.pyx: closed mesh; colat/rr/maxAlpha/cos_gamma/effGrav [S],
   phis [S], *cellArea the common cell area */
int or_construct_closed_mesh(size_t S, double R_eq, double zeta, double eps,
                             int shape, double *colat, double *phis,
                             double *rr, double *cellArea, double *maxAlpha,
                             double *cos_gamma, double *effGrav);
/* integrator_for_time_invariance.pyx (blackbody): flux[N_E] */
int or_integrate_time_invariant(double omega, double r_s, double inclination,
                                size_t S, double cellArea, const double *radial,
                                const double *r_s_over_r, const double *colat,
                                const double *phis, double logT, size_t N_R,
                                const double *deflection, const double *cos_alpha,
                                const double *maxDeflection, const double *cos_gammaArray,
                                size_t N_E, const double *energies,
                                int image_order_limit, double *flux);

/* energy_integrator.pyx: out[N_in, N_P] (already transposed) */
void or_energy_integrator(size_t N_E, size_t N_P, const double *signal,
                          const double *log10_energies, size_t N_edges,
                          const double *log10_edges, double *out);

/* default_background_marginalisation.pyx */
void or_precomputation(size_t N_ch, size_t N_b, const int *counts,
                       double *precomp);
int or_eval_marginal_likelihood(double exposure_time, size_t N_b,
                                const double *phases, size_t N_ch,
                                const double *counts, size_t n_comp,
                                const double *const *components,
                                const size_t *comp_N_P,
                                const double *const *component_phases,
                                const double *phase_shifts,
                                const double *precomp, const double *support,
                                double epsabs, double epsrel, double epsilon,
                                double sigmas, double llzero,
                                const unsigned char *allow_negative,
                                double slim, const double *background,
                                double *lnl, double *STAR, double *B_ML,
                                double *B_ML_supp);
/* _poisson_likelihood_given_background.pyx: lnL and EXPEC [N_ch, N_b];
   background [N_ch, N_b] count rates; precomp [N_ch] or NULL */
int or_poisson_likelihood_given_background(double exposure_time, size_t N_b,
                                           const double *phases, size_t N_ch,
                                           const double *counts, size_t n_comp,
                                           const double *const *components,
                                           const size_t *comp_N_P,
                                           const double *const *component_phases,
                                           const double *phase_shifts,
                                           const double *background,
                                           const double *precomp,
                                           const unsigned char *allow_negative,
                                           double *lnl, double *EXPEC);
/* _gaussian_likelihood_given_background_IQU.pyx: lnL and EXPEC [N_ch, N_b];
   errors [N_ch, N_b] per channel and bin */
int or_gaussian_likelihood_given_background(double exposure_time, size_t N_b,
                                            const double *phases, size_t N_ch,
                                            const double *counts, const double *errors,
                                            size_t n_comp,
                                            const double *const *components,
                                            const size_t *comp_N_P,
                                            const double *const *component_phases,
                                            const double *phase_shifts,
                                            const double *background,
                                            const unsigned char *allow_negative,
                                            double *lnl, double *EXPEC);
/* _gaussian_likelihood_QnUn.pyx: lnL of counts/errors [N_rows, N_b] against
   one model row star [N_b] */
double or_gaussian_likelihood_QnUn(size_t N_rows, size_t N_b, const double *counts,
                                   const double *errors, const double *star);
/* phase_integrator.pyx: out [N_rows, N_b]; returns 1 for fewer than two phases */
int or_phase_integrator(double T_exp, size_t N_b, const double *phases, size_t N_rows,
                        const double *signal, size_t NP, const double *signal_phases,
                        double shift, int allow_neg, double *out);
#endif
