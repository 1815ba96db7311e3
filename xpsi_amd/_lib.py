"""ctypes binding to libxpsi_hip.so (declarations: include/xpsi_hip.h).

The library is required: there is no CPU fallback.  Loading fails loudly
when the shared object has not been built.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxpsi_hip.so")

_lib = None

d, i32, sz, vp = C.c_double, C.c_int, C.c_size_t, C.c_void_p


class RegionConfig(C.Structure):
    _fields_ = [("cede", i32), ("is_antiphased", i32), ("sqrt_num_cells", i32),
                ("min_sqrt_num_cells", i32), ("max_sqrt_num_cells", i32), ("omit", i32)]


class ModelConfig(C.Structure):
    _fields_ = [("n_regions", i32), ("region", RegionConfig * 2),
                ("num_rays", i32), ("num_leaves", i32), ("num_phases", i32),
                ("image_order_limit", i32), ("atm_ext", i32),
                ("atm_axes", vp * 4), ("atm_N", i32 * 4), ("atm_I", vp), ("star_shape", i32),
                ("frequency", d), ("mode_frequency", d),
                ("N_E", i32), ("energies", vp),
                ("N_in", i32), ("energy_edges", vp), ("attenuation", vp),
                ("N_ch", i32), ("matrix", vp),
                ("N_b", i32), ("data_phases", vp), ("counts", vp),
                ("exposure_time", d), ("support", vp),
                ("epsabs", d), ("epsrel", d), ("epsilon", d), ("sigmas", d),
                ("llzero", d), ("slim", d),
                ("max_batch", i32), ("max_sqrt_cells", i32),
                ("precision", i32), ("fold_kernel", i32),
                ("elsewhere", i32), ("else_sqrt_num_cells", i32), ("else_num_rays", i32),
                ("else_image_order_limit", i32), ("fold_terms", i32)]


class Request(C.Structure):
    _fields_ = [("mass", d), ("radius", d), ("distance", d), ("cos_inclination", d),
                ("region", (d * 8) * 2), ("sqrt_num_cells", i32 * 2),
                ("elsewhere_temperature", d), ("omit", (d * 3) * 2)]


class XHError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise XHError("libxpsi_hip.so is not built; run `python -m xpsi_amd.build` "
                      "(or __graft_entry__.build()) first")
    # PyTorch allocates the drop-ins' device buffers.  Load it first so that
    # this library's libamdhip64 dependency resolves to the HIP runtime torch
    # already holds: a second runtime instance in the same process finds no
    # device once torch's has opened it.
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    L.xh_last_error.restype = C.c_char_p
    L.xh_integrate_workspace_bytes.restype = sz
    L.xh_integrate_workspace_bytes.argtypes = [i32] * 5
    L.xh_integrate.restype = i32
    L.xh_integrate.argtypes = ([d, d, d, i32] + [vp] * 7 + [i32] + [vp] * 5 +
                               [i32, vp, i32, vp, i32, vp, i32, vp, vp, vp, i32, i32, d, vp, vp, sz,
                                C.POINTER(i32), vp])
    L.xh_integrate_stokes_workspace_bytes.restype = sz
    L.xh_integrate_stokes_workspace_bytes.argtypes = [i32, i32, i32]
    L.xh_integrate_stokes.restype = i32
    L.xh_integrate_stokes.argtypes = ([d, d, i32] + [vp] * 6 + [i32, vp, i32] + [vp] * 5 +
                                      [i32, vp, i32, vp, i32, vp, i32, i32, i32, d, vp, vp, vp, vp, vp, sz,
                                       C.POINTER(i32), vp])
    L.xh_integrate_general_workspace_bytes.restype = sz
    L.xh_integrate_general_workspace_bytes.argtypes = [i32, i32]
    L.xh_integrate_general.restype = i32
    L.xh_integrate_general.argtypes = ([d, d, i32] + [vp] * 6 + [i32, vp, i32] + [vp] * 5 +
                                       [i32, vp, i32, vp, i32, vp, i32, i32, i32, vp, vp, vp, vp, vp, sz,
                                        C.POINTER(i32), vp])
    L.xh_integrate_time_invariant_workspace_bytes.restype = sz
    L.xh_integrate_time_invariant_workspace_bytes.argtypes = [i32, i32]
    L.xh_integrate_time_invariant.restype = i32
    L.xh_integrate_time_invariant.argtypes = ([d, d, d, i32, d] + [vp] * 4 + [i32, i32] + [vp] * 4 +
                                              [i32, vp, i32, vp, vp, sz, C.POINTER(i32), vp])
    L.xh_energy_integrator.restype = i32
    L.xh_energy_integrator.argtypes = [i32, i32, vp, vp, i32, vp, vp, vp]
    L.xh_fold.restype = i32
    L.xh_fold.argtypes = [i32, i32, i32, vp, i32, vp, vp, vp]
    L.xh_precomputation.restype = i32
    L.xh_precomputation.argtypes = [i32, i32, vp, vp]
    L.xh_marginal_workspace_bytes.restype = sz
    L.xh_marginal_workspace_bytes.argtypes = [i32]
    L.xh_eval_marginal_likelihood.restype = i32
    L.xh_eval_marginal_likelihood.argtypes = (
        [d, i32, vp, i32, vp, i32, vp, i32, vp, vp, vp, vp, d, d, d, d, d, i32, d,
         vp, C.POINTER(d), vp, vp, vp, vp, sz, vp])
    L.xh_eval_marginal_likelihood_ex.restype = i32
    L.xh_eval_marginal_likelihood_ex.argtypes = (
        [d, i32, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, d, d, d, d, d, vp, d,
         vp, C.POINTER(d), vp, vp, vp, vp, sz, vp])
    L.xh_poisson_likelihood_given_background_ex.restype = i32
    L.xh_poisson_likelihood_given_background_ex.argtypes = (
        [d, i32, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp, C.POINTER(d), vp, vp, sz, vp])
    L.xh_gaussian_workspace_bytes.restype = sz
    L.xh_gaussian_workspace_bytes.argtypes = [i32, i32]
    L.xh_gaussian_likelihood_given_background.restype = i32
    L.xh_gaussian_likelihood_given_background.argtypes = (
        [d, i32, vp, i32, vp, vp, i32, vp, i32, vp, vp, vp, i32, C.POINTER(d), vp, vp, sz, vp])
    L.xh_gaussian_likelihood_given_background_ex.restype = i32
    L.xh_gaussian_likelihood_given_background_ex.argtypes = (
        [d, i32, vp, i32, vp, vp, i32, vp, vp, vp, vp, vp, vp, C.POINTER(d), vp, vp, sz, vp])
    L.xh_gaussian_likelihood_QnUn.restype = i32
    L.xh_gaussian_likelihood_QnUn.argtypes = [i32, i32, vp, vp, vp, C.POINTER(d), vp, sz, vp]
    L.xh_phase_integrator.restype = i32
    L.xh_phase_integrator.argtypes = [d, i32, vp, i32, vp, i32, vp, d, i32, vp, vp]
    L.xh_compute_rays_workspace_bytes.restype = sz
    L.xh_compute_rays_workspace_bytes.argtypes = [i32, i32]
    L.xh_compute_rays.restype = i32
    L.xh_compute_rays.argtypes = [i32, d, vp, vp, i32, vp, vp, vp, vp, vp, sz, C.POINTER(i32), vp]
    L.xh_allocate_cells_workspace_bytes.restype = sz
    L.xh_allocate_cells_workspace_bytes.argtypes = []
    L.xh_allocate_cells.restype = i32
    L.xh_allocate_cells.argtypes = [i32, i32, i32] + [d] * 6 + [i32] + [d] * 8 + [vp, vp, sz, vp]
    L.xh_cell_mesh_workspace_bytes.restype = sz
    L.xh_cell_mesh_workspace_bytes.argtypes = [i32]
    L.xh_construct_cell_mesh.restype = i32
    L.xh_construct_cell_mesh.argtypes = ([i32, d, i32] + [d] * 5 + [i32] + [d] * 5 + [vp] * 7 +
                                         [vp, sz, C.POINTER(i32), vp])
    L.xh_atm_init.restype = vp
    L.xh_atm_init.argtypes = [i32, vp, vp, vp]
    L.xh_atm_eval_I.restype = i32
    L.xh_atm_eval_I.argtypes = [vp, i32, vp, vp, vp, vp, vp]
    L.xh_atm_norm.restype = d
    L.xh_atm_norm.argtypes = [vp]
    L.xh_atm_free.restype = i32
    L.xh_atm_free.argtypes = [vp]
    L.xh_synthesise_workspace_bytes.restype = sz
    L.xh_synthesise_workspace_bytes.argtypes = [i32, i32]
    L.xh_synthesise_exposure.restype = i32
    L.xh_synthesise_exposure.argtypes = [d, i32, vp, i32, i32, vp, i32, vp, vp, d, vp, i32, C.c_ulonglong,
                                         vp, vp, C.POINTER(d), vp, sz, vp]
    L.xh_synthesise_given_total_count_number.restype = i32
    L.xh_synthesise_given_total_count_number.argtypes = [i32, vp, d, i32, i32, vp, i32, vp, vp, d, vp, i32,
                                                         C.c_ulonglong, vp, vp, vp, vp, sz, vp]
    L.xh_prior_inverse_sample.restype = i32
    L.xh_prior_inverse_sample.argtypes = [i32, i32, vp, vp, i32, i32, d, vp, vp, vp]
    L.xh_poisson_workspace_bytes.restype = sz
    L.xh_poisson_workspace_bytes.argtypes = [i32]
    L.xh_poisson_likelihood_given_background.restype = i32
    L.xh_poisson_likelihood_given_background.argtypes = (
        [d, i32, vp, i32, vp, i32, vp, i32, vp, vp, vp, vp, i32, C.POINTER(d), vp, vp, sz, vp])
    L.xh_eval_marginal_likelihood_noalloc.restype = i32
    L.xh_eval_marginal_likelihood_noalloc.argtypes = (
        [d, i32, vp, i32, vp, i32, vp, i32, vp, vp, vp, vp, d, d, d, d, d, i32, d,
         vp, C.POINTER(d), vp, vp, vp, vp])
    L.xh_pipeline_create.restype = vp
    L.xh_pipeline_create.argtypes = [C.POINTER(ModelConfig)]
    L.xh_pipeline_destroy.restype = None
    L.xh_pipeline_destroy.argtypes = [vp]
    L.xh_pipeline_eval.restype = i32
    L.xh_pipeline_eval.argtypes = [vp, i32, C.POINTER(Request), vp, vp]
    L.xh_pipeline_launch.restype = i32
    L.xh_pipeline_launch.argtypes = [vp, i32, C.POINTER(Request), vp]
    L.xh_pipeline_sync.restype = i32
    L.xh_pipeline_sync.argtypes = [vp]
    L.xh_pipeline_results.restype = C.POINTER(d)
    L.xh_pipeline_results.argtypes = [vp]
    L.xh_pipeline_set_timing.restype = i32
    L.xh_pipeline_set_timing.argtypes = [vp, i32]
    L.xh_pipeline_stage_ms.restype = i32
    L.xh_pipeline_stage_ms.argtypes = [vp, vp, i32]
    L.xh_pipeline_members_per_request.restype = i32
    L.xh_pipeline_members_per_request.argtypes = [vp]
    L.xh_pipeline_work.restype = i32
    L.xh_pipeline_work.argtypes = [vp, i32, vp]
    L.xh_set_fold_energy_chunk.restype = i32
    L.xh_set_fold_energy_chunk.argtypes = [i32]
    L.xh_set_fold_kernel.restype = i32
    L.xh_set_fold_kernel.argtypes = [i32]
    L.xh_set_moment_shape.restype = i32
    L.xh_set_moment_shape.argtypes = [i32]
    L.xh_set_fold_rings.restype = i32
    L.xh_set_fold_rings.argtypes = [i32]
    L.xh_set_wave_fold_rings.restype = i32
    L.xh_set_wave_fold_rings.argtypes = [i32]
    L.xh_set_wave_fold_chain.restype = i32
    L.xh_set_wave_fold_chain.argtypes = [i32]
    L.xh_set_wave_fold_shared.restype = i32
    L.xh_set_wave_fold_shared.argtypes = [i32]
    L.xh_set_phase_interpolant.restype = i32
    L.xh_set_phase_interpolant.argtypes = [i32]
    L.xh_get_phase_interpolant.restype = i32
    L.xh_get_phase_interpolant.argtypes = []
    L.xh_set_lane_phases.restype = i32
    L.xh_set_lane_phases.argtypes = [i32]
    L.xh_pipeline_use_graphs.restype = i32
    L.xh_pipeline_use_graphs.argtypes = [vp, i32]
    L.xh_server_create.restype = vp
    L.xh_server_create.argtypes = [C.POINTER(ModelConfig), i32, i32, d, i32]
    L.xh_server_submit.restype = i32
    L.xh_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_longlong)]
    L.xh_server_poll.restype = i32
    L.xh_server_poll.argtypes = [vp, vp, vp, vp, i32]
    L.xh_server_stats.restype = i32
    L.xh_server_stats.argtypes = [vp, vp, i32]
    L.xh_server_destroy.restype = None
    L.xh_server_destroy.argtypes = [vp]
    L.xh_server_warm.restype = i32
    L.xh_server_warm.argtypes = [vp]
    L.xh_pipeline_capture.restype = i32
    L.xh_pipeline_capture.argtypes = [vp, i32, vp]
    L.xh_server_set_policy.restype = i32
    L.xh_server_set_policy.argtypes = [vp, d]
    L.xh_server_footprint.restype = d
    L.xh_server_footprint.argtypes = [vp, C.POINTER(Request)]
    L.xh_pipeline_fetch.restype = i32
    L.xh_pipeline_fetch.argtypes = [vp, i32, i32, i32, vp, sz]
    _lib = L
    return L


def check(rc, what):
    if rc != 0:
        raise XHError("%s failed (%d): %s" % (what, rc, lib().xh_last_error().decode()))


EXPORTED = ["xh_last_error", "xh_integrate_workspace_bytes", "xh_integrate",
            "xh_integrate_general_workspace_bytes", "xh_integrate_general",
            "xh_integrate_stokes_workspace_bytes", "xh_integrate_stokes",
            "xh_integrate_time_invariant_workspace_bytes", "xh_integrate_time_invariant",
            "xh_energy_integrator", "xh_fold", "xh_precomputation",
            "xh_marginal_workspace_bytes", "xh_eval_marginal_likelihood", "xh_eval_marginal_likelihood_ex",
            "xh_eval_marginal_likelihood_noalloc",
            "xh_poisson_workspace_bytes", "xh_poisson_likelihood_given_background", "xh_gaussian_workspace_bytes",
            "xh_gaussian_likelihood_given_background", "xh_gaussian_likelihood_given_background_ex",
            "xh_gaussian_likelihood_QnUn", "xh_phase_integrator",
            "xh_poisson_likelihood_given_background_ex",
            "xh_compute_rays_workspace_bytes", "xh_compute_rays", "xh_allocate_cells_workspace_bytes",
            "xh_allocate_cells", "xh_cell_mesh_workspace_bytes", "xh_construct_cell_mesh",
            "xh_prior_inverse_sample", "xh_synthesise_workspace_bytes", "xh_synthesise_exposure",
            "xh_synthesise_given_total_count_number", "xh_atm_init", "xh_atm_eval_I", "xh_atm_norm",
            "xh_atm_free",
            "xh_pipeline_create", "xh_pipeline_destroy", "xh_pipeline_eval",
            "xh_pipeline_launch", "xh_pipeline_sync", "xh_pipeline_results",
            "xh_pipeline_fetch", "xh_pipeline_set_timing", "xh_pipeline_stage_ms",
            "xh_pipeline_members_per_request", "xh_pipeline_work",
            "xh_set_fold_energy_chunk", "xh_set_fold_kernel", "xh_set_moment_shape", "xh_set_fold_rings", "xh_set_wave_fold_rings", "xh_set_wave_fold_chain", "xh_set_wave_fold_shared", "xh_set_phase_interpolant", "xh_get_phase_interpolant", "xh_set_lane_phases", "xh_pipeline_use_graphs", "xh_server_create",
            "xh_server_submit", "xh_server_poll", "xh_server_stats", "xh_server_destroy",
            "xh_server_set_policy", "xh_server_footprint", "xh_server_warm", "xh_pipeline_capture"]
