"""MI355X-native X-PSI pulse-profile likelihood.

Hand-written HIP kernels for gfx950 (libxpsi_hip.so, C ABI in
include/xpsi_hip.h) behind the reference package's call shapes:
`xpsi_amd.cellmesh.integrator_for_azimuthal_invariance.integrate`,
`xpsi_amd.tools.energy_integrator`, `xpsi_amd.likelihoods.
default_background_marginalisation.{precomputation, eval_marginal_likelihood}`
and the batched whole-likelihood `xpsi_amd.pipeline.LikelihoodPipeline`.
"""
from ._lib import lib, XHError  # noqa: F401

__version__ = "0.1.0"
