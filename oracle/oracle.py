"""CPU oracle: ctypes bindings to oracle/_build/liboracle.so and a
restatement of the reference model layer's orchestration.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, always as the checker.

The orchestration follows
  xpsi/Spacetime.py:111-190 (derived quantities),
  xpsi/HotRegion.py:774-1017 (mesh, rays, lag calibration, cell parameters),
  xpsi/HotRegion.py:1115-1241 (integration per super/cede member),
  xpsi/Likelihood.py:297-432 (divide by d^2, register, signal call),
  xpsi/Signal.py:379-453, 1092-1197 (energy grid, energy integral, fold),
  examples/examples_fast/Modules/CustomSignal.py:52-67 (marginal likelihood).
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
# "checker": one thread, no FMA contraction (the parity reference);
# "fast": the same sources built with the reference's flags and OpenMP (the
# benchmark's CPU baseline)
_BUILD = "checker"
_FILES = {"checker": "liboracle.so", "fast": "liboracle_fast.so"}

# xpsi/global_imports.py:67-88
c_light = 2.99792458e8
kpc = 3.08567758e19
keV = 1.60217662e-16
GM_sun = 0.5 * 2.95325024e3
G_SI = 6.6740831e-11
km = 1.0e3

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def select_build(name):
    """Use the "checker" or the "fast" (OpenMP, -O3 -march=native) build;
    takes effect for calls made afterwards."""
    global _LIB, _BUILD
    if name not in _FILES:
        raise ValueError(name)
    if name != _BUILD:
        _BUILD, _LIB = name, None


def set_threads(n):
    """OpenMP threads of the fast build (returns the count now in use)."""
    return lib().or_set_threads(int(n))


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", _FILES[_BUILD])
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        d, sz, i = C.c_double, C.c_size_t, C.c_int
        L.or_set_threads.restype = i
        L.or_set_threads.argtypes = [i]
        L.or_interval.restype = sz
        L.or_interval.argtypes = [_dp, sz, d]
        L.or_steffen_init.argtypes = [_dp, _dp, sz, _dp, _dp, _dp, _dp]
        L.or_steffen_eval.restype = d
        L.or_steffen_eval.argtypes = [_dp, sz, _dp, _dp, _dp, _dp, d]
        L.or_steffen_deriv.restype = d
        L.or_steffen_deriv.argtypes = [_dp, sz, _dp, _dp, _dp, d]
        L.or_akima_init.argtypes = [_dp, _dp, sz, i, _dp, _dp, _dp, _dp]
        L.or_akima_eval.restype = d
        L.or_akima_eval.argtypes = [_dp, _dp, sz, _dp, _dp, _dp, d]
        L.or_cspline_periodic_init.argtypes = [_dp, _dp, sz, _dp, _dp, _dp, _dp]
        L.or_set_phase_interpolant.restype = i
        L.or_set_phase_interpolant.argtypes = [i]
        L.or_get_phase_interpolant.restype = i
        L.or_phase_init.argtypes = [_dp, _dp, sz, _dp, _dp, _dp, _dp]
        L.or_akima_integ.restype = d
        L.or_akima_integ.argtypes = [_dp, _dp, sz, _dp, _dp, _dp, d, d]
        L.or_effective_gravity.restype = d
        L.or_effective_gravity.argtypes = [d, d, d, d, i]
        L.or_allocate_cells.argtypes = [sz, sz, sz] + [d] * 6 + [i] + [d] * 8 + [_dp]
        L.or_construct_cell_mesh.restype = i
        L.or_construct_cell_mesh.argtypes = ([i, sz, sz] + [d] * 5 + [i] + [d] * 5 + [_dp] * 7)
        L.or_compute_rays.restype = i
        L.or_compute_rays.argtypes = [sz, d, _dp, _dp, sz, _dp, _dp, _dp, _dp]
        L.or_integrate.restype = i
        L.or_integrate.argtypes = ([d] * 4 + [sz, sz] + [_dp] * 6 + [_ip, sz] +
                                   [_dp] * 5 + [sz, _dp, sz, _dp, sz, _dp, i,
                                                C.c_void_p, i, C.c_void_p, _dp])
        L.or_integrate_general.restype = i
        L.or_integrate_general.argtypes = ([d, d, sz, sz] + [_dp] * 6 + [_ip, sz] +
                                           [_dp] * 5 + [sz, _dp, sz, _dp, sz, _dp, i,
                                                        C.c_void_p, _dp])
        L.or_integrate_general_ext.restype = i
        L.or_integrate_general_ext.argtypes = ([d, d, sz, sz] + [_dp] * 6 + [sz, _ip, sz] + [_dp] * 5 +
                                               [sz, _dp, sz, _dp, sz, _dp, i, i, i, C.c_void_p, _dp,
                                                C.c_void_p, C.c_void_p])
        L.or_integrate_stokes.restype = i
        L.or_integrate_stokes.argtypes = ([d, d, sz, sz] + [_dp] * 6 + [sz, _ip, sz] + [_dp] * 5 +
                                          [sz, _dp, sz, _dp, sz, _dp, i, i, i, d, C.c_void_p,
                                           _dp, C.c_void_p, C.c_void_p])
        L.or_hot_I.restype = d
        L.or_hot_I.argtypes = [i, i, d, d, _dp]
        L.or_pol_angle.restype = d
        L.or_pol_angle.argtypes = [d] * 17
        L.or_disk_block.restype = i
        L.or_disk_block.argtypes = [d] * 8
        L.or_num4d_sequence.argtypes = [C.c_void_p, sz, _dp, _dp, _dp, _dp]
        L.or_construct_closed_mesh.restype = i
        L.or_construct_closed_mesh.argtypes = ([sz, d, d, d, i, _dp, _dp, _dp,
                                                C.POINTER(C.c_double), _dp, _dp, _dp])
        L.or_integrate_time_invariant.restype = i
        L.or_integrate_time_invariant.argtypes = ([d, d, d, sz, d] + [_dp] * 4 +
                                                  [d, sz] + [_dp] * 4 + [sz, _dp, i, _dp])
        L.or_energy_integrator.argtypes = [sz, sz, _dp, _dp, sz, _dp, _dp]
        L.or_precomputation.argtypes = [sz, sz, _ip, _dp]
        L.or_eval_marginal_likelihood.restype = i
        L.or_eval_marginal_likelihood.argtypes = (
            [d, sz, _dp, sz, _dp, sz, C.c_void_p, C.c_void_p, C.c_void_p, _dp,
             _dp, _dp, d, d, d, d, d, C.c_void_p, d, C.c_void_p,
             C.POINTER(C.c_double), _dp, _dp, _dp])
        L.or_poisson_likelihood_given_background.restype = i
        L.or_poisson_likelihood_given_background.argtypes = (
            [d, sz, _dp, sz, _dp, sz, C.c_void_p, C.c_void_p, C.c_void_p, _dp, _dp,
             C.c_void_p, C.c_void_p, C.POINTER(C.c_double), _dp])
        L.or_gaussian_likelihood_given_background.restype = i
        L.or_gaussian_likelihood_given_background.argtypes = (
            [d, sz, _dp, sz, _dp, _dp, sz, C.c_void_p, C.c_void_p, C.c_void_p, _dp, _dp,
             C.c_void_p, C.POINTER(C.c_double), _dp])
        L.or_gaussian_likelihood_QnUn.restype = d
        L.or_gaussian_likelihood_QnUn.argtypes = [sz, sz, _dp, _dp, _dp]
        L.or_phase_integrator.restype = i
        L.or_phase_integrator.argtypes = [d, sz, _dp, sz, _dp, sz, _dp, d, i, _dp]
        _LIB = L
    return _LIB


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ----------------------------------------------------------------- numerics
def steffen(x, y):
    x, y = _c(x), _c(y)
    n = len(x)
    co = [np.zeros(n) for _ in range(4)]
    lib().or_steffen_init(x, y, n, *co)
    return x, co


def steffen_eval(sp, xq, deriv=False):
    x, co = sp
    if deriv:
        return lib().or_steffen_deriv(x, len(x), co[0], co[1], co[2], float(xq))
    return lib().or_steffen_eval(x, len(x), *co, float(xq))


def akima(x, y, periodic=True):
    x, y = _c(x), _c(y)
    n = len(x)
    co = [np.zeros(n) for _ in range(3)]
    lib().or_akima_init(x, y, n, int(periodic), *co, np.zeros(n + 3))
    return x, y, co


PHASE_INTERPOLANTS = {"Akima": 0, "Steffen": 1, "Cubic": 2}


def set_phase_interpolant(name):
    """xpsi.tools.core.set_phase_interpolant: "Akima" (periodic, default),
    "Steffen" or "Cubic" (periodic cubic spline), for every later call."""
    if lib().or_set_phase_interpolant(PHASE_INTERPOLANTS[name]) != 0:
        raise ValueError(name)


def get_phase_interpolant():
    k = lib().or_get_phase_interpolant()
    return [n for n, v in PHASE_INTERPOLANTS.items() if v == k][0]


def phase_spline(x, y):
    """Coefficients (x, y, [b, c, d]) of the current phase interpolant, for
    akima_eval and akima_integ (the shared piece form)."""
    x, y = _c(x), _c(y)
    n = len(x)
    co = [np.zeros(n) for _ in range(3)]
    lib().or_phase_init(x, y, n, *co, np.zeros(3 * n + 3))
    return x, y, co


def akima_eval(sp, xq):
    x, y, co = sp
    return lib().or_akima_eval(x, y, len(x), *co, float(xq))


def akima_integ(sp, a, b):
    x, y, co = sp
    return lib().or_akima_integ(x, y, len(x), *co, float(a), float(b))


# ------------------------------------------------------------ cell mesh/rays
def allocate_cells(num_cells, min_sqrt, max_sqrt, mass, r_s, R_eq, Omega,
                   zeta, epsilon, shape, superRadius, cedeRadius,
                   superColatitude, cedeColatitude, superAzimuth,
                   holeColatitude, holeRadius, holeAzimuth):
    out = np.zeros(4)
    lib().or_allocate_cells(int(num_cells), int(min_sqrt), int(max_sqrt),
                            mass, r_s, R_eq, Omega, zeta, epsilon, int(shape),
                            superRadius, cedeRadius, superColatitude,
                            cedeColatitude, superAzimuth, holeColatitude,
                            holeRadius, holeAzimuth, out)
    return tuple(out)


def construct_cell_mesh(polar, numCell, S, mass, r_s, R_eq, zeta, epsilon,
                        shape, cedeRadius, cedeColatitude, superRadius,
                        superColatitude, superAzimuth):
    S = int(S)
    theta, phi, area = (np.zeros((S, S)) for _ in range(3))
    r, maxA, cg, eg = (np.zeros(S) for _ in range(4))
    lib().or_construct_cell_mesh(int(polar), int(numCell), S, mass, r_s, R_eq,
                                 zeta, epsilon, int(shape), cedeRadius,
                                 cedeColatitude, superRadius, superColatitude,
                                 superAzimuth, theta, phi, r, area, maxA, cg, eg)
    return theta, phi, r, area, maxA, cg, eg


def compute_rays(r_s, r_s_over_r, maxAlpha, N_R):
    u = _c(r_s_over_r)
    N_u = len(u)
    defl, cosa, lag = (np.zeros((N_u, N_R)) for _ in range(3))
    maxD = np.zeros(N_u)
    t = lib().or_compute_rays(N_u, r_s, u, maxAlpha, N_R, defl, cosa, lag, maxD)
    return t, defl, cosa, lag, maxD


class _Atm(C.Structure):
    _fields_ = [("ndims", C.c_size_t), ("params", C.c_void_p * 4),
                ("N", C.c_size_t * 4), ("intensity", C.c_void_p)]


def integrate(omega, r_s, inclination, cellArea, radial, r_s_over_r, theta,
              phi, cellParams, radiates, deflection, cos_alpha, lag,
              maxDeflection, cos_gamma, energies, leaves, phases, atm_ext=1,
              atmosphere=None, image_order_limit=None, R=0.0, correction=None):
    """integrator_for_azimuthal_invariance.integrate; `correction` is the
    [S, S, 2] elsewhere parameter array (blackbody) or None."""
    S0, S1 = cellArea.shape
    N_R = deflection.shape[1]
    flux = np.zeros((len(energies), len(phases)))
    keep = []
    atm_ptr = None
    if atm_ext == 2:
        atm_ptr = _atm_struct(atmosphere, keep)
    t = lib().or_integrate(R, omega, r_s, inclination, S0, S1, _c(cellArea),
                           _c(radial), _c(r_s_over_r), _c(theta), _c(phi),
                           _c(cellParams), _c(radiates, np.int32), N_R,
                           _c(deflection), _c(cos_alpha), _c(lag),
                           _c(maxDeflection), _c(cos_gamma), len(energies),
                           _c(energies), len(leaves), _c(leaves), len(phases),
                           _c(phases), int(atm_ext), atm_ptr,
                           int(image_order_limit or 0),
                           None if correction is None else correction.ctypes.data, flux)
    return t, flux


def integrate_general(omega, inclination, cellArea, radial, r_s_over_r, theta, phi,
                      cellParams, radiates, deflection, cos_alpha, lag, maxDeflection,
                      cos_gamma, energies, leaves, phases, image_order_limit=None,
                      correction=None):
    """cellmesh/integrator.integrate for the blackbody: cellParams [S, S, 2]
    per cell (log10 T, ...), `correction` the [S, S, 2] elsewhere parameter
    array or None.  Returns (status, flux [N_E, N_P])."""
    S0, S1 = cellArea.shape
    N_R = deflection.shape[1]
    flux = np.zeros((len(energies), len(phases)))
    corr = None if correction is None else _c(correction)
    t = lib().or_integrate_general(omega, inclination, S0, S1, _c(cellArea), _c(radial),
                                   _c(r_s_over_r), _c(theta), _c(phi), _c(cellParams),
                                   _c(radiates, np.int32), N_R, _c(deflection), _c(cos_alpha),
                                   _c(lag), _c(maxDeflection), _c(cos_gamma), len(energies),
                                   _c(energies), len(leaves), _c(leaves), len(phases),
                                   _c(phases), int(image_order_limit or 0),
                                   None if corr is None else corr.ctypes.data, flux)
    return t, flux


def integrate_general_stokes(omega, inclination, cellArea, radial, r_s_over_r, theta, phi, cellParams,
                             radiates, deflection, cos_alpha, lag, maxDeflection, cos_gamma, energies,
                             leaves, phases, atm_ext=1, beam_opt=0, image_order_limit=None,
                             correction=None, stokes=True):
    """integratorIQU.integrate (stokes=True: (status, I, Q, U)) or
    integrator.integrate with the atmosphere and beaming options (stokes=False:
    (status, I)); cellParams [S, S, n], correction [S, S, 2] or None."""
    S0, S1 = cellArea.shape
    N_R = deflection.shape[1]
    P = _c(cellParams)
    fI = np.zeros((len(energies), len(phases)))
    fQ = np.zeros_like(fI) if stokes else None
    fU = np.zeros_like(fI) if stokes else None
    corr = None if correction is None else _c(correction)
    t = lib().or_integrate_general_ext(omega, inclination, S0, S1, _c(cellArea), _c(radial), _c(r_s_over_r),
                                       _c(theta), _c(phi), P, P.shape[2], _c(radiates, np.int32), N_R,
                                       _c(deflection), _c(cos_alpha), _c(lag), _c(maxDeflection),
                                       _c(cos_gamma), len(energies), _c(energies), len(leaves), _c(leaves),
                                       len(phases), _c(phases), int(image_order_limit or 0), int(atm_ext),
                                       int(beam_opt), None if corr is None else corr.ctypes.data, fI,
                                       None if fQ is None else fQ.ctypes.data,
                                       None if fU is None else fU.ctypes.data)
    if t == 2:
        raise ValueError("option outside the oracle's restatement")
    return (t, fI, fQ, fU) if stokes else (t, fI)


def integrate_stokes(omega, inclination, cellArea, radial, r_s_over_r, theta, phi, cellParams,
                     radiates, deflection, cos_alpha, lag, maxDeflection, cos_gamma, energies,
                     leaves, phases, atm_ext=1, beam_opt=0, image_order_limit=None, R_in=1e6,
                     correction=None, stokes=True):
    """integratorIQU_for_azimuthal_invariance.integrate (stokes=True: returns
    (status, I, Q, U)) or the plain invariant integrator under the same options
    (stokes=False: (status, I)).  cellParams [S, S, n] (n = 2, or 7 with
    beaming: log10 T, log g, a, b, c, d, n_mu); correction [S, S, 2] or None."""
    S0, S1 = cellArea.shape
    N_R = deflection.shape[1]
    P = _c(cellParams)
    fI = np.zeros((len(energies), len(phases)))
    fQ = np.zeros_like(fI) if stokes else None
    fU = np.zeros_like(fI) if stokes else None
    corr = None if correction is None else _c(correction)
    t = lib().or_integrate_stokes(omega, inclination, S0, S1, _c(cellArea), _c(radial), _c(r_s_over_r),
                                  _c(theta), _c(phi), P, P.shape[2], _c(radiates, np.int32), N_R,
                                  _c(deflection), _c(cos_alpha), _c(lag), _c(maxDeflection), _c(cos_gamma),
                                  len(energies), _c(energies), len(leaves), _c(leaves), len(phases),
                                  _c(phases), int(atm_ext), int(beam_opt), int(image_order_limit or 0),
                                  float(R_in), None if corr is None else corr.ctypes.data, fI,
                                  None if fQ is None else fQ.ctypes.data,
                                  None if fU is None else fU.ctypes.data)
    if t == 2:
        raise ValueError("option outside the oracle's restatement")
    return (t, fI, fQ, fU) if stokes else (t, fI)


def hot_I(atm, beam_opt, E, mu, VEC):
    return lib().or_hot_I(int(atm), int(beam_opt), float(E), float(mu), _c(VEC))


def pol_angle(*args):
    return lib().or_pol_angle(*[float(a) for a in args])


def disk_block(*args):
    return lib().or_disk_block(*[float(a) for a in args])


def _atm_struct(atmosphere, keep):
    arrs = [_c(a) for a in atmosphere]
    keep.append(arrs)
    st = _Atm()
    st.ndims = 4
    for k in range(4):
        st.params[k] = arrs[k].ctypes.data
        st.N[k] = len(arrs[k])
    st.intensity = arrs[4].ctypes.data
    keep.append(st)
    return C.addressof(st)


def num4d_sequence(atmosphere, E, mu, logT, logg):
    """hot_Num4D.pyx:248-438 evaluated for the queries (E[q], mu[q], logT[q],
    logg[q]) in order, from the state an integrate() call starts with;
    returns the intensities without the norm."""
    keep = []
    ptr = _atm_struct(atmosphere, keep)
    E, mu = _c(E), _c(mu)
    vec = _c(np.stack([np.broadcast_to(logT, E.shape), np.broadcast_to(logg, E.shape)], axis=1))
    out = np.zeros(len(E))
    lib().or_num4d_sequence(ptr, len(E), E, mu, vec, out)
    return out


def closed_mesh(st, S):
    """global_mesh.pyx:21-147 for the star `st` with S rings of S cells."""
    S = int(S)
    colat, phis, r, maxA, cg, eg = (np.zeros(S) for _ in range(6))
    ca = C.c_double(0.0)
    t = lib().or_construct_closed_mesh(S, st["R"], st["zeta"], st["epsilon"], st["shape"],
                                       colat, phis, r, C.byref(ca), maxA, cg, eg)
    assert t == 0
    return dict(S=S, colat=colat, phis=phis, r=r, cellArea=ca.value, maxAlpha=maxA,
                cos_gamma=cg, effGrav=eg)


def elsewhere_embed(st, es):
    """Elsewhere.embed (xpsi/Elsewhere.py:254-384): closed mesh and rays.
    `es`: sqrt_num_cells (default 64), num_rays (default 1000)."""
    m = closed_mesh(st, es.get("sqrt_num_cells", 64))
    u = np.ascontiguousarray(st["r_s"] / m["r"])
    t, defl, cosa, lag, maxD = compute_rays(st["r_s"], u, m["maxAlpha"],
                                            int(es.get("num_rays", 1000)))
    if t == 1:
        return None, "RayError"
    m.update(r_s_over_r=u, deflection=defl, cos_alpha=cosa, maxDeflection=maxD)
    return m, None


def elsewhere_integrate(st, m, energies, logT, image_order_limit=None):
    """Elsewhere.integrate (xpsi/Elsewhere.py:386-441): spectrum[N_E]."""
    out = np.zeros(len(energies))
    t = lib().or_integrate_time_invariant(st["Omega"], st["r_s"], st["i"], m["S"],
                                          m["cellArea"], m["r"], m["r_s_over_r"],
                                          m["colat"], m["phis"], float(logT),
                                          m["deflection"].shape[1], m["deflection"],
                                          m["cos_alpha"], m["maxDeflection"],
                                          m["cos_gamma"], len(energies), _c(energies),
                                          int(image_order_limit or 0), out)
    return t, out


# ------------------------------------------------------------ signal / lnL
def energy_integrator(signal, log10_energies, log10_edges):
    signal = _c(signal)
    N_E, N_P = signal.shape
    out = np.zeros((len(log10_edges) - 1, N_P))
    lib().or_energy_integrator(N_E, N_P, signal, _c(log10_energies),
                               len(log10_edges), _c(log10_edges), out)
    return out


def precomputation(counts):
    counts = _c(counts, np.int32)
    out = np.zeros(counts.shape[0])
    lib().or_precomputation(counts.shape[0], counts.shape[1], counts, out)
    return out


def eval_marginal_likelihood(exposure_time, phases, counts, components,
                             component_phases, phase_shifts, precomp, support,
                             epsabs=0.0, epsrel=1.0e-8, epsilon=1.0e-3,
                             sigmas=10.0, llzero=-1.0e90, allow_negative=False,
                             slim=20.0, background=None):
    comps = [_c(cmp) for cmp in components]
    cph = [_c(p) for p in component_phases]
    nc = len(comps)
    N_ch, N_b = counts.shape
    cp = (C.c_void_p * nc)(*[cmp.ctypes.data for cmp in comps])
    pp = (C.c_void_p * nc)(*[p.ctypes.data for p in cph])
    nps = (C.c_size_t * nc)(*[cmp.shape[1] for cmp in comps])
    if isinstance(allow_negative, bool):
        an = np.full(nc, int(allow_negative), dtype=np.uint8)
    else:
        an = np.asarray(allow_negative, dtype=np.uint8)
    bg = None if background is None else _c(background)
    STAR = np.zeros((N_ch, N_b))
    bml, bmls = np.zeros(N_ch), np.zeros(N_ch)
    lnl = C.c_double(0.0)
    lib().or_eval_marginal_likelihood(
        float(exposure_time), N_b, _c(phases), N_ch, _c(counts), nc, cp, nps,
        pp, _c(phase_shifts), _c(precomp), _c(support), epsabs, epsrel,
        epsilon, sigmas, llzero, an.ctypes.data, slim,
        None if bg is None else bg.ctypes.data, C.byref(lnl), STAR, bml, bmls)
    return lnl.value, STAR, bml, bmls


def poisson_likelihood_given_background(exposure_time, phases, counts, components,
                                        component_phases, phase_shifts, background,
                                        neg_sum_ln_data_factorial=None, allow_negative=False):
    """_poisson_likelihood_given_background.pyx:14-113: (lnL, EXPEC)."""
    comps = [_c(cmp) for cmp in components]
    cph = [_c(p) for p in component_phases]
    nc = len(comps)
    N_ch, N_b = counts.shape
    cp = (C.c_void_p * nc)(*[cmp.ctypes.data for cmp in comps])
    pp = (C.c_void_p * nc)(*[p.ctypes.data for p in cph])
    nps = (C.c_size_t * nc)(*[cmp.shape[1] for cmp in comps])
    if isinstance(allow_negative, bool):
        an = np.full(nc, int(allow_negative), dtype=np.uint8)
    else:
        an = np.asarray(allow_negative, dtype=np.uint8)
    pre = None if neg_sum_ln_data_factorial is None else _c(neg_sum_ln_data_factorial)
    EXPEC = np.zeros((N_ch, N_b))
    lnl = C.c_double(0.0)
    lib().or_poisson_likelihood_given_background(
        float(exposure_time), N_b, _c(phases), N_ch, _c(counts), nc, cp, nps, pp,
        _c(phase_shifts), _c(background), None if pre is None else pre.ctypes.data,
        an.ctypes.data, C.byref(lnl), EXPEC)
    return lnl.value, EXPEC


def gaussian_likelihood_given_background(exposure_time, phases, counts, errors, components,
                                         component_phases, phase_shifts, background, allow_negative=False):
    """_gaussian_likelihood_given_background_IQU.pyx:14-91: (lnL, EXPEC)."""
    comps = [_c(cmp) for cmp in components]
    cph = [_c(p) for p in component_phases]
    nc = len(comps)
    N_ch, N_b = counts.shape
    cp = (C.c_void_p * nc)(*[cmp.ctypes.data for cmp in comps])
    pp = (C.c_void_p * nc)(*[p.ctypes.data for p in cph])
    nps = (C.c_size_t * nc)(*[cmp.shape[1] for cmp in comps])
    if isinstance(allow_negative, bool):
        an = np.full(nc, int(allow_negative), dtype=np.uint8)
    else:
        an = np.asarray(allow_negative, dtype=np.uint8)
    EXPEC = np.zeros((N_ch, N_b))
    lnl = C.c_double(0.0)
    lib().or_gaussian_likelihood_given_background(
        float(exposure_time), N_b, _c(phases), N_ch, _c(counts), _c(errors), nc, cp, nps, pp,
        _c(phase_shifts), _c(background), an.ctypes.data, C.byref(lnl), EXPEC)
    return lnl.value, EXPEC


def gaussian_likelihood_QnUn(phases, counts, errors, star):
    """_gaussian_likelihood_QnUn.pyx:14-44: (lnL, star)."""
    counts = _c(counts)
    N_rows, N_b = counts.shape
    star = _c(star)
    return float(lib().or_gaussian_likelihood_QnUn(N_rows, N_b, counts, _c(errors), star)), star


def phase_integrator(exposure_time, phases, signal, signal_phases, phase_shift, allow_negative=0):
    """phase_integrator.pyx:23-121: [rows, intervals]."""
    signal = _c(signal)
    phases = _c(phases)
    out = np.zeros((signal.shape[0], len(phases) - 1))
    rc = lib().or_phase_integrator(float(exposure_time), len(phases) - 1, phases, signal.shape[0], signal,
                                   signal.shape[1], _c(signal_phases), float(phase_shift), int(bool(allow_negative)),
                                   out)
    if rc:
        raise ValueError("phase_integrator needs two or more signal phases")
    return out


# ---------------------------------------------------------- model pipeline
def spacetime(mass, radius, distance, cos_inclination, frequency,
              star_shape="AGM_14"):
    """Derived quantities, xpsi/Spacetime.py:111-190."""
    r_g = mass * GM_sun
    M = mass * GM_sun * c_light ** 2 / G_SI
    R = radius * km
    Omega = 2.0 * math.pi * frequency
    return dict(mass=mass, M=M, r_g=r_g, r_s=2.0 * r_g, R=R, Omega=Omega,
                f=frequency, i=math.acos(cos_inclination),
                d=distance * kpc, d_sq=(distance * kpc) ** 2, zeta=r_g / R,
                epsilon=Omega ** 2 * R ** 3 / (G_SI * M),
                shape={"AGM_14": 0, "sphere": 1}[star_shape])


def construct_energy_array(num_energies, edges):
    """Signal.py:1092-1197 for one signal (no coverage gaps)."""
    L, U = edges[0], edges[-1]
    e = np.logspace(np.log10(L), np.log10(U), int(math.ceil(num_energies)),
                    base=10.0)[::-1]
    return np.ascontiguousarray(e[::-1])


def hot_region_embed(st, hr, mode_frequency):
    """HotRegion.py:774-1017: mesh, rays, calibrated lag, cell parameters.

    `hr` holds the region parameters and resolution settings:
    super_colatitude, super_radius, super_temperature, phase_shift, and
    optionally cede_colatitude/cede_radius/cede_azimuth/cede_temperature,
    omit_*; sqrt_num_cells, min/max_sqrt_num_cells, num_rays, num_leaves,
    num_phases, is_antiphased, image_order_limit, atm_ext.
    Returns (members, error) with members a list of dicts.
    """
    sc, sr = hr["super_colatitude"], hr["super_radius"]
    cede = hr.get("cede_radius", 0.0) > 0.0 or hr.get("cede", False)
    cr = hr.get("cede_radius", 0.0) if cede else 0.0
    cc = hr.get("cede_colatitude", sc) if cede else sc
    ca = hr.get("cede_azimuth", 0.0) if cede else 0.0
    oc = hr.get("omit_colatitude", sc)
    orad = hr.get("omit_radius", 0.0)
    oa = hr.get("omit_azimuth", 0.0)
    nc = int(hr.get("sqrt_num_cells", 32)) ** 2
    alloc = allocate_cells(nc, hr.get("min_sqrt_num_cells", 10),
                           hr.get("max_sqrt_num_cells", 80), st["M"],
                           st["r_s"], st["R"], st["Omega"], st["zeta"],
                           st["epsilon"], st["shape"], sr, cr, sc, cc, -ca,
                           oc, orad, oa)
    members = []
    specs = [(alloc[0], alloc[1], sr, sc, orad, oc, oa, -oa,
              hr["super_temperature"])]
    if cr > 0.0:
        specs.append((alloc[2], alloc[3], cr, cc, sr, sc, -ca, ca,
                      hr.get("cede_temperature")))
    for (S, ncell, radm, colm, rad_cut, col_cut, azi_cut, dphi, temp) in specs:
        polar = (colm - radm < 0.0) or (colm + radm > math.pi)
        th, ph, r, area, maxA, cg, eg = construct_cell_mesh(
            polar, ncell, S, st["mass"], st["r_s"], st["R"], st["zeta"],
            st["epsilon"], st["shape"], radm, colm, rad_cut, col_cut, azi_cut)
        ph = ph + dphi
        if hr.get("is_antiphased", False):
            ph = ph + math.pi
        u = np.ascontiguousarray(st["r_s"] / r)
        t, defl, cosa, lag, maxD = compute_rays(st["r_s"], u, maxA,
                                                int(hr.get("num_rays", 200)))
        if t == 1:
            return None, "RayError"
        R_i = st["R"] / st["r_s"]
        Cc = (1.0 / u - R_i) + np.log((1.0 / u - 1.0) / (R_i - 1.0))
        Cc *= st["r_s"] / c_light
        lag = (lag - Cc[:, None]) * (mode_frequency * 2.0 * math.pi)
        radiates = (area > 0.0).astype(np.int32)
        params = np.ones((int(S), int(S), 2))
        params[..., 0] = temp
        params[..., 1] = eg[:, None]
        members.append(dict(S=int(S), theta=th, phi=ph, r=r, cellArea=area,
                            maxAlpha=maxA, cos_gamma=cg, effGrav=eg,
                            r_s_over_r=u, deflection=defl, cos_alpha=cosa,
                            lag=np.ascontiguousarray(lag), maxDeflection=maxD,
                            radiates=radiates, params=params))
    return members, None


def hot_region_integrate(st, hr, members, energies, atmosphere=None, else_logT=None):
    """HotRegion.integrate per member; with else_logT, every cell subtracts
    the blackbody elsewhere intensity (HotRegion.py:1056-1069)."""
    N_L = int(hr.get("num_leaves", 64))
    N_P = int(hr.get("num_phases") or N_L)
    leaves = np.linspace(0.0, 2.0 * math.pi, N_L)
    phases = 2.0 * math.pi * np.linspace(0.0, 1.0, N_P)
    out = []
    for m in members:
        corr = None
        if else_logT is not None:
            corr = np.ones_like(m["params"])
            corr[..., 0] = else_logT
            corr[..., 1] = m["effGrav"][:, None]
        t, flux = integrate(st["Omega"], st["r_s"], st["i"], m["cellArea"],
                            m["r"], m["r_s_over_r"], m["theta"], m["phi"],
                            m["params"], m["radiates"], m["deflection"],
                            m["cos_alpha"], m["lag"], m["maxDeflection"],
                            m["cos_gamma"], energies, leaves, phases,
                            atm_ext=hr.get("atm_ext", 1),
                            atmosphere=atmosphere,
                            image_order_limit=hr.get("image_order_limit"),
                            R=st["R"], correction=corr)
        if t == 1:
            return None
        out.append(flux)
    return out


def loglikelihood(model, p=None, return_parts=False):
    """One likelihood evaluation (Likelihood._driver, CustomSignal.__call__).

    `model` is a dict: spacetime (frequency, star_shape), hot_regions (list
    of dicts as for hot_region_embed, each with its parameter values),
    instrument (matrix[N_ch, N_in], energy_edges[N_in+1]), data (counts,
    phases, exposure_time), num_energies, support, marginal settings,
    optional interstellar callable and atmosphere.
    """
    sp = model["spacetime"]
    st = spacetime(sp["mass"], sp["radius"], sp["distance"],
                   sp["cos_inclination"], sp["frequency"],
                   sp.get("star_shape", "AGM_14"))
    llzero = model.get("llzero", -1.0e90)
    edges = model["instrument"]["energy_edges"]
    energies = construct_energy_array(model["num_energies"], edges)
    signals, phase_sets, shifts = [], [], []
    es = model.get("elsewhere")
    spectrum, else_logT = None, None
    if es is not None:
        # Photosphere.embed/integrate (xpsi/Photosphere.py:446-593)
        em, err = elsewhere_embed(st, es)
        if err is not None:
            return llzero * 0.5 if not return_parts else (llzero * 0.5, None)
        else_logT = es["elsewhere_temperature"]
        t, spectrum = elsewhere_integrate(st, em, energies, else_logT,
                                          es.get("image_order_limit"))
        if t == 1:
            return llzero * 0.5 if not return_parts else (llzero * 0.5, None)
    for g, hr in enumerate(model["hot_regions"]):
        members, err = hot_region_embed(st, hr, sp["frequency"])
        if err is not None:
            return llzero * 0.5 if not return_parts else (llzero * 0.5, None)
        fluxes = hot_region_integrate(st, hr, members, energies,
                                      model.get("atmosphere"), else_logT)
        if fluxes is None:
            return llzero * 0.5 if not return_parts else (llzero * 0.5, None)
        if spectrum is not None and g == 0:
            # the time-invariant spectrum joins the first component (Photosphere.py:590-592)
            fluxes[0] = fluxes[0] + spectrum[:, None]
        integrated = None
        for f in fluxes:
            tmp = energy_integrator(f / st["d_sq"], np.log10(energies),
                                    np.log10(edges))
            integrated = tmp if integrated is None else integrated + tmp
        if model.get("interstellar") is not None:
            mids = 0.5 * (edges[:-1] + edges[1:])
            integrated = integrated * model["interstellar"](mids)[:, None]
        folded = model["instrument"]["matrix"] @ integrated
        signals.append(np.ascontiguousarray(folded))
        N_P = int(hr.get("num_phases") or hr.get("num_leaves", 64))
        phase_sets.append(np.linspace(0.0, 1.0, N_P))
        shifts.append(hr["phase_shift"])
    data = model["data"]
    counts = np.ascontiguousarray(data["counts"], dtype=np.float64)
    pre = precomputation(counts.astype(np.int32))
    support = model.get("support")
    if support is None:
        support = -1.0 * np.ones((counts.shape[0], 2))
        support[:, 0] = 0.0
    ms = model.get("marginal", {})
    lnl, STAR, bml, bmls = eval_marginal_likelihood(
        data["exposure_time"], data["phases"], counts, signals, phase_sets,
        np.asarray(shifts, dtype=np.float64), pre, support,
        ms.get("epsabs", 0.0), ms.get("epsrel", 1.0e-8),
        ms.get("epsilon", 1.0e-3), ms.get("sigmas", 10.0), llzero)
    if lnl <= llzero:
        lnl = llzero * 0.5
    if return_parts:
        return lnl, dict(energies=energies, signals=signals, STAR=STAR)
    return lnl
