"""Oracle components checked against mathematics (no reference output needed).

Each check is chosen so a plausible slip (dropped term, wrong index or sign,
transposed operand) fails: interpolants must reproduce data and be exact on
the function class they preserve; the Gauss-Kronrod rule must integrate
polynomials up to its degree exactly; areas and rays must hit closed forms.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate as si

from oracle import oracle as O


def test_interval_convention():
    x = np.array([0.0, 1.0, 2.0, 3.0])
    L = O.lib()
    assert L.or_interval(x, 4, 0.0) == 0
    assert L.or_interval(x, 4, 0.999) == 0
    assert L.or_interval(x, 4, 1.0) == 1
    assert L.or_interval(x, 4, 3.0) == 2      # x_max maps to the last interval


def test_steffen_nodes_linear_and_monotone():
    rng = np.random.default_rng(1)
    x = np.cumsum(rng.uniform(0.1, 1.0, 30))
    y = np.cumsum(rng.uniform(0.0, 1.0, 30))       # monotone data
    sp = O.steffen(x, y)
    for k in range(len(x)):
        assert O.steffen_eval(sp, x[k]) == pytest.approx(y[k], abs=1e-12)
    fine = np.linspace(x[0], x[-1], 3001)
    vals = np.array([O.steffen_eval(sp, v) for v in fine])
    assert np.all(np.diff(vals) >= -1e-12)          # Steffen preserves monotonicity
    lin = O.steffen(x, 2.0 * x + 1.0)
    for v in fine[::97]:
        assert O.steffen_eval(lin, v) == pytest.approx(2.0 * v + 1.0, rel=1e-13)
        assert O.steffen_eval(lin, v, deriv=True) == pytest.approx(2.0, rel=1e-12)


def test_akima_nodes_linear_and_integral():
    rng = np.random.default_rng(2)
    x = np.cumsum(rng.uniform(0.1, 1.0, 25))
    y = np.sin(x)
    for periodic in (False, True):
        sp = O.akima(x, y, periodic)
        for k in range(len(x)):
            assert O.akima_eval(sp, x[k]) == pytest.approx(y[k], abs=1e-13)
    # natural ends extrapolate slopes linearly, so linear data are reproduced
    lin = O.akima(x, 3.0 * x - 2.0, periodic=False)
    a, b = x[0] + 0.3, x[-1] - 0.2
    assert O.akima_integ(lin, a, b) == pytest.approx(1.5 * (b * b - a * a) - 2.0 * (b - a), rel=1e-12)
    # integral of the interpolant equals a fine quadrature of its evaluations
    sp = O.akima(x, y, True)
    ref = si.quad(lambda t: O.akima_eval(sp, t), x[2] + 0.1, x[-3], limit=400, points=list(x[3:-3]))[0]
    assert O.akima_integ(sp, x[2] + 0.1, x[-3]) == pytest.approx(ref, rel=1e-10)


def test_akima_flat_run_is_zero():
    # zero runs (hidden phases) must stay exactly zero inside the run
    x = np.linspace(0.0, 1.0, 12)
    y = np.array([0, 0, 0, 0, 1.0, 2.0, 1.5, 0, 0, 0, 0, 0])
    sp = O.akima(x, y, True)
    for v in np.linspace(x[0], x[3], 17):
        assert O.akima_eval(sp, v) == 0.0
    for v in np.linspace(x[8], x[11], 17):
        assert O.akima_eval(sp, v) == 0.0


_FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


def _gk61(f, a, b):
    L = O.lib()
    L.or_gk61.argtypes = [_FN, C.c_void_p, C.c_double, C.c_double] + [C.POINTER(C.c_double)] * 4
    cb = _FN(lambda x, p: f(x))
    r, e, ra, rs = (C.c_double() for _ in range(4))
    L.or_gk61(cb, None, a, b, C.byref(r), C.byref(e), C.byref(ra), C.byref(rs))
    return r.value, e.value


def _qag(f, a, b, epsrel):
    L = O.lib()
    L.or_qag61.argtypes = [_FN, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                           C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.or_qag61.restype = C.c_int
    cb = _FN(lambda x, p: f(x))
    r, e = C.c_double(), C.c_double()
    st = L.or_qag61(cb, None, a, b, 0.0, epsrel, 100, C.byref(r), C.byref(e))
    return st, r.value


def test_gk61_polynomial_exactness():
    # 61-point Kronrod rule is exact up to degree 91; degree 92+ is not
    for deg in (0, 1, 2, 17, 60, 91):
        r, _ = _gk61(lambda x, d=deg: x ** d, 0.0, 1.0)
        assert r == pytest.approx(1.0 / (deg + 1), rel=1e-13)
    # a single pass cannot resolve a kink (the adaptive layer exists for that)
    r, e = _gk61(lambda x: abs(x - 0.3), 0.0, 1.0)
    assert abs(r - 0.29) > 1e-8 and e > abs(r - 0.29)


def test_qag_known_integrals():
    st, r = _qag(lambda x: 1.0 / math.sqrt(x) if x > 0 else 0.0, 0.0, 1.0, 1e-10)
    assert r == pytest.approx(2.0, rel=1e-9)
    st, r = _qag(lambda x: math.exp(-x * x), -5.0, 5.0, 1e-12)
    assert st == 0 and r == pytest.approx(math.sqrt(math.pi) * math.erf(5.0), rel=1e-12)
    st, r = _qag(lambda x: abs(x - 0.3), 0.0, 1.0, 1e-10)   # kink
    assert r == pytest.approx(0.5 * (0.09 + 0.49), rel=1e-9)


def test_spherical_mesh_area_closed_form():
    # on a sphere the radiating cells tile the cap: sum of areas = 2 pi R^2 (1 - cos rho)
    R, rho, colat = 12e3, 0.4, 0.9
    st = O.spacetime(1.4, 12.0, 1.0, 0.5, 300.0, "sphere")
    al = O.allocate_cells(32 * 32, 10, 80, st["M"], st["r_s"], st["R"], st["Omega"], st["zeta"],
                          st["epsilon"], 1, rho, 0.0, colat, colat, 0.0, colat, 0.0, 0.0)
    th, ph, r, area, maxA, cg, eg = O.construct_cell_mesh(False, al[1], al[0], 1.4, st["r_s"], R,
                                                          st["zeta"], st["epsilon"], 1, rho, colat,
                                                          0.0, colat, 0.0)
    assert area.sum() == pytest.approx(2.0 * math.pi * R * R * (1.0 - math.cos(rho)), rel=1e-9)
    assert np.all(cg == 1.0)                        # no surface tilt on a sphere
    # polar cap version
    th, ph, r, area, *_ = O.construct_cell_mesh(True, al[1], al[0], 1.4, st["r_s"], R, st["zeta"],
                                                st["epsilon"], 1, rho, 0.2, 0.0, 0.2, 0.0)
    assert area.sum() == pytest.approx(2.0 * math.pi * R * R * (1.0 - math.cos(rho)), rel=1e-8)


def test_rays_limits():
    # radial ray: no deflection, no lag; weak field: deflection -> emission angle
    u = np.array([1e-9, 0.3])
    maxA = np.array([1.5, 1.5])
    t, defl, cosa, lag, maxD = O.compute_rays(1.0, u, maxA, 20)
    assert t == 0
    assert defl[:, 0].tolist() == [0.0, 0.0] and lag[:, 0].tolist() == [0.0, 0.0]
    assert np.allclose(defl[0], np.arccos(cosa[0]), rtol=1e-6, atol=1e-8)
    # strong field deflects more than the emission angle, monotonically
    assert np.all(defl[1, 1:] > np.arccos(cosa[1, 1:]))
    assert np.all(np.diff(defl[1]) > 0)


def test_precomputation_is_minus_sum_lnfact():
    from scipy.special import gammaln
    rng = np.random.default_rng(3)
    counts = rng.integers(0, 200, size=(7, 9)).astype(np.int32)
    pre = O.precomputation(counts)
    assert np.allclose(pre, -gammaln(counts + 1.0).sum(axis=1), rtol=1e-13)


# ------------------------------------------ the phase interpolant choice
# xpsi/tools/core.pyx:21-116: "Akima" (periodic, default), "Steffen", "Cubic"
# (periodic cubic spline).  The periodic cubic spline is the unique C2
# periodic interpolant, so scipy's CubicSpline(bc_type="periodic") pins it.
@pytest.fixture
def phase_interpolant():
    yield O.set_phase_interpolant
    O.set_phase_interpolant("Akima")


def _periodic_data(seed, n):
    rng = np.random.default_rng(seed)
    x = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0.0, 1.0, n - 2)]))
    y = np.cos(2 * np.pi * x) + 0.3 * np.sin(6 * np.pi * x) + rng.normal(0.0, 0.05, n)
    y[-1] = y[0]
    return x, y


def test_periodic_cubic_spline_matches_scipy(phase_interpolant):
    from scipy.interpolate import CubicSpline
    for seed, n in ((1, 5), (2, 17), (3, 64)):
        x, y = _periodic_data(seed, n)
        ref = CubicSpline(x, y, bc_type="periodic")
        phase_interpolant("Cubic")
        sp = O.phase_spline(x, y)
        xq = np.random.default_rng(seed + 10).uniform(0.0, 1.0, 200)
        got = np.array([O.akima_eval(sp, q) for q in xq])
        assert np.max(np.abs(got - ref(xq))) < 1e-12 * np.max(np.abs(y))
        for a, b in ((0.1, 0.7), (0.0, 1.0), (0.33, 0.34)):
            assert abs(O.akima_integ(sp, a, b) - ref.integrate(a, b)) < 1e-12 * np.max(np.abs(y))


def test_steffen_as_phase_interpolant(phase_interpolant):
    x, y = _periodic_data(4, 33)
    st = O.steffen(x, y)
    phase_interpolant("Steffen")
    sp = O.phase_spline(x, y)
    for q in np.linspace(0.0, 1.0, 101):
        assert abs(O.akima_eval(sp, q) - O.steffen_eval(st, q)) < 1e-14 * np.max(np.abs(y))
    # the integral of the same cubic pieces, against quadrature
    val, _ = si.quad(lambda q: O.steffen_eval(st, q), 0.2, 0.9, points=list(x[(x > 0.2) & (x < 0.9)]),
                     limit=200, epsabs=1e-14, epsrel=1e-13)
    assert abs(O.akima_integ(sp, 0.2, 0.9) - val) < 1e-12


def test_phase_interpolants_in_the_integrator(phase_interpolant):
    # every interpolant converges to the same pulse with the leaf count, and
    # the choice reaches the integrator (different at a finite leaf count)
    st = O.spacetime(1.4, 12.0, 1.0, math.cos(1.0), 300.0, "AGM_14")
    hr = dict(super_colatitude=1.2, super_radius=0.8, super_temperature=6.3, phase_shift=0.0,
              sqrt_num_cells=24, min_sqrt_num_cells=10, max_sqrt_num_cells=100, num_rays=512, num_leaves=64)
    members, err = O.hot_region_embed(st, hr, st["f"])
    m = members[0]
    energies = np.array([0.5, 2.0])
    phases = 2.0 * math.pi * np.linspace(0.0, 1.0, 32)

    def flux(N_L):
        leaves = np.linspace(0.0, 2.0 * math.pi, N_L)
        t, f = O.integrate(st["Omega"], st["r_s"], st["i"], m["cellArea"], m["r"], m["r_s_over_r"], m["theta"],
                           m["phi"], m["params"], m["radiates"], m["deflection"], m["cos_alpha"], m["lag"],
                           m["maxDeflection"], m["cos_gamma"], energies, leaves, phases, R=st["R"])
        assert t == 0
        return f

    ref = flux(512)
    out = {}
    for name in ("Akima", "Steffen", "Cubic"):
        phase_interpolant(name)
        out[name] = [np.max(np.abs(flux(n) - ref)) / np.max(ref) for n in (48, 96)]
    for name, (e48, e96) in out.items():
        assert e96 < e48 and e96 < 1e-3, (name, e48, e96)
    assert out["Steffen"][0] != out["Akima"][0] and out["Cubic"][0] != out["Akima"][0]
