"""Oracle functions that no reference output covers, pinned by mathematics.

- The 4-D table atmosphere (xpsi/surface_radiation_field/hot_Num4D.pyx:248-438):
  cubic Lagrange interpolation on four nodes per axis reproduces any
  tensor-product cubic polynomial exactly, in range and when it extrapolates,
  so a table sampled from such a polynomial P must return P at every query.
  The base-node state machine (:300-351) then shows only through its clamp:
  a query beyond an axis edge is clamped to the edge on the call that moves
  the base node to the end of the axis, and extrapolated on later calls.
- The ceding annulus (xpsi/cellmesh/mesh.pyx:23-417, polar_mesh.pyx:23-430):
  on a spherical star its cells tile the cede cap minus the superseding cap.
- The blackbody integrator (integrator_for_azimuthal_invariance.pyx:70-665):
  in flat space, without rotation, a polar cap or ring seen pole-on has the
  closed-form flux norm * B(E) * pi R^2 (sin^2 rho_out - sin^2 rho_in) / (E keV)
  at every phase.
- The integrator without azimuthal invariance (xpsi/cellmesh/integrator.pyx
  :48-667), per-cell temperatures: seen pole-on in flat space a cap split into
  two azimuthal halves at two temperatures gives the mean of the two closed
  forms, and a per-cell elsewhere correction the difference of two; on an
  inclined, rotating, oblate star with one temperature it converges with the
  leaf count to the invariant integrator (independent interpolation: GEOM and
  Z there against the specific-intensity profile here); and rolling a
  temperature map by one cell along a parallel of a pole-centred mesh shifts
  the pulse by that cell's azimuth step.
- The Stokes integrator and the hot-region options
  (integratorIQU_for_azimuthal_invariance.pyx:68-746, common_functions.pyx,
  hot_BB_burst.pyx, hot_wrapper.pyx:115-233): the reference's own checks of
  the polarisation angle (its non-relativistic limit, xpsi/tests/
  test_cellmesh.py:38-95) and of disk blocking (a disk inside the star blocks
  every southern ray, :97-137); with the blackbody it reproduces the plain
  integrator bit for bit with Q = U = 0; pole-on in flat space the burst
  atmosphere's and the beamed intensity's fluxes are closed forms (integrals
  of mu^n over the cap), Q and U cancel, and beam option 3's quadrature
  converges to option 2; a small spot's Q/I and U/I follow
  PD(mu) cos 2chi_0, PD(mu) sin 2chi_0 around the rotation; a northern spot
  is never blocked and a southern one behind a disk inside the star is dark.
- The general integrator's Stokes form and options (integratorIQU.pyx:51-717,
  integrator.pyx with atmospheres and beaming): with the blackbody it is the
  plain general integrator bit for bit with Q = U = 0; pole-on the burst and
  beamed fluxes are the closed forms; the reference forms one polarisation
  angle per parallel and image (from the leaf loop's last values), so for one
  radiating parallel and one image U/Q is the same at every energy and phase,
  and seen pole-on in flat space (mu eta = cos theta on the whole parallel)
  sqrt(Q^2 + U^2) = |PD(cos theta)| I.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

K_B_OVER_KEV = 1.38064852e-23 / 1.60217662e-16
BB_NORM = 1.0e-7 * 5.040366110812353e22
KEV = 1.60217662e-16


def _poly_table(seed, shape, negative=False):
    """Axes (non-uniform) and a table sampled from a sum of three products of
    per-axis cubics; returns (atmosphere tuple, P callable)."""
    rng = np.random.default_rng(seed)
    lims = [(5.1, 6.8), (13.7, 15.0), (0.05, 1.0), (-1.3, 2.0)]
    axes = []
    for (a, b), n in zip(lims, shape):
        x = np.sort(rng.uniform(0.0, 1.0, n - 2))
        axes.append(np.concatenate([[a], a + (b - a) * x, [b]]))
    centre = [0.5 * (a + b) for a, b in lims]
    coef = rng.uniform(-1.0, 1.0, (3, 4, 4))
    const = 0.0 if negative else 40.0

    def P(x0, x1, x2, x3):
        xs = [x0, x1, x2, x3]
        tot = const
        for k in range(3):
            term = 1.0
            for d in range(4):
                t = xs[d] - centre[d]
                c = coef[k, d]
                term = term * (c[0] + t * (c[1] + t * (c[2] + t * c[3])))
            tot = tot + term
        return tot

    grid = np.meshgrid(*axes, indexing="ij")
    table = np.ascontiguousarray(P(*grid).ravel())
    return (axes[0], axes[1], axes[2], axes[3], table), P


def _energy(logT, x3):
    return K_B_OVER_KEV * 10.0 ** logT * 10.0 ** x3


def test_num4d_reproduces_cubic_polynomials_in_range():
    atm, P = _poly_table(11, (6, 5, 7, 9))
    ax = atm[:4]
    rng = np.random.default_rng(12)
    n = 400
    # random order: the base node walks up and down every axis
    q = [rng.uniform(a[0], a[-1], n) for a in ax]
    E = _energy(q[0], q[3])
    out = O.num4d_sequence(atm, E, q[2], q[0], q[1])
    x3 = np.log10(E / (K_B_OVER_KEV * 10.0 ** q[0]))
    ref = P(q[0], q[1], q[2], x3) * 10.0 ** (3.0 * q[0])
    assert np.all(ref > 0.0)
    np.testing.assert_allclose(out, ref, rtol=1e-11)
    # a transposed table (axes 2 and 3 swapped in memory) must not pass
    bad = np.ascontiguousarray(atm[4].reshape(6, 5, 7, 9).transpose(0, 1, 3, 2).ravel())
    bad_atm = (ax[0], ax[1], ax[3], ax[2], bad)
    out_bad = O.num4d_sequence(bad_atm, E, q[2], q[0], q[1])
    assert np.max(np.abs(out_bad / ref - 1.0)) > 1e-3


def test_num4d_negative_values_clip_to_zero():
    atm, P = _poly_table(13, (5, 5, 6, 7), negative=True)
    ax = atm[:4]
    rng = np.random.default_rng(14)
    q = [rng.uniform(a[0], a[-1], 200) for a in ax]
    E = _energy(q[0], q[3])
    out = O.num4d_sequence(atm, E, q[2], q[0], q[1])
    x3 = np.log10(E / (K_B_OVER_KEV * 10.0 ** q[0]))
    ref = P(q[0], q[1], q[2], x3) * 10.0 ** (3.0 * q[0])
    neg = ref < 0.0
    assert neg.any() and (~neg).any()
    assert np.all(out[neg] == 0.0)
    np.testing.assert_allclose(out[~neg], ref[~neg], rtol=1e-9)


def test_num4d_clamp_then_extrapolate():
    atm, P = _poly_table(15, (6, 5, 7, 9))
    ax = atm[:4]
    lT, lg, mu = 6.0, 14.3, 0.6
    x = ax[3]
    top, bot = x[-1], x[0]
    seq = [
        (x[1] * 0.5 + x[2] * 0.5, None),    # in range, base node stays at 0
        (top + 0.3, top),                   # moves the base node to the end: clamped
        (top + 0.3, None),                  # base node already at the end: extrapolated
        (top + 0.1, None),                  # extrapolated
        (0.5 * (x[3] + x[4]), None),        # back in range
        (top + 0.25, top),                  # moves to the end again: clamped
        (bot - 0.2, bot),                   # moves the base node to 0: clamped
        (bot - 0.2, None),                  # base node already at 0: extrapolated
        (bot - 0.05, None),
    ]
    x3q = np.array([s[0] for s in seq])
    E = _energy(lT, x3q)
    out = O.num4d_sequence(atm, E, np.full(len(seq), mu), lT, lg)
    for k, (q, clamp) in enumerate(seq):
        xe = np.log10(E[k] / (K_B_OVER_KEV * 10.0 ** lT))
        want = P(lT, lg, mu, clamp if clamp is not None else xe) * 10.0 ** (3.0 * lT)
        assert out[k] == pytest.approx(want, rel=1e-11), (k, q, clamp)
    # the clamped and the extrapolated values genuinely differ
    assert abs(out[1] / out[2] - 1.0) > 1e-4 and abs(out[6] / out[7] - 1.0) > 1e-4
    # a fresh state starts at base node 0: a first query below the axis extrapolates
    out0 = O.num4d_sequence(atm, _energy(lT, np.array([bot - 0.2])), np.array([mu]), lT, lg)
    xe = np.log10(_energy(lT, bot - 0.2) / (K_B_OVER_KEV * 10.0 ** lT))
    assert out0[0] == pytest.approx(P(lT, lg, mu, xe) * 10.0 ** (3.0 * lT), rel=1e-11)


def test_num4d_temperature_clamp_scales_with_the_clamped_value():
    # the result is multiplied by 10^(3 logT) of the clamped logT, while
    # log10(E/kT) is formed from the query's own logT (hot_Num4D.pyx:295-298,438)
    atm, P = _poly_table(16, (6, 5, 7, 9))
    ax = atm[:4]
    lg, mu = 14.1, 0.4
    lTs = np.array([5.5, ax[0][-1] + 0.1])
    x3 = np.array([0.3, 0.3])
    E = _energy(lTs, x3)
    out = O.num4d_sequence(atm, E, np.full(2, mu), lTs, lg)
    lt_c = ax[0][-1]
    want = P(lt_c, lg, mu, 0.3) * 10.0 ** (3.0 * lt_c)
    assert out[1] == pytest.approx(want, rel=1e-10)


def _sphere():
    return O.spacetime(1.4, 12.0, 1.0, 0.5, 300.0, "sphere")


@pytest.mark.parametrize("case", ["concentric", "offset", "polar"])
def test_ceding_annulus_area_closed_form(case):
    st = _sphere()
    if case == "concentric":
        hr = dict(super_colatitude=0.9, super_radius=0.15, cede_colatitude=0.9, cede_radius=0.35)
    elif case == "offset":
        hr = dict(super_colatitude=0.95, super_radius=0.12, cede_colatitude=1.0, cede_radius=0.4,
                  cede_azimuth=0.08)
    else:
        hr = dict(super_colatitude=0.15, super_radius=0.1, cede_colatitude=0.2, cede_radius=0.35)
    hr.update(super_temperature=6.3, cede_temperature=6.0, cede=True, phase_shift=0.0,
              sqrt_num_cells=32, num_rays=64, num_leaves=32)
    members, err = O.hot_region_embed(st, hr, 300.0)
    assert err is None and len(members) == 2
    R = st["R"]
    sup = members[0]["cellArea"].sum()
    ced = members[1]["cellArea"].sum()
    cap = lambda rho: 2.0 * math.pi * R * R * (1.0 - math.cos(rho))
    assert sup == pytest.approx(cap(hr["super_radius"]), rel=1e-8)
    assert ced == pytest.approx(cap(hr["cede_radius"]) - cap(hr["super_radius"]), rel=1e-8)
    # the hole is really cut: the annulus has cells that do not radiate
    assert np.count_nonzero(members[1]["cellArea"] == 0.0) > 0


def _flat_pole_on(rho_in, rho_out, S_sqrt, logT):
    """A pole-centred cap (rho_in = 0) or ring on a nearly massless,
    non-rotating sphere seen from the pole, integrated by the oracle."""
    st = O.spacetime(1e-7, 12.0, 1.0, 1.0, 1e-9, "sphere")
    if rho_in > 0.0:
        hr = dict(super_colatitude=0.0, super_radius=rho_in, cede_colatitude=0.0,
                  cede_radius=rho_out, cede=True, cede_temperature=logT)
    else:
        hr = dict(super_colatitude=0.0, super_radius=rho_out)
    hr.update(super_temperature=logT, phase_shift=0.0, sqrt_num_cells=S_sqrt,
              min_sqrt_num_cells=10, max_sqrt_num_cells=400, num_rays=400, num_leaves=32)
    members, err = O.hot_region_embed(st, hr, st["f"])
    assert err is None
    energies = np.array([0.5, 1.0, 2.0, 4.0])
    fl = O.hot_region_integrate(st, hr, members, energies)
    flux = fl[-1]
    kT = K_B_OVER_KEV * 10.0 ** logT
    B = energies ** 3 / np.expm1(energies / kT)
    R = st["R"]
    exact = BB_NORM * B * math.pi * R * R * (math.sin(rho_out) ** 2 - math.sin(rho_in) ** 2)
    exact = exact / (energies * KEV)
    return flux, exact


def test_blackbody_flat_pole_on_closed_form():
    errs = []
    for S in (16, 48):
        flux, exact = _flat_pole_on(0.0, 0.5, S, 6.2)
        # the same at every phase (axisymmetric view), equal to the closed form
        assert np.all(np.abs(flux / flux[:, :1] - 1.0) < 1e-9)
        errs.append(np.max(np.abs(flux[:, 0] / exact - 1.0)))
    # converges with the mesh: the residual is the cell quadrature's error
    assert errs[1] < 2e-5 and errs[1] < errs[0] / 4.0
    flux, exact = _flat_pole_on(0.2, 0.5, 48, 6.2)
    assert np.all(np.abs(flux / flux[:, :1] - 1.0) < 1e-9)
    assert np.max(np.abs(flux[:, 0] / exact - 1.0)) < 3e-5


def test_energy_integrator_power_law():
    # the integral over an input bin of 10^logE F ln10 d(logE) is the integral
    # of F dE; for F = E^-1.5 that is 2 (lo^-0.5 - hi^-0.5), to the spline's error
    E = np.logspace(np.log10(0.3), np.log10(10.0), 200)
    F = E ** -1.5
    sig = np.repeat(F[:, None], 3, axis=1) * np.array([1.0, 2.0, 0.5])[None, :]
    edges = np.logspace(np.log10(0.35), np.log10(12.0), 30)
    out = O.energy_integrator(sig, np.log10(E), np.log10(edges))
    hi = np.minimum(edges[1:], E[-1])
    exact = 2.0 * (edges[:-1] ** -0.5 - hi ** -0.5)
    inside = edges[:-1] < E[-1]
    # the reference's periodic Akima end conditions distort the last two
    # energy intervals (a non-periodic spectrum); away from them the bins agree
    away = (edges[:-1] > E[2]) & (edges[1:] < E[-3])
    np.testing.assert_allclose(out[away, 0], exact[away], rtol=2e-6)
    np.testing.assert_allclose(out[:, 1], 2.0 * out[:, 0], rtol=1e-14)   # linear in the signal
    assert np.all(out[~inside] == 0.0)                                   # beyond the grid


def _marginal_brute(T, N_b, counts, star_rate, support):
    """sum over channels of log integral over B of prod_j Poisson(d_j | c_j(B)),
    c_j = T star_rate_j + (T / N_b) B, on the support (upper < 0: unbounded)."""
    from scipy import integrate as si
    from scipy.special import gammaln
    tot = 0.0
    for i in range(counts.shape[0]):
        d = counts[i]
        lam = lambda B: T * star_rate[i] + (T / N_b) * B
        logf = lambda B: float(np.sum(d * np.log(lam(B)) - lam(B)))
        lo = max(support[i, 0], 0.0)
        hi = support[i, 1] if support[i, 1] > 0.0 else 1e3
        # the integrand's peak and width, then quadrature over +-40 widths of it
        from scipy import optimize as so
        Bp = so.minimize_scalar(lambda b: -logf(b), bounds=(lo, hi), method="bounded",
                                options=dict(xatol=1e-12)).x
        w = math.sqrt(1.0 / max(np.sum(d / lam(Bp) ** 2) * (T / N_b) ** 2, 1e-300))
        a, b = max(lo, Bp - 40.0 * w), min(hi, Bp + 40.0 * w)
        A = logf(Bp)
        val, _ = si.quad(lambda x: math.exp(logf(x) - A), a, b, points=[Bp], limit=400,
                         epsabs=0.0, epsrel=1e-12)
        tot += math.log(val) + A - float(np.sum(gammaln(d + 1.0)))
    return tot


@pytest.mark.parametrize("bounded", [False, True])
def test_marginal_likelihood_brute_force(bounded):
    # expected counts from the phase-bin integral of a shifted periodic pulse
    # (bins wrap at the cycle's end), then the background integral by quadrature
    rng = np.random.default_rng(21)
    T, N_b, N_ch, NP = 1000.0, 8, 3, 48
    ph = np.linspace(0.0, 1.0, NP)
    amp = np.array([0.02, 0.05, 0.01])
    pulses = amp[:, None] * (1.0 + 0.6 * np.sin(2.0 * math.pi * ph))[None, :]
    shift = 0.3
    edges = np.linspace(0.0, 1.0, N_b + 1)
    # the bin integral of each channel's pulse, by quadrature of the interpolant
    from scipy import integrate as si
    star = np.zeros((N_ch, N_b))
    for i in range(N_ch):
        sp = O.akima(ph, pulses[i], True)
        f = lambda x: O.akima_eval(sp, x - math.floor(x))
        for j in range(N_b):
            a, b = edges[j] + shift, edges[j + 1] + shift
            pts = [math.floor(a) + 1.0] if math.floor(a) != math.floor(b) else None
            star[i, j] = si.quad(f, a, b, points=pts, limit=200, epsrel=1e-13)[0]
    B_true = np.array([0.2, 0.1, 0.05])
    lam = T * star + (T / N_b) * B_true[:, None]
    counts = rng.poisson(lam).astype(np.float64)
    if bounded:
        support = np.stack([0.5 * B_true, 1.5 * B_true], axis=1)
    else:
        support = np.zeros((N_ch, 2))
        support[:, 1] = -1.0
    pre = O.precomputation(counts.astype(np.int32))
    lnl, STAR, bml, bmls = O.eval_marginal_likelihood(T, edges, counts, [pulses], [ph],
                                                      np.array([shift]), pre, support)
    ref = _marginal_brute(T, N_b, counts, star, support)
    assert lnl == pytest.approx(ref, abs=1e-7)
    # a dropped shift moves the answer by far more than the bar
    lnl0, *_ = O.eval_marginal_likelihood(T, edges, counts, [pulses], [ph], np.array([0.0]),
                                          pre, support)
    assert abs(lnl0 - ref) > 1e-3


def _two_component_case():
    """Two pulses on one phase grid with their shifts, and each channel's
    phase-bin integrals of the shifted Akima interpolants by Gauss-Legendre on
    every cubic piece (exact) between the knots."""
    rng = np.random.default_rng(31)
    T, N_b, N_ch, NP = 2000.0, 10, 4, 40
    ph = np.linspace(0.0, 1.0, NP)
    amp = np.array([0.03, 0.01, 0.05, 0.02])
    pulses = amp[:, None] * (1.0 + 0.7 * np.cos(2.0 * math.pi * (ph - 0.2)))[None, :]
    second = 0.5 * amp[:, None] * (1.0 + 0.4 * np.sin(4.0 * math.pi * ph))[None, :]
    shifts = np.array([0.15, -0.4])
    edges = np.linspace(0.0, 1.0, N_b + 1)
    star = np.zeros((N_ch, N_b))
    for comp, sh in ((pulses, shifts[0]), (second, shifts[1])):
        for i in range(N_ch):
            sp = O.akima(ph, comp[i], True)
            f = lambda x: O.akima_eval(sp, x - math.floor(x))
            for j in range(N_b):
                a, b = edges[j] + sh, edges[j + 1] + sh
                knots = np.concatenate([ph + k for k in range(math.floor(a), math.floor(b) + 1)])
                cuts = np.unique(np.concatenate([[a, b], knots[(knots > a) & (knots < b)]]))
                gx, gw = np.polynomial.legendre.leggauss(4)
                tot = 0.0
                for lo, hi in zip(cuts[:-1], cuts[1:]):
                    xs = 0.5 * (hi - lo) * gx + 0.5 * (hi + lo)
                    tot += 0.5 * (hi - lo) * sum(w * f(x) for x, w in zip(xs, gw))
                star[i, j] += tot
    return rng, T, N_b, N_ch, ph, pulses, second, shifts, edges, star


def test_poisson_likelihood_given_background_brute_force():
    # expected counts = exposure x (phase-bin integral of the shifted pulse +
    # background rate); lnL = sum d log(lambda) - lambda - ln d!
    from scipy.special import gammaln
    rng, T, N_b, N_ch, ph, pulses, second, shifts, edges, star = _two_component_case()
    bg = rng.uniform(0.001, 0.004, (N_ch, N_b))
    lam = T * (star + bg)
    counts = rng.poisson(lam).astype(np.float64)
    pre = O.precomputation(counts.astype(np.int32))
    lnl, EXPEC = O.poisson_likelihood_given_background(T, edges, counts, [pulses, second], [ph, ph],
                                                       shifts, bg, pre)
    np.testing.assert_allclose(EXPEC, lam, rtol=1e-11)
    ref = float(np.sum(counts * np.log(lam) - lam - gammaln(counts + 1.0)))
    assert lnl == pytest.approx(ref, abs=1e-8)
    # without the factorial term
    lnl0, _ = O.poisson_likelihood_given_background(T, edges, counts, [pulses, second], [ph, ph],
                                                    shifts, bg)
    assert lnl0 == pytest.approx(ref + float(np.sum(gammaln(counts + 1.0))), abs=1e-8)
    # a bin with data but no expected counts: a draw in [-1e90, -1e89]
    zero = np.zeros_like(pulses)
    c2 = np.zeros_like(counts)
    c2[0, 0] = 3.0
    lz, _ = O.poisson_likelihood_given_background(T, edges, c2, [zero], [ph], shifts[:1],
                                                  np.zeros_like(bg))
    assert -1e90 <= lz <= -1e89
    # every such bin adds its own draw
    c2[1, 3] = 1.0
    lz2, _ = O.poisson_likelihood_given_background(T, edges, c2, [zero], [ph], shifts[:1],
                                                   np.zeros_like(bg))
    assert -2e90 <= lz2 <= -2e89


def test_gaussian_likelihood_given_background_against_normal_logpdf():
    # EXPEC as the Poisson statistic's (the brute-force phase-bin integrals);
    # lnL = sum of the normal log-densities of the counts about EXPEC
    from scipy.stats import norm
    rng, T, N_b, N_ch, ph, pulses, second, shifts, edges, star = _two_component_case()
    bg = rng.uniform(0.001, 0.004, (N_ch, N_b))
    lam = T * (star + bg)
    counts = lam + rng.normal(0.0, 1.0, lam.shape) * np.sqrt(lam)
    errors = np.sqrt(lam) * rng.uniform(0.8, 1.2, lam.shape)
    lnl, EXPEC = O.gaussian_likelihood_given_background(T, edges, counts, errors, [pulses, second], [ph, ph],
                                                        shifts, bg)
    np.testing.assert_allclose(EXPEC, lam, rtol=1e-11)
    ref = float(np.sum(norm.logpdf(counts, loc=lam, scale=errors)))
    assert lnl == pytest.approx(ref, rel=1e-12, abs=1e-9)
    # per-component allow_negative: the second component negated integrates
    # below zero in every bin; it is dropped without its flag, and with it the
    # sum is clipped at zero before the background is added
    first = _part(ph, pulses, shifts[0], edges)
    minus = -_part(ph, second, shifts[1], edges)
    l_drop, E_drop = O.gaussian_likelihood_given_background(T, edges, counts, errors, [pulses, -second],
                                                            [ph, ph], shifts, bg, allow_negative=[False, False])
    np.testing.assert_allclose(E_drop, T * (first + bg), rtol=1e-11)
    l_keep, E_keep = O.gaussian_likelihood_given_background(T, edges, counts, errors, [pulses, -second],
                                                            [ph, ph], shifts, bg, allow_negative=[False, True])
    np.testing.assert_allclose(E_keep, T * (np.maximum(first + minus, 0.0) + bg), rtol=1e-11)
    assert l_drop != l_keep


def _part(ph, comp, sh, edges):
    """Each channel's phase-bin integrals of one shifted Akima pulse (4-point
    Gauss-Legendre on every cubic piece, exact)."""
    N_ch, N_b = comp.shape[0], len(edges) - 1
    out = np.zeros((N_ch, N_b))
    gx, gw = np.polynomial.legendre.leggauss(4)
    for i in range(N_ch):
        sp = O.akima(ph, comp[i], True)
        f = lambda x: O.akima_eval(sp, x - math.floor(x))
        for j in range(N_b):
            a, b = edges[j] + sh, edges[j + 1] + sh
            knots = np.concatenate([ph + k for k in range(math.floor(a), math.floor(b) + 1)])
            cuts = np.unique(np.concatenate([[a, b], knots[(knots > a) & (knots < b)]]))
            for lo, hi in zip(cuts[:-1], cuts[1:]):
                xs = 0.5 * (hi - lo) * gx + 0.5 * (hi + lo)
                out[i, j] += 0.5 * (hi - lo) * sum(w * f(x) for x, w in zip(xs, gw))
    return out


def test_phase_integrator_brute_force():
    # each row's shifted Akima pulse integrated over the phase intervals
    # (4-point Gauss-Legendre on every cubic piece, exact), times the
    # exposure; a negative pulse gives zeros without allow_negative and its
    # negative integrals with it (no clip, unlike the expected counts)
    rng, T, N_b, N_ch, ph, pulses, second, shifts, edges, star = _two_component_case()
    for sh in (0.0, 0.37, -0.81):
        got = O.phase_integrator(T, edges, pulses, ph, sh)
        np.testing.assert_allclose(got, T * _part(ph, pulses, sh, edges), rtol=1e-12)
    neg = -pulses
    np.testing.assert_array_equal(O.phase_integrator(T, edges, neg, ph, 0.2), 0.0)
    np.testing.assert_allclose(O.phase_integrator(T, edges, neg, ph, 0.2, allow_negative=True),
                               -T * _part(ph, pulses, 0.2, edges), rtol=1e-12)


def test_gaussian_likelihood_QnUn_against_normal_logpdf():
    # one model row for every data row: sum over rows and bins of the normal
    # log-density of d_ij about star_j with scale errors_ij
    from scipy.stats import norm
    rng = np.random.default_rng(5)
    N_rows, N_b = 3, 16
    star = rng.uniform(-0.3, 0.3, N_b)
    counts = star[None, :] + rng.normal(0.0, 0.05, (N_rows, N_b))
    errors = rng.uniform(0.03, 0.08, (N_rows, N_b))
    lnl, out = O.gaussian_likelihood_QnUn(np.linspace(0.0, 1.0, N_b + 1), counts, errors, star)
    ref = float(np.sum(norm.logpdf(counts, loc=star[None, :], scale=errors)))
    assert lnl == pytest.approx(ref, rel=1e-12)
    np.testing.assert_array_equal(out, star)


def _general(st, hr, params=None, correction=None, N_L=64, phases=None, invariant=False):
    """Both integrators on member 0 of `hr`; params/correction are callables
    of the member returning [S, S, 2] arrays (default: the member's own)."""
    members, err = O.hot_region_embed(st, hr, st["f"])
    assert err is None
    m = members[0]
    energies = np.array([0.5, 1.0, 2.0, 4.0])
    leaves = np.linspace(0.0, 2.0 * math.pi, N_L)
    if phases is None:
        phases = 2.0 * math.pi * np.linspace(0.0, 1.0, N_L)
    P = m["params"] if params is None else params(m)
    C = None if correction is None else correction(m)
    geo = (m["cellArea"], m["r"], m["r_s_over_r"], m["theta"], m["phi"])
    rays = (m["deflection"], m["cos_alpha"], m["lag"], m["maxDeflection"], m["cos_gamma"])
    t, f = O.integrate_general(st["Omega"], st["i"], *geo, P, m["radiates"], *rays, energies, leaves,
                               phases, correction=C)
    assert t == 0
    if not invariant:
        return f, m
    t2, f2 = O.integrate(st["Omega"], st["r_s"], st["i"], *geo, m["params"], m["radiates"], *rays,
                         energies, leaves, phases, R=st["R"], correction=C)
    assert t2 == 0
    return f, f2


def _flat_cap(S):
    st = O.spacetime(1e-7, 12.0, 1.0, 1.0, 1e-9, "sphere")
    hr = dict(super_colatitude=0.0, super_radius=0.5, super_temperature=6.2, phase_shift=0.0,
              sqrt_num_cells=S, min_sqrt_num_cells=10, max_sqrt_num_cells=400, num_rays=400,
              num_leaves=32)
    return st, hr


def _closed(st, logT, rho=0.5):
    energies = np.array([0.5, 1.0, 2.0, 4.0])
    kT = K_B_OVER_KEV * 10.0 ** logT
    B = energies ** 3 / np.expm1(energies / kT)
    return BB_NORM * B * math.pi * st["R"] ** 2 * math.sin(rho) ** 2 / (energies * KEV)


def test_general_integrator_two_temperature_halves_pole_on():
    st, hr = _flat_cap(48)

    def halves(m):
        P = m["params"].copy()
        P[:, : P.shape[1] // 2, 0] = 6.0
        P[:, P.shape[1] // 2:, 0] = 6.3
        return P

    f, m = _general(st, hr, params=halves, N_L=32)
    assert np.all(np.abs(f / f[:, :1] - 1.0) < 1e-9)
    exact = 0.5 * (_closed(st, 6.0) + _closed(st, 6.3))
    assert np.max(np.abs(f[:, 0] / exact - 1.0)) < 3e-5
    # one temperature everywhere: the invariant integrator's value, to rounding
    g, h = _general(st, hr, N_L=32, invariant=True)
    assert np.max(np.abs(g / h - 1.0)) < 1e-11


def test_general_integrator_per_cell_correction_pole_on():
    st, hr = _flat_cap(48)

    def corr(m):
        C = np.ones_like(m["params"])
        C[..., 0] = 6.0
        return C

    f, f_inv = _general(st, hr, correction=corr, N_L=32, invariant=True)
    exact = _closed(st, 6.2) - _closed(st, 6.0)
    assert np.max(np.abs(f[:, 0] / exact - 1.0)) < 3e-5
    assert np.max(np.abs(f / f_inv - 1.0)) < 1e-11


def test_general_integrator_converges_to_invariant_inclined():
    st = O.spacetime(1.4, 12.0, 1.0, math.cos(1.0), 300.0, "AGM_14")
    hr = dict(super_colatitude=1.2, super_radius=0.8, super_temperature=6.3, phase_shift=0.0,
              sqrt_num_cells=32, min_sqrt_num_cells=10, max_sqrt_num_cells=100, num_rays=512,
              num_leaves=64)
    errs = []
    for N_L in (64, 128):
        f, f_inv = _general(st, hr, N_L=N_L, invariant=True)
        errs.append(np.max(np.abs(f - f_inv)) / np.max(np.abs(f_inv)))
    assert errs[0] < 2e-6 and errs[1] < errs[0] / 3.0


def test_general_integrator_rolled_map_shifts_phase():
    st = O.spacetime(1.4, 12.0, 1.0, math.cos(1.0), 300.0, "AGM_14")
    hr = dict(super_colatitude=0.0, super_radius=0.6, super_temperature=6.2, phase_shift=0.0,
              sqrt_num_cells=24, min_sqrt_num_cells=10, max_sqrt_num_cells=100, num_rays=512,
              num_leaves=64)
    members, _ = O.hot_region_embed(st, hr, st["f"])
    S = members[0]["phi"].shape[1]
    dphi = 2.0 * math.pi / S
    assert np.allclose(np.diff(members[0]["phi"], axis=1), dphi)
    phases = dphi * np.arange(S + 1)

    def sector(m, roll=0):
        P = m["params"].copy()
        P[:, :, 0] = 6.0
        P[:, S // 3: S // 3 + 4, 0] = 6.35
        return np.roll(P, roll, axis=1)

    f, _ = _general(st, hr, params=sector, phases=phases)
    g, _ = _general(st, hr, params=lambda m: sector(m, 1), phases=phases)
    # the map is not axisymmetric: the pulse has a shape to shift
    assert np.max(f) / np.min(f) > 1.05
    assert np.max(np.abs(g[:, :-1] - f[:, 1:])) < 1e-10 * np.max(np.abs(f))
    # and where the sector sits: the invariant integrator restricted to the
    # sector's cells at its temperature plus the rest at theirs
    m = members[0]
    energies = np.array([0.5, 1.0, 2.0, 4.0])
    leaves = np.linspace(0.0, 2.0 * math.pi, 64)
    geo = (m["cellArea"], m["r"], m["r_s_over_r"], m["theta"], m["phi"])
    rays = (m["deflection"], m["cos_alpha"], m["lag"], m["maxDeflection"], m["cos_gamma"])
    hot = np.zeros_like(m["radiates"])
    hot[:, S // 3: S // 3 + 4] = 1
    parts = []
    for logT, mask in ((6.35, hot), (6.0, 1 - hot)):
        P = m["params"].copy()
        P[..., 0] = logT
        t, fl = O.integrate(st["Omega"], st["r_s"], st["i"], *geo, P, mask * m["radiates"], *rays,
                            energies, leaves, phases, R=st["R"])
        assert t == 0
        parts.append(fl)
    ref = parts[0] + parts[1]
    assert np.max(np.abs(f - ref)) < 5e-6 * np.max(np.abs(ref))


# ---------------------------------------------------------------- Stokes and options
def _stokes(st, m, params=None, atm=1, beam=0, R_in=1e6, N_L=64, phases=None, stokes=True):
    energies = np.array([0.5, 1.0, 2.0, 4.0])
    leaves = np.linspace(0.0, 2.0 * math.pi, N_L)
    if phases is None:
        phases = 2.0 * math.pi * np.linspace(0.0, 1.0, N_L)
    P = m["params"] if params is None else params
    return O.integrate_stokes(st["Omega"], st["i"], m["cellArea"], m["r"], m["r_s_over_r"], m["theta"],
                              m["phi"], P, m["radiates"], m["deflection"], m["cos_alpha"], m["lag"],
                              m["maxDeflection"], m["cos_gamma"], energies, leaves, phases, atm_ext=atm,
                              beam_opt=beam, R_in=R_in, stokes=stokes)


def _member(st, **kw):
    hr = dict(super_temperature=6.3, phase_shift=0.0, sqrt_num_cells=32, min_sqrt_num_cells=10,
              max_sqrt_num_cells=100, num_rays=512, num_leaves=64)
    hr.update(kw)
    members, err = O.hot_region_embed(st, hr, st["f"])
    assert err is None
    return members[0]


def test_pol_angle_non_relativistic_limit():
    rng = np.random.default_rng(1234)
    for _ in range(10):
        th = rng.uniform(0.001, math.pi - 0.001)
        ph = rng.uniform(0.0, 2.0 * math.pi)
        inc = rng.uniform(0.001, math.pi - 0.001)
        chi = O.pol_angle(ph, 0.3, math.sqrt(1 - 0.09), 0.4, math.sqrt(1 - 0.16), math.sin(th), math.cos(th),
                          math.sin(inc), math.cos(inc), 0.0, 1.0, 1.0, 0.5, 1.0, 0.0, 1.0, 0.5)
        ref = math.atan2(-math.sin(th) * math.sin(ph),
                         math.sin(inc) * math.cos(th) - math.cos(inc) * math.sin(th) * math.cos(ph))
        assert abs(chi - ref) < 1e-12
    assert math.isfinite(O.pol_angle(0.1, 0.2, 0.98, 0.3, 0.95, 0.4, 0.92, 0.5, 0.86, 0.6, 0.8, 1.0, 0.7, 1.0,
                                     0.2, 1.1, 0.9))


def test_disk_blocks_southern_rays_inside_the_star():
    rng = np.random.default_rng(1234)
    for _ in range(10):
        # the reference test's draws, in its order
        radius = rng.uniform(5.0, 16.0)
        R_in = rng.uniform(0.1, radius * 0.999)
        inc = rng.uniform(0.001, math.pi / 2 - 0.001)
        psi = rng.uniform(0.001, math.pi - 0.001)
        th = rng.uniform(math.pi / 2 + 0.01, math.pi - 0.01)
        sin_alpha = rng.uniform(0.1, 0.9)
        r_s_over_r = rng.uniform(0.01, 0.2)
        args = (R_in, math.cos(inc), math.cos(psi), math.cos(th), r_s_over_r, radius, sin_alpha)
        assert O.disk_block(*args, th / math.pi) == 0
        # the northern hemisphere is never blocked
        assert O.disk_block(*args, (math.pi - th) / math.pi) == 1


def test_stokes_blackbody_is_the_plain_integrator():
    st = O.spacetime(1.4, 12.0, 1.0, math.cos(1.0), 300.0, "AGM_14")
    m = _member(st, super_colatitude=1.2, super_radius=0.8)
    t, I, Q, U = _stokes(st, m)
    assert t == 0
    energies = np.array([0.5, 1.0, 2.0, 4.0])
    leaves = np.linspace(0.0, 2.0 * math.pi, 64)
    phases = 2.0 * math.pi * np.linspace(0.0, 1.0, 64)
    t2, ref = O.integrate(st["Omega"], st["r_s"], st["i"], m["cellArea"], m["r"], m["r_s_over_r"], m["theta"],
                          m["phi"], m["params"], m["radiates"], m["deflection"], m["cos_alpha"], m["lag"],
                          m["maxDeflection"], m["cos_gamma"], energies, leaves, phases, R=st["R"])
    assert t2 == 0
    assert np.array_equal(I, ref)
    assert not np.any(Q) and not np.any(U)


def test_burst_and_beamed_flat_pole_on_closed_forms():
    st, hr = _flat_cap(48)
    members, err = O.hot_region_embed(st, hr, st["f"])
    m = members[0]
    energies = np.array([0.5, 1.0, 2.0, 4.0])
    kT = K_B_OVER_KEV * 10.0 ** 6.2
    base = BB_NORM * energies ** 3 / np.expm1(energies / kT) * 2.0 * math.pi * st["R"] ** 2 / (energies * KEV)
    rho = 0.5
    mu1 = math.sin(rho) ** 2 / 2.0            # integral of mu over the cap / (2 pi R^2)
    mu2 = (1.0 - math.cos(rho) ** 3) / 3.0    # integral of mu^2
    t, I, Q, U = _stokes(st, m, atm=3, N_L=32)
    assert t == 0
    assert np.max(np.abs(I[:, 0] / (base * (0.421 * mu1 + 0.868 * mu2)) - 1.0)) < 3e-5
    assert np.all(np.abs(I / I[:, :1] - 1.0) < 1e-9)
    assert np.max(np.abs(Q)) < 1e-8 * np.max(I) and np.max(np.abs(U)) < 1e-8 * np.max(I)
    P = np.zeros(m["params"].shape[:2] + (7,))
    P[..., :2] = m["params"]
    P[..., 2] = 0.3          # a: 1 + a mu (c = 0, so E^c = 1)
    P[..., 6] = 100
    t, I1 = _stokes(st, m, params=P, beam=1, N_L=32, stokes=False)
    assert np.max(np.abs(I1[:, 0] / (base * (mu1 + 0.3 * mu2)) - 1.0)) < 3e-5
    t, I2 = _stokes(st, m, params=P, beam=2, N_L=32, stokes=False)
    anorm = 0.5 / (0.5 + 0.3 / 3.0)
    assert np.max(np.abs(I2[:, 0] / (anorm * base * (mu1 + 0.3 * mu2)) - 1.0)) < 3e-5
    # energy-dependent terms: 1 + a E^c mu + b E^d mu^2 (E' = E here: no redshift)
    mu3 = (1.0 - math.cos(rho) ** 4) / 4.0
    P[..., 3], P[..., 4], P[..., 5] = 0.4, 1.0, 0.5
    t, I1e = _stokes(st, m, params=P, beam=1, N_L=32, stokes=False)
    assert np.max(np.abs(I1e[:, 0] / (base * (mu1 + 0.3 * energies * mu2 + 0.4 * energies ** 0.5 * mu3)) - 1.0)) < 3e-5
    P[..., 3], P[..., 4], P[..., 5] = 0.0, 0.0, 0.0
    errs = []
    for n in (10, 100, 1000):
        P[..., 6] = n
        t, I3 = _stokes(st, m, params=P, beam=3, N_L=32, stokes=False)
        errs.append(np.max(np.abs(I3 / I2 - 1.0)))
    assert errs[2] < 1e-6 and errs[1] < errs[0] / 50.0 and errs[2] < errs[1] / 50.0


def test_small_spot_polarisation_follows_pd_and_chi0():
    st = O.spacetime(1e-7, 12.0, 1.0, math.cos(1.0), 1e-9, "sphere")
    th0 = 1.1
    m = _member(st, super_colatitude=th0, super_radius=0.01, sqrt_num_cells=16, max_sqrt_num_cells=64,
                num_rays=400)
    ph = 2.0 * math.pi * np.linspace(0.0, 1.0, 64)
    t, I, Q, U = _stokes(st, m, atm=3, phases=ph)
    assert t == 0
    inc = st["i"]
    cpsi = math.cos(inc) * math.cos(th0) + math.sin(inc) * math.sin(th0) * np.cos(ph)
    PD = 0.1171 * (cpsi - 1.0) / (1.0 + 3.582 * cpsi)
    chi0 = np.arctan2(-math.sin(th0) * np.sin(ph),
                      math.sin(inc) * math.cos(th0) - math.cos(inc) * math.sin(th0) * np.cos(ph))
    good = cpsi > 0.3
    assert np.count_nonzero(good) > 20
    assert np.max(np.abs(Q[1][good] / I[1][good] - PD[good] * np.cos(2 * chi0[good]))) < 3e-5
    assert np.max(np.abs(U[1][good] / I[1][good] - PD[good] * np.sin(2 * chi0[good]))) < 3e-5


def test_disk_blocking_in_the_integrator():
    st = O.spacetime(1.4, 12.0, 1.0, math.cos(0.5), 300.0, "AGM_14")
    south = _member(st, super_colatitude=2.2, super_radius=0.4, sqrt_num_cells=24)
    t, I0 = _stokes(st, south, stokes=False)
    assert np.max(I0) > 0.0
    t, Ib = _stokes(st, south, R_in=5e3, stokes=False)
    assert t == 0 and not np.any(Ib)
    # more of the southern rays pass as the disk's inner edge recedes
    sums = [_stokes(st, south, R_in=r, stokes=False)[1].sum() for r in (1.3e4, 1.6e4, 2.0e4)]
    assert 0.0 < sums[0] < sums[1] <= sums[2] <= I0.sum() * (1 + 1e-12)
    north = _member(st, super_colatitude=0.8, super_radius=0.5)
    assert np.array_equal(_stokes(st, north, R_in=5e3, stokes=False)[1], _stokes(st, north, stokes=False)[1])


# ------------------------------------------------ general integrator: Stokes and options
def _gen_stokes(st, m, params=None, atm=1, beam=0, radiates=None, N_L=64, image_order_limit=None,
                stokes=True):
    energies = np.array([0.5, 1.0, 2.0, 4.0])
    leaves = np.linspace(0.0, 2.0 * math.pi, N_L)
    phases = 2.0 * math.pi * np.linspace(0.0, 1.0, N_L)
    P = m["params"] if params is None else params
    R = m["radiates"] if radiates is None else radiates
    return O.integrate_general_stokes(st["Omega"], st["i"], m["cellArea"], m["r"], m["r_s_over_r"],
                                      m["theta"], m["phi"], P, R, m["deflection"], m["cos_alpha"], m["lag"],
                                      m["maxDeflection"], m["cos_gamma"], energies, leaves, phases, atm_ext=atm,
                                      beam_opt=beam, image_order_limit=image_order_limit, stokes=stokes)


def test_general_stokes_blackbody_is_the_general_integrator():
    st = O.spacetime(1.4, 12.0, 1.0, math.cos(1.0), 300.0, "AGM_14")
    m = _member(st, super_colatitude=1.2, super_radius=0.8)
    P = m["params"].copy()
    P[..., 0] += 0.1 * np.sin(m["phi"])
    t, I, Q, U = _gen_stokes(st, m, params=P)
    assert t == 0
    f, _ = _general(st, dict(super_colatitude=1.2, super_radius=0.8, super_temperature=6.3, phase_shift=0.0,
                             sqrt_num_cells=32, min_sqrt_num_cells=10, max_sqrt_num_cells=100, num_rays=512,
                             num_leaves=64), params=lambda mm: P)
    assert np.array_equal(I, f)
    assert not np.any(Q) and not np.any(U)


def test_general_burst_and_beamed_pole_on_closed_forms():
    st, hr = _flat_cap(48)
    members, err = O.hot_region_embed(st, hr, st["f"])
    m = members[0]
    energies = np.array([0.5, 1.0, 2.0, 4.0])
    kT = K_B_OVER_KEV * 10.0 ** 6.2
    base = BB_NORM * energies ** 3 / np.expm1(energies / kT) * 2.0 * math.pi * st["R"] ** 2 / (energies * KEV)
    rho = 0.5
    mu1, mu2 = math.sin(rho) ** 2 / 2.0, (1.0 - math.cos(rho) ** 3) / 3.0
    t, I, Q, U = _gen_stokes(st, m, atm=3, N_L=32)
    assert t == 0
    assert np.max(np.abs(I[:, 0] / (base * (0.421 * mu1 + 0.868 * mu2)) - 1.0)) < 3e-5
    P = np.zeros(m["params"].shape[:2] + (7,))
    P[..., :2] = m["params"]
    P[..., 2], P[..., 6] = 0.3, 100
    t, I1 = _gen_stokes(st, m, params=P, beam=1, N_L=32, stokes=False)
    assert np.max(np.abs(I1[:, 0] / (base * (mu1 + 0.3 * mu2)) - 1.0)) < 3e-5


def test_general_stokes_one_angle_per_parallel_and_image():
    # one radiating parallel of an inclined, rotating star: U / Q the same everywhere
    st = O.spacetime(1.4, 12.0, 1.0, math.cos(1.0), 300.0, "AGM_14")
    m = _member(st, super_colatitude=1.0, super_radius=0.6)
    rad = np.zeros_like(m["radiates"])
    row = int(np.argmax(m["radiates"].sum(axis=1)))
    rad[row] = m["radiates"][row]
    t, I, Q, U = _gen_stokes(st, m, atm=3, radiates=rad, image_order_limit=1)
    assert t == 0
    big = np.abs(Q) > 1e-3 * np.max(np.abs(Q))
    ratio = U[big] / Q[big]
    assert np.max(np.abs(ratio - ratio[0])) < 1e-9 * max(1.0, abs(ratio[0]))
    # pole-on in flat space: mu eta = cos theta on the parallel, so |P| = |PD| I
    st0, hr0 = _flat_cap(32)
    m0 = O.hot_region_embed(st0, hr0, st0["f"])[0][0]
    rad0 = np.zeros_like(m0["radiates"])
    r0 = m0["radiates"].shape[0] // 2
    rad0[r0] = m0["radiates"][r0]
    t, I0, Q0, U0 = _gen_stokes(st0, m0, atm=3, radiates=rad0, N_L=32, image_order_limit=1)
    assert t == 0
    c = math.cos(m0["theta"][r0, 0])
    PD = 0.1171 * (c - 1.0) / (1.0 + 3.582 * c)
    assert np.max(np.abs(np.hypot(Q0, U0) / (abs(PD) * I0) - 1.0)) < 1e-6
