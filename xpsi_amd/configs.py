"""Model descriptions of the benchmark workloads.

config1: one circular blackbody spot on an oblate (AGM_14) star, M = 1.4 Msun,
R_eq = 12 km, 300 Hz, inclination 60 deg, colatitude 45 deg, radius 0.4 rad,
log10 T = 6.1; sqrt_num_cells 32, 200 rays, 64 leaves, 128 energies over
0.3-10 keV, 32 phase bins, 100 channels with Gaussian redistribution
(sigma = 5% E) and an effective-area bump at 1 keV; Poisson counts for 1e6 s
plus 5 background counts per channel and bin (numpy default_rng(42)).

The response recipe (assumptions recorded alongside the data):
  input bins: edges logspace(log10 0.3, log10 10, N_in + 1) keV;
  channels: edges logspace(log10 0.3, log10 10, N_ch + 1) keV;
  M[c, e] = A(E_e) * (Phi((hi_c - E_e)/(0.05 E_e)) - Phi((lo_c - E_e)/(0.05 E_e)))
  with E_e the input-bin midpoint and A(E) = 1000 cm^2 exp(-(log10 E)^2 / (2 * 0.25^2)).
"""
import math
import os

import numpy as np
from scipy.special import ndtr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def gaussian_response(N_ch, N_in, lo=0.3, hi=10.0, sigma_frac=0.05, area=1000.0, width=0.25):
    in_edges = np.logspace(np.log10(lo), np.log10(hi), N_in + 1)
    ch_edges = np.logspace(np.log10(lo), np.log10(hi), N_ch + 1)
    Ee = 0.5 * (in_edges[:-1] + in_edges[1:])
    A = area * np.exp(-np.log10(Ee) ** 2 / (2.0 * width ** 2))
    sig = sigma_frac * Ee
    M = A[None, :] * (ndtr((ch_edges[1:, None] - Ee[None, :]) / sig[None, :])
                      - ndtr((ch_edges[:-1, None] - Ee[None, :]) / sig[None, :]))
    return np.ascontiguousarray(M), in_edges, ch_edges


def config1(counts=None):
    M, in_edges, _ = gaussian_response(100, 128)
    if counts is None:
        path = os.path.join(GOLDEN, "config1.npz")
        counts = np.load(path)["counts"] if os.path.exists(path) else np.zeros((100, 32))
    return dict(
        spacetime=dict(mass=1.4, radius=12.0, distance=1.0, cos_inclination=0.5,
                       frequency=300.0, star_shape="AGM_14"),
        hot_regions=[dict(super_colatitude=math.pi / 4.0, super_radius=0.4,
                          super_temperature=6.1, phase_shift=0.0, sqrt_num_cells=32,
                          min_sqrt_num_cells=10, max_sqrt_num_cells=80, num_rays=200,
                          num_leaves=64, is_antiphased=False, image_order_limit=None)],
        instrument=dict(matrix=M, energy_edges=in_edges),
        data=dict(counts=counts, phases=np.linspace(0.0, 1.0, 33), exposure_time=1.0e6),
        num_energies=128, support=None,
        marginal=dict(epsabs=0.0, epsrel=1.0e-8, epsilon=1.0e-3, sigmas=10.0),
        llzero=-1.0e90)


def prior_accepts(mass, radius, frequency):
    """Compactness and shape cuts of examples/examples_fast/Modules/CustomPrior.py:49-66
    (with the spacetime derived quantities of xpsi/Spacetime.py:111-190)."""
    GM = 0.5 * 2.95325024e3
    c, G = 2.99792458e8, 6.6740831e-11
    if not radius <= 16.0:
        return False
    r_g = mass * GM
    M = mass * GM * c * c / G
    R = radius * 1.0e3
    Omega = 2.0 * math.pi * frequency
    zeta = r_g / R
    eps = Omega ** 2.0 * R ** 3.0 / (G * M)
    R_r_s = R / (2.0 * r_g)
    R_p = 1.0 + eps * (-0.788 + 1.030 * zeta)
    if R_p < 1.505 / R_r_s:
        return False
    arg = -1.0 / (3.0 * eps * (-0.788 + 1.030 * zeta))
    if arg >= 0.0 and math.sqrt(arg) < 1.0:
        return False
    return True


def config3_draws(n=4096, seed=3):
    """Parameter vectors of the batched workload: uniform prior draws
    (M 1-3, R 8-16 km, cos i 0-1, colatitude 0-pi, radius 0.05-1 rad,
    log T 5.5-6.7, phase U[0,1), distance 1 kpc), sqrt_num_cells from
    {24, 32, 48, 64, 96} and num_energies from {64, 128, 256}; drawn with
    numpy.random.default_rng(seed) in that order."""
    rng = np.random.default_rng(seed)
    M = rng.uniform(1.0, 3.0, n)
    R = rng.uniform(8.0, 16.0, n)
    ci = rng.uniform(0.0, 1.0, n)
    colat = rng.uniform(0.0, math.pi, n)
    rho = rng.uniform(0.05, 1.0, n)
    logT = rng.uniform(5.5, 6.7, n)
    phase = rng.uniform(0.0, 1.0, n)
    snc = rng.choice([24, 32, 48, 64, 96], n)
    ne = rng.choice([64, 128, 256], n)
    out = []
    for k in range(n):
        out.append(dict(mass=M[k], radius=R[k], distance=1.0, cos_inclination=ci[k],
                        num_energies=int(ne[k]),
                        hot_regions=[dict(phase_shift=phase[k], super_colatitude=colat[k],
                                          super_radius=rho[k], super_temperature=logT[k],
                                          sqrt_num_cells=int(snc[k]))]))
    return out


def config3_model(num_energies, counts=None):
    """Model of the batched workload for one energy count: config-1 response,
    data and atmosphere; mesh bound raised to 128 cells per side."""
    m = config1(counts)
    m["num_energies"] = int(num_energies)
    h = m["hot_regions"][0]
    h["max_sqrt_num_cells"] = 128
    h["image_order_limit"] = None
    return m


def num4d_table(seed=2):
    """Synthetic four-dimensional atmosphere table (hot_Num4D.pyx layout):
    axes log10 T (35, 5.1-6.8), log10 g (14, 13.7-15.0), mu (67, k/67 for
    k = 1..67) and log10(E / kT) (166, -1.3-2.0); values I_nu / T^3 (cgs per
    Hz) of a blackbody times (1 + 0.5 mu) times (1 + 0.05 n), with n a sum of
    four low-frequency sinusoids over the normalised grid coordinates
    (numpy default_rng(seed) draws their integer frequencies in {1, 2} and
    phases), scaled so that max |n| <= 1.  Flattened C order (T, g, mu, E)."""
    logT = np.linspace(5.1, 6.8, 35)
    logg = np.linspace(13.7, 15.0, 14)
    mu = np.arange(1, 68) / 67.0
    logE = np.linspace(-1.3, 2.0, 166)
    rng = np.random.default_rng(seed)
    freqs = rng.integers(1, 3, size=(4, 4))
    phs = rng.uniform(0.0, 2.0 * math.pi, size=4)
    uT = np.linspace(0, 1, 35)[:, None, None, None]
    ug = np.linspace(0, 1, 14)[None, :, None, None]
    um = np.linspace(0, 1, 67)[None, None, :, None]
    ue = np.linspace(0, 1, 166)[None, None, None, :]
    n = np.zeros((35, 14, 67, 166))
    for q in range(4):
        f = freqs[q]
        n += np.sin(2.0 * math.pi * (f[0] * uT + f[1] * ug + f[2] * um + f[3] * ue) + phs[q])
    n /= 4.0
    k_cgs, h_cgs, c_cgs = 1.380649e-16, 6.62607015e-27, 2.99792458e10
    x = 10.0 ** logE
    planck = 2.0 * k_cgs ** 3 / (c_cgs ** 2 * h_cgs ** 2) * x ** 3 / np.expm1(x)
    buf = planck[None, None, None, :] * (1.0 + 0.5 * mu)[None, None, :, None] * (1.0 + 0.05 * n)
    return (logT, logg, mu, logE, np.ascontiguousarray(buf.reshape(-1)))


def config2(atmosphere="Num4D", counts=None):
    """Primary hot region: superseding spot with a concentric ceding ring;
    secondary: the same region antipodal (colatitude pi - theta, half a cycle
    later).  theta 0.9, super radius 0.15, cede radius 0.35, log T 6.3 (super)
    and 6.0 (cede); sqrt_num_cells 64 (max 128), 128 leaves, 256 energies over
    0.3-10 keV, 300 channels x 256 inputs, 32 phase bins, 1e6 s."""
    M, in_edges, _ = gaussian_response(300, 256)
    atm_ext = 2 if atmosphere == "Num4D" else 1
    name = "config2_%s.npz" % atmosphere
    if counts is None:
        path = os.path.join(GOLDEN, name)
        counts = np.load(path)["counts"] if os.path.exists(path) else np.zeros((300, 32))
    regions = []
    for k in range(2):
        regions.append(dict(super_colatitude=0.9 if k == 0 else math.pi - 0.9,
                            super_radius=0.15, cede_radius=0.35, cede_colatitude=None,
                            cede_azimuth=0.0, super_temperature=6.3, cede_temperature=6.0,
                            phase_shift=0.0, cede=True, sqrt_num_cells=64,
                            min_sqrt_num_cells=10, max_sqrt_num_cells=128, num_rays=200,
                            num_leaves=128, is_antiphased=(k == 1), image_order_limit=None,
                            atm_ext=atm_ext))
    for r in regions:
        r["cede_colatitude"] = r["super_colatitude"]
    return dict(
        spacetime=dict(mass=1.4, radius=12.0, distance=1.0, cos_inclination=0.5,
                       frequency=300.0, star_shape="AGM_14"),
        hot_regions=regions,
        instrument=dict(matrix=M, energy_edges=in_edges),
        data=dict(counts=counts, phases=np.linspace(0.0, 1.0, 33), exposure_time=1.0e6),
        num_energies=256, support=None,
        atmosphere=num4d_table() if atm_ext == 2 else None,
        marginal=dict(epsabs=0.0, epsrel=1.0e-8, epsilon=1.0e-3, sigmas=10.0),
        llzero=-1.0e90)


def interstellar(N_H=1.0e21, sigma0=2.0e-22, index=-2.5):
    """Interstellar attenuation exp(-N_H sigma(E)), sigma = sigma0 (E/keV)^index
    cm^2, evaluated at the input-bin midpoints (xpsi/Signal.py:435)."""
    def attenuation(E):
        return np.exp(-N_H * sigma0 * np.asarray(E, dtype=np.float64) ** index)
    return attenuation


def config5(scale="full", counts=None):
    """Two hot regions of the config-2 geometry with blackbody emission (a
    superseding spot with a concentric ceding ring, and its antipode half a
    cycle later), the rest of the surface ("elsewhere") at log T 5.8 on a
    closed mesh, and interstellar absorption with N_H = 1e21 cm^-2.

    scale "full": sqrt_num_cells 256 (max 320), 256 leaves, 512 energies,
    elsewhere 64 cells per side and 1000 rays, 300 channels x 256 inputs,
    64 phase bins.  scale "test": sqrt_num_cells 16, 32 leaves, 64 energies,
    elsewhere 16 cells per side and 200 rays, 60 channels x 64 inputs, 16 bins
    (parity-test size)."""
    full = scale == "full"
    n_ch, n_in = (300, 256) if full else (60, 64)
    M, in_edges, _ = gaussian_response(n_ch, n_in)
    n_b = 64 if full else 16
    if counts is None:
        path = os.path.join(GOLDEN, "config5_%s.npz" % scale)
        counts = np.load(path)["counts"] if os.path.exists(path) else np.zeros((n_ch, n_b))
    regions = []
    for k in range(2):
        th = 0.9 if k == 0 else math.pi - 0.9
        regions.append(dict(super_colatitude=th, super_radius=0.15, cede_radius=0.35,
                            cede_colatitude=th, cede_azimuth=0.0, super_temperature=6.3,
                            cede_temperature=6.0, phase_shift=0.0, cede=True,
                            sqrt_num_cells=256 if full else 16, min_sqrt_num_cells=10,
                            max_sqrt_num_cells=320 if full else 32, num_rays=200 if full else 100,
                            num_leaves=256 if full else 32, is_antiphased=(k == 1),
                            image_order_limit=None, atm_ext=1))
    return dict(
        spacetime=dict(mass=1.4, radius=12.0, distance=1.0, cos_inclination=0.5,
                       frequency=300.0, star_shape="AGM_14"),
        hot_regions=regions,
        elsewhere=dict(sqrt_num_cells=64 if full else 16, num_rays=1000 if full else 200,
                       image_order_limit=None, elsewhere_temperature=5.8),
        interstellar=interstellar(),
        instrument=dict(matrix=M, energy_edges=in_edges),
        data=dict(counts=counts, phases=np.linspace(0.0, 1.0, n_b + 1), exposure_time=1.0e6),
        num_energies=512 if full else 64, support=None,
        marginal=dict(epsabs=0.0, epsrel=1.0e-8, epsilon=1.0e-3, sigmas=10.0),
        llzero=-1.0e90)


def config5_draws(n=1024, seed=5):
    """Parameter vectors of the large workload: uniform draws of M 1-3,
    R 8-16 km, cos i 0-1, primary colatitude 0-pi (the secondary region is
    its antipode), super radius 0.05-0.5 rad, cede radius super + 0.05-0.5,
    super log T 6.0-6.7, cede log T 5.9-6.5, phase U[0,1); elsewhere at
    log T 5.8, distance 1 kpc; numpy.random.default_rng(seed) in that order."""
    rng = np.random.default_rng(seed)
    M = rng.uniform(1.0, 3.0, n)
    R = rng.uniform(8.0, 16.0, n)
    ci = rng.uniform(0.0, 1.0, n)
    th = rng.uniform(0.0, math.pi, n)
    rs = rng.uniform(0.05, 0.5, n)
    rc = rs + rng.uniform(0.05, 0.5, n)
    Ts = rng.uniform(6.0, 6.7, n)
    Tc = rng.uniform(5.9, 6.5, n)
    ph = rng.uniform(0.0, 1.0, n)
    out = []
    for k in range(n):
        regs = []
        for g in range(2):
            t = th[k] if g == 0 else math.pi - th[k]
            regs.append(dict(phase_shift=ph[k], super_colatitude=t, super_radius=rs[k],
                             super_temperature=Ts[k], cede_colatitude=t, cede_radius=rc[k],
                             cede_azimuth=0.0, cede_temperature=Tc[k]))
        out.append(dict(mass=M[k], radius=R[k], distance=1.0, cos_inclination=ci[k],
                        hot_regions=regs, elsewhere_temperature=5.8))
    return out


def config4_model(atmosphere="Num4D"):
    """Model of the serving workload: the config-2 model (two regions with
    ceding rings, 4-D table atmosphere, 256 energies, 300 channels, 128
    leaves); requests carry their own mesh resolution."""
    return config2(atmosphere)


def config4_draws(n=512, seed=4):
    """Request stream of the serving workload: uniform draws of M 1-3,
    R 8-16 km, cos i 0-1, primary colatitude 0-pi (the secondary region is
    its antipode), super radius 0.05-0.3 rad, cede radius super + 0.05-0.3,
    super log T 5.9-6.35, cede log T 5.8-6.1 (inside the table's range),
    phase U[0,1), distance 1 kpc, and sqrt_num_cells from {32, 48, 64, 96}
    for both regions; numpy.random.default_rng(seed) in that order."""
    rng = np.random.default_rng(seed)
    M = rng.uniform(1.0, 3.0, n)
    R = rng.uniform(8.0, 16.0, n)
    ci = rng.uniform(0.0, 1.0, n)
    th = rng.uniform(0.0, math.pi, n)
    rs = rng.uniform(0.05, 0.3, n)
    rc = rs + rng.uniform(0.05, 0.3, n)
    Ts = rng.uniform(5.9, 6.35, n)
    Tc = rng.uniform(5.8, 6.1, n)
    ph = rng.uniform(0.0, 1.0, n)
    snc = rng.choice([32, 48, 64, 96], n)
    out = []
    for k in range(n):
        regs = []
        for g in range(2):
            t = th[k] if g == 0 else math.pi - th[k]
            regs.append(dict(phase_shift=ph[k], super_colatitude=t, super_radius=rs[k],
                             super_temperature=Ts[k], cede_colatitude=t, cede_radius=rc[k],
                             cede_azimuth=0.0, cede_temperature=Tc[k],
                             sqrt_num_cells=int(snc[k])))
        out.append(dict(mass=M[k], radius=R[k], distance=1.0, cos_inclination=ci[k],
                        hot_regions=regs))
    return out
