"""Pins of the oracle's flow model (or_flow_step, restating blackholes/src/emission.cuh:99-175) by an
independent evaluation and by physical identities.

The independent model below is written from the physics, not from the restatement: the Kerr metric
as a 4x4 Boyer-Lindquist tensor (lapse from its inverse), the fluid four-velocity u = u^t (1, 0, 0,
Omega) normalised through the tensor, photon and field four-vectors contracted with it, K2 from
scipy (Leung et al. 2011 emissivity via oracle/mlp_ref.leung_lnf_scipy), the Planck function, and
the Mino-time path length. The flow's profiles (density, temperature, field strength, Keplerian
rotation) are the reference's model definitions (emission.cuh:100-147) and are stated once here.

What each check catches:
- test_against_independent_model: (dI, dtau) of both oracle builds against the model, to 3e-4
  with scipy's K2 (the float K2 table's interpolation error) and to 1e-10 with the same float
  table. A wrong metric component, a sign in k.u, a transposed pitch-angle projection, a wrong
  constant or a dropped factor fails by orders of magnitude.
- test_kirchhoff_and_thin_limit: dI / dtau -> g^3 B_nu(T_e, nu_f) as n0 -> 0 (Kirchhoff's law with
  the transfer quadrature dI = g^3 B_nu (1 - e^-tau)), with B_nu from the Planck formula alone.
- test_linear_in_step: dtau is exactly proportional to the Mino-time step h.
- test_turbulence_scales_density: a uniform turbulence factor T equals multiplying n0 by T (the
  factor multiplies the electron density, not the temperature or the field).
- test_velocity_clamp: near the horizon the Keplerian rate is clamped to 0.999 of the timelike
  bound |Omega - omega| < lapse / sqrt(g_phph); the points used exercise the clamp.
- test_early_outs: each early return of emission.cuh:106,109,131 gives exactly (0, 0), and points
  just inside each bound do not."""
import numpy as np
import pytest
from scipy.special import kve

from helpers import setup, shot
from mlp_ref import leung_lnf_scipy, table_lerp_f32
from oracle_ctypes import Oracle

# CGS constants of the model (blackholes/src/common.hpp:17-26)
C_C, C_MP, C_ME, C_QE, C_KB, C_H = 2.99792458e10, 1.67262192e-24, 9.1093837e-28, 4.80320471e-10, 1.380649e-16, \
    6.62607015e-27


@pytest.fixture(scope="module")
def scene():
    cfg = setup(shot("c2_radio_1024.ini"), with_grid=False)
    return cfg


def metric(a, r, u):
    """Covariant Kerr metric in Boyer-Lindquist (t, r, theta, phi), u = cos(theta)."""
    s2 = 1.0 - u * u
    sig = r * r + a * a * u * u
    dlt = r * r - 2.0 * r + a * a
    g = np.zeros((4, 4))
    g[0, 0] = -(1.0 - 2.0 * r / sig)
    g[0, 3] = g[3, 0] = -2.0 * a * r * s2 / sig
    g[3, 3] = (r * r + a * a + 2.0 * a * a * r * s2 / sig) * s2
    g[1, 1] = sig / dlt
    g[2, 2] = sig
    return g, sig


def planck(nu, T):
    return 2.0 * C_H * nu ** 3 / C_C ** 2 / np.expm1(C_H * nu / (C_KB * T))


def model(P, r, u, lam, E, h, k2_table=None, n0=None, clamp_info=None, pitch=None):
    """(dI, dtau) of one Mino-time step h at (r, u = cos theta) for a photon of conserved
    lambda = L/E, with E the camera normalisation: the photon momentum normalised to unit conserved
    energy is p = (-1, ., ., lambda) (covariant), and nu_fluid / nu_obs = E (-p.u_fluid)."""
    a = P["a"]
    n0 = P["n0"] if n0 is None else n0
    s2 = max(1.0 - u * u, 1e-12)
    R, z = r * np.sqrt(s2), r * u
    if r > P["r_flow_max"] or R < 1e-3:
        return 0.0, 0.0
    vert = np.exp(-0.5 * (z / (P["h_r"] * R)) ** 2)
    if vert < 1e-7:
        return 0.0, 0.0
    ne = n0 * r ** -P["n_index"] * vert * (1.0 - np.exp(-(r - P["rh"]) ** 2 / 0.25))
    if ne < 1e-30:
        return 0.0, 0.0
    theta_e = P["theta0"] * r ** -P["theta_index"]
    # field strength from plasma beta against the model's ion pressure n_e m_p c^2 / (6 r)
    B = np.sqrt(8.0 * np.pi * ne * C_MP * C_C ** 2 / (6.0 * r * P["beta_p"]))
    g, sig = metric(a, r, u)
    ginv = np.linalg.inv(g)
    lapse = np.sqrt(-1.0 / ginv[0, 0])
    omega = -g[0, 3] / g[3, 3]
    vbound = lapse / np.sqrt(g[3, 3])
    om = P["flow_kep"] / (max(R, P["r_isco"]) ** 1.5 + a)
    om_c = omega + np.clip(om - omega, -0.999 * vbound, 0.999 * vbound)
    if clamp_info is not None:
        clamp_info.append(om_c != om)
    U = np.array([1.0, 0.0, 0.0, om_c])
    nrm = U @ g @ U
    if nrm >= 0.0:
        return 0.0, 0.0
    U = U / np.sqrt(-nrm)                      # u.u = -1
    p = np.array([-1.0, 0.0, 0.0, lam])        # covariant; r and theta parts do not enter below
    e_f = -(p @ U)                             # photon energy in the fluid frame (per unit conserved energy)
    nu_f = P["nu_obs"] * E * e_f
    gshift = P["nu_obs"] / nu_f
    u_cov = g @ U
    b = np.array([0.0, 0.0, 0.0, 1.0]) + u_cov[3] * U     # toroidal field: e_phi projected orthogonal to u
    cos_b = (p @ b) / (np.sqrt(b @ g @ b) * e_f)
    sin_b = np.sqrt(max(1.0 - cos_b * cos_b, 1e-6))
    assert e_f > 0.0
    if pitch is not None:
        pitch.append(sin_b)
    nu_c = C_QE * B / (2.0 * np.pi * C_ME * C_C)
    if k2_table is None:
        lnfj = leung_lnf_scipy(np.log10(nu_f / nu_c), theta_e, np.arcsin(sin_b))
    else:
        tab, l0, dl = k2_table
        lk2 = float(table_lerp_f32(tab, l0, dl, np.float32(np.log(theta_e))))
        lk2_true = np.log(kve(2, 1.0 / theta_e)) - 1.0 / theta_e
        lnfj = leung_lnf_scipy(np.log10(nu_f / nu_c), theta_e, np.arcsin(sin_b)) + lk2_true - lk2
    j = ne * C_QE ** 2 * nu_c / C_C * np.exp(lnfj)
    Bnu = planck(nu_f, theta_e * C_ME * C_C ** 2 / C_KB)
    dl = e_f * sig * h * P["rg_cm"]            # proper length in the fluid frame: (-p.u) d(affine), d(affine) = Sigma d(Mino)
    tau = j / Bnu * dl
    return gshift ** 3 * Bnu * -np.expm1(-tau), tau


def params(cfg):
    S = cfg["OS"]
    return {n: getattr(S, n) for n in ("a", "rh", "r_isco", "n0", "n_index", "theta_index", "theta0", "h_r", "beta_p",
                                        "r_flow_max", "nu_obs", "rg_cm", "flow_kep")}


def sample_points(rng, n, rh, near=False, a=0.94, near_width=0.6):
    """Points in the flow with photons that can be there: for lambda, the smallest Carter constant
    the polar potential allows at u (eta >= lambda^2 u^2 / (1 - u^2) - a^2 u^2, eta >= 0) must leave
    the radial potential (r^2 + a^2 - a lambda)^2 - Delta (eta + (lambda - a)^2) non-negative, on its
    future-directed branch r^2 + a^2 - a lambda > 0."""
    out = []
    while len(out) < n:
        r = rng.uniform(rh + 0.005, rh + near_width) if near else np.exp(rng.uniform(np.log(rh + 0.05), np.log(59.0)))
        u, lam = rng.uniform(-0.45, 0.45), rng.uniform(-7.0, 7.0)
        eta = max(0.0, lam * lam * u * u / (1.0 - u * u) - a * a * u * u)
        pr = r * r + a * a - a * lam
        if pr <= 0.0 or pr * pr - (r * r - 2 * r + a * a) * (eta + (lam - a) ** 2) < 0.0:
            continue
        out.append((r, u, rng.uniform(0, 2 * np.pi), rng.uniform(-2000, 0), lam,
                    rng.uniform(0.8, 1.25), np.exp(rng.uniform(np.log(1e-3), np.log(0.1)))))
    return tuple(np.array(c) for c in zip(*out))


@pytest.mark.parametrize("kind", ["libm", "bhm"])
def test_against_independent_model(scene, kind):
    O = Oracle(kind)
    P = params(scene)
    rng = np.random.default_rng(11)
    pts = list(zip(*sample_points(rng, 300, P["rh"]))) + list(zip(*sample_points(rng, 100, P["rh"], near=True)))
    table = (scene["k2"], scene["l0"], scene["dl"])
    worst_loose, worst_tight, nonzero = 0.0, 0.0, 0
    for r, u, phi, t, lam, E, h in pts:
        dI, dtau = O.flow_step(scene["OS"], r, u, phi, t, lam, E, h)
        sinb = []
        mI, mtau = model(P, r, u, lam, E, h, pitch=sinb)
        tI, ttau = model(P, r, u, lam, E, h, k2_table=table)
        if mtau == 0.0:
            assert dI == 0.0 and dtau == 0.0
            continue
        nonzero += 1
        for got, want_loose, want_tight in ((dI, mI, tI), (dtau, mtau, ttau)):
            worst_loose = max(worst_loose, abs(got / want_loose - 1.0))
            # sin(theta_B) from 1 - cos^2 loses digits as cos -> 1 (relative error ~ 1e-16 / sin^2)
            if sinb[0] > 0.1:
                worst_tight = max(worst_tight, abs(got / want_tight - 1.0))
    assert nonzero > 350
    assert worst_loose < 3e-4, f"{kind}: {worst_loose:.3e} against the scipy-K2 model"
    assert worst_tight < 1e-10, f"{kind}: {worst_tight:.3e} against the float-table model"


def test_kirchhoff_and_thin_limit(scene, oracle_libm):
    """dI / dtau = g^3 B_nu (1 - e^-tau) / tau -> g^3 B_nu as tau -> 0, with g and B_nu computed
    from the fluid velocity and the Planck formula alone (no emissivity)."""
    P = params(scene)
    rng = np.random.default_rng(12)
    S = scene["OS"]
    n0_saved = S.n0
    try:
        S.n0 = P["n0"] * 1e-9
        for r, u, phi, t, lam, E, h in zip(*sample_points(rng, 100, P["rh"])):
            dI, dtau = oracle_libm.flow_step(S, r, u, phi, t, lam, E, h)
            if dtau == 0.0:
                continue
            assert dtau < 1e-6
            # g and nu_f from the four-velocity, B_nu from Planck
            g, _ = metric(P["a"], r, u)
            ginv = np.linalg.inv(g)
            omega = -g[0, 3] / g[3, 3]
            vb = np.sqrt(-1.0 / ginv[0, 0]) / np.sqrt(g[3, 3])
            R = r * np.sqrt(1 - u * u)
            om = omega + np.clip(P["flow_kep"] / (max(R, P["r_isco"]) ** 1.5 + P["a"]) - omega, -0.999 * vb, 0.999 * vb)
            U = np.array([1.0, 0, 0, om])
            U = U / np.sqrt(-(U @ g @ U))
            nu_f = P["nu_obs"] * E * (U[0] - lam * U[3])
            Te = P["theta0"] * r ** -P["theta_index"] * C_ME * C_C ** 2 / C_KB
            want = (P["nu_obs"] / nu_f) ** 3 * planck(nu_f, Te)
            assert abs(dI / dtau / want - 1.0) < 1e-6 + dtau
    finally:
        S.n0 = n0_saved


def test_linear_in_step(scene, oracle_bhm):
    P = params(scene)
    rng = np.random.default_rng(13)
    for r, u, phi, t, lam, E, h in zip(*sample_points(rng, 50, P["rh"])):
        _, t1 = oracle_bhm.flow_step(scene["OS"], r, u, phi, t, lam, E, h)
        _, t2 = oracle_bhm.flow_step(scene["OS"], r, u, phi, t, lam, E, 2.0 * h)
        _, t4 = oracle_bhm.flow_step(scene["OS"], r, u, phi, t, lam, E, 0.25 * h)
        if t1 == 0.0:
            continue
        assert abs(t2 / (2.0 * t1) - 1.0) < 1e-15 and abs(4.0 * t4 / t1 - 1.0) < 1e-15


def test_turbulence_scales_density(scene, oracle_bhm):
    """A uniform turbulence grid of value T gives what n0 * T gives without turbulence."""
    from oracle_ctypes import make_scene
    from helpers import scene_dict
    T = 1.75   # exact in float
    grid = np.full((8, 8, 8), T, dtype=np.float32)
    d = scene_dict(scene["S"])
    S_turb = make_scene(d, k2=scene["k2"], grid=grid, turb_half=60.0)
    d2 = dict(d)
    d2["n0"] = d["n0"] * T
    d2["flow_turb_amp"] = 0.0
    S_scaled = make_scene(d2, k2=scene["k2"])
    for S in (S_turb, S_scaled):
        S.k2_l0, S.k2_dl = scene["l0"], scene["dl"]
    assert S_turb.flow_turb_amp > 0 and S_turb.turb_n == 8
    rng = np.random.default_rng(14)
    n = 0
    for r, u, phi, t, lam, E, h in zip(*sample_points(rng, 60, scene["OS"].rh)):
        a = oracle_bhm.flow_step(S_turb, r, u, phi, t, lam, E, h)
        b = oracle_bhm.flow_step(S_scaled, r, u, phi, t, lam, E, h)
        if b[1] == 0.0:
            continue
        n += 1
        assert abs(a[0] / b[0] - 1.0) < 1e-12 and abs(a[1] / b[1] - 1.0) < 1e-12
    assert n > 40


def test_velocity_clamp(scene, oracle_libm):
    """Near the horizon the Keplerian rate exceeds the timelike bound and is clamped; the oracle
    agrees with the model there (the clamp decides u^t, the Doppler factor and the pitch angle)."""
    P = params(scene)
    rng = np.random.default_rng(15)
    clamped = []
    table = (scene["k2"], scene["l0"], scene["dl"])
    S = scene["OS"]
    n0_saved = S.n0
    try:
        S.n0 = P["n0"] = n0_saved * 1e4   # a denser flow, so that the softened inner edge still emits
        for r, u, phi, t, lam, E, h in zip(*sample_points(rng, 300, P["rh"], near=True, near_width=0.1)):
            info, sinb = [], []
            mI, mtau = model(P, r, u, lam, E, h, k2_table=table, clamp_info=info, pitch=sinb)
            if not info or not info[0] or mtau == 0.0:
                continue
            clamped.append(r)
            dI, dtau = oracle_libm.flow_step(S, r, u, phi, t, lam, E, h)
            # g^3 B_nu = dI / (1 - e^-tau) depends on the clamped velocity through the Doppler factor
            # only. tau also depends on the pitch angle, whose sine comes from 1 - cos^2 (photons
            # near the horizon run almost along the toroidal field), and on u^t: at 0.999 of the
            # light-speed bound, 1 / sqrt(lapse^2 - g_phph (Omega - omega)^2) cancels ~500-fold,
            # and the Wien tail exp(-X^(1/3)) amplifies that again; a wrong clamp factor or sign
            # changes tau at order one.
            assert abs(dI / -np.expm1(-dtau) / (mI / -np.expm1(-mtau)) - 1.0) < 1e-10
            assert abs(dtau / mtau - 1.0) < 1e-6
    finally:
        S.n0 = n0_saved
    assert len(clamped) > 20, f"only {len(clamped)} sampled points exercised the clamp"


def test_early_outs(scene, oracle_bhm):
    P = params(scene)
    S = scene["OS"]
    f = oracle_bhm.flow_step
    args = (0.3, -500.0, 2.0, 1.0, 0.01)   # phi, t, lam, E, h
    # r beyond the flow's edge (emission.cuh:106)
    assert f(S, P["r_flow_max"] * (1 + 1e-12), 0.0, *args) == (0.0, 0.0)
    assert f(S, P["r_flow_max"] * (1 - 1e-12), 0.0, *args)[1] > 0.0
    # on the axis: cylindrical radius below 1e-3 (emission.cuh:106)
    u_axis = np.sqrt(1.0 - (0.9e-3 / 10.0) ** 2)
    assert f(S, 10.0, u_axis, *args) == (0.0, 0.0)
    # far above the mid-plane: exp(-zh^2 / 2) below 1e-7 (emission.cuh:109)
    zh_edge = np.sqrt(-2.0 * np.log(1e-7))
    u_out = np.cos(np.arctan(1.0 / (P["h_r"] * zh_edge * 1.001)))
    u_in = np.cos(np.arctan(1.0 / (P["h_r"] * zh_edge * 0.999)))
    assert f(S, 10.0, u_out, *args) == (0.0, 0.0)
    assert f(S, 10.0, u_in, *args)[1] > 0.0
    # electron density below 1e-30 (emission.cuh:131): without the guard a zero density (n0 = 0, or
    # r = r_+ where the inner softening vanishes) gives B = 0, X = inf and 0 * inf = NaN
    n0 = S.n0
    try:
        for S.n0 in (0.0, 1e-40):
            assert f(S, 10.0, 0.0, *args) == (0.0, 0.0)
        S.n0 = 1e-27
        assert all(np.isfinite(f(S, 10.0, 0.0, *args)))
    finally:
        S.n0 = n0
