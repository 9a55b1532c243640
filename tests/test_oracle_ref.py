"""Pins the CPU oracle to the reference's own compiled host code.

Golden vectors in tests/golden/ref_vectors.npz come from oracle/_ref/libref.so, which is the
reference's kerr.hpp / shot.cpp / scene_build.cpp / config.cpp compiled unmodified
(tests/golden/make_golden.py). The glibc build of the oracle must reproduce them bit for bit; the
bh_math build (the GPU's elementary functions) must agree wherever no transcendental is involved,
and to the stated tolerance elsewhere.
"""
import math
import os

import numpy as np
import pytest

from helpers import ROOT

D2R = math.pi / 180


def bits_equal(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.int64), b.view(np.int64))


@pytest.mark.parametrize("kind", ["libm", "bhm"])
def test_kerr_point_horizon_isco(golden, kind, request):
    o = request.getfixturevalue(f"oracle_{kind}")
    for row, want in zip(golden["kp_in"], golden["kp_out"]):
        assert bits_equal(o.kerr_point(*row), want)
    for a, rh, ri in zip(golden["horizon_in"], golden["horizon_out"], golden["isco_out"]):
        assert o.lib.or_horizon_radius(a) == rh
        assert o.lib.or_isco_radius(a) == ri


@pytest.mark.parametrize("kind", ["libm", "bhm"])
def test_camera_photon(golden, kind, request):
    o = request.getfixturevalue(f"oracle_{kind}")
    for row, want in zip(golden["photon_in"], golden["photon_out"]):
        a, kp, phi0, t0, d, beta = row[0], row[1:12], row[12], row[13], row[14:17], row[17:20]
        assert bits_equal(o.camera_photon(a, kp, phi0, t0, d, beta), want)


@pytest.mark.parametrize("kind", ["libm", "bhm"])
def test_rhs_dp54_projection_bit_exact(golden, kind, request):
    o = request.getfixturevalue(f"oracle_{kind}")
    for row, want in zip(golden["rhs_in"], golden["rhs_out"]):
        assert bits_equal(o.geodesic_rhs(row[0], row[1], row[2], row[3:9]), want)
    for row, want in zip(golden["dp54_in"], golden["dp54_out"]):
        a, lam, eta, y, k1, h, tol = row[0], row[1], row[2], row[3:9], row[9:15], row[15], row[16]
        err, yo, k7 = o.dp54_step(a, lam, eta, y, k1, h, tol)
        assert bits_equal(np.concatenate([[err], yo, k7]), want)
    for row, want in zip(golden["proj_in"], golden["proj_out"]):
        y, k = o.project_radial(row[0], row[1], row[2], row[3:9], row[9:15])
        assert bits_equal(np.concatenate([y, k]), want)


def test_vacuum_fate_libm_exact(golden, oracle_libm):
    fates = np.array([oracle_libm.vacuum_fate(0.94, p, 1e-10, 1100.0, 20000) for p in golden["fate_photons"]])
    assert np.array_equal(fates, golden["fate_tol1e-10"])
    assert set(np.unique(fates)) == {-1, 1}


def test_vacuum_fate_bhm(golden, oracle_bhm):
    # the step controller's x^-0.2 differs from glibc's pow by a few ulp; fates of rays that are not
    # within ~1e-9 of the shadow edge cannot change
    fates = np.array([oracle_bhm.vacuum_fate(0.94, p, 1e-10, 1100.0, 20000) for p in golden["fate_photons"]])
    assert np.array_equal(fates, golden["fate_tol1e-10"])


@pytest.mark.parametrize("kind", ["libm", "bhm"])
def test_critical_curve_point(golden, kind, request):
    o = request.getfixturevalue(f"oracle_{kind}")
    for (a, th, rp), (ok, al, be) in zip(golden["ccp_in"], golden["ccp_out"]):
        got = o.critical_curve_point(a, th, rp)
        assert (got is not None) == bool(ok)
        if got:
            assert got == (al, be)


def test_make_frame(golden, oracle_libm):
    for row, want in zip(golden["mf_in"], golden["mf_out"]):
        F = oracle_libm.make_frame(*row[:16], orbit=bool(row[16]), W=int(row[17]))
        got = np.concatenate([[F.r, F.th, F.ph, F.t], list(F.kp), list(F.F), list(F.R), list(F.U), list(F.beta),
                              [F.cx, F.cy, F.pix, 0.0]])
        assert bits_equal(got, want)


def test_find_critical_point(golden, oracle_libm, oracle_bhm):
    for row, want in zip(golden["fcp_in"], golden["fcp_out"]):
        a, r, incl, phi, roll, psi = row
        F = oracle_libm.make_frame(a, r, incl, phi, roll=roll, fov=30.0, W=1)
        assert oracle_libm.find_critical_point(a, F, psi) == want
        Fb = oracle_bhm.make_frame(a, r, incl, phi, roll=roll, fov=30.0, W=1)
        assert abs(oracle_bhm.find_critical_point(a, Fb, psi) - want) < 1e-9


def test_zoom_target_known_answers(golden):
    # values printed by the reference for the c3 window and for shot 02 (zoom target line)
    fcp = golden["fcp_out"]
    assert round(fcp[-2], 12) == 4.759876947514
    assert round(fcp[-1], 12) == 4.303047749780


def test_selftest_known_answers(oracle_libm, oracle_bhm):
    lines = [l for l in open(os.path.join(ROOT, "tests", "golden", "selftest_reference.txt")) if not l.startswith("#")]
    for o in (oracle_libm, oracle_bhm):
        rows, drift, steps = o.selftest()
        got = ["a=%.2f incl=%4.0f  ray-traced edges %+.6f %+.6f   analytic %+.6f %+.6f  (M)" % tuple(r) for r in rows]
        got.append("radial constraint |R - r'^2| / r^4, worst over %d steps (no projection): %.2e" % (steps, drift))
        assert got == [l.strip() for l in lines]


def test_k2_table(golden, oracle_libm):
    tab, l0, dl = oracle_libm.k2_table(512)
    assert np.array_equal(tab, golden["k2_table"])
    assert l0 == np.float32(math.log(0.02)) and dl == np.float32((math.log(1e4) - math.log(0.02)) / 511)


def test_ref_library_live_when_present(golden):
    """With the reference sources present, the live reference build reproduces the fixtures."""
    path = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref/libref.so not built (reference sources absent)")
    from oracle_ctypes import Ref
    ref = Ref()
    for row, want in zip(golden["dp54_in"][:50], golden["dp54_out"][:50]):
        a, lam, eta, y, k1, h, tol = row[0], row[1], row[2], row[3:9], row[9:15], row[15], row[16]
        err, yo, k7 = ref.dp54_step(a, lam, eta, y, k1, h, tol)
        assert bits_equal(np.concatenate([[err], yo, k7]), want)


def test_turbulence_lookup_closed_forms(oracle_bhm):
    """The oracle's turbulence lookup against what the interpolation must reproduce: a grid that is
    affine in (x, y, z) and affine across time slices is returned exactly (up to rounding) at any
    point and time inside the grid, and times outside the slices clamp to the first/last slice."""
    from oracle_ctypes import make_scene
    n, nt, half, t0, dt = 16, 3, 10.0, 5.0, 2.0
    dx = 2 * half / n
    c = (np.arange(n) + 0.5) * dx - half   # cell centres
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    slope = np.array([0.3, -0.2, 0.1])
    base = 2.0 + slope[0] * x + slope[1] * y + slope[2] * z
    grid = np.stack([base + 0.5 * k for k in range(nt)]).astype(np.float64)
    S = make_scene({"turb_n": n}, grid=grid.astype(np.float32), turb_half=half, turb_t0=t0, turb_dt=dt)
    rng = np.random.default_rng(3)
    for _ in range(200):
        p = rng.uniform(-half + dx, half - dx, 3)
        r = float(np.linalg.norm(p))
        u = p[2] / r
        phi = math.atan2(p[1], p[0])
        t = rng.uniform(t0 - 3, t0 + (nt - 1) * dt + 3)
        k = min(max((t - t0) / dt, 0.0), nt - 1.0)
        want = 2.0 + slope @ p + 0.5 * k
        got = oracle_bhm.turb_factor(S, r, u, phi, t)
        assert abs(got - want) < 1e-5, (p, t, got, want)
    # single slice: time is ignored
    S1 = make_scene({"turb_n": n}, grid=grid[0].astype(np.float32), turb_half=half)
    assert oracle_bhm.turb_factor(S1, 3.0, 0.2, 0.7, 0.0) == oracle_bhm.turb_factor(S1, 3.0, 0.2, 0.7, 99.0)
