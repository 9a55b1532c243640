"""Pins the surrogate's CPU oracle (oracle/mlp_ref.py) without any GPU:
- the FP64 backward pass against central finite differences of the loss;
- the exact normalised coefficients (K2 through the tracer's float table) against the Leung et al.
  formula with K2 from scipy, over the region that carries the 230 GHz frame's flux and the box;
- the K2 table of the product's host layer against scipy's modified Bessel function.
"""
import numpy as np

import bhrt_amd as bh
import mlp_ref as M
from helpers import shot


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(size=M.NPARAM) * 0.3
    xn = rng.uniform(-1, 1, (16, 3))
    t = rng.normal(size=(16, 2))
    w = rng.uniform(0.5, 2.0, 16)
    norm = np.array([-1.5, 1.95, -3.45, 15.5, 2.65, 3.45, 0.3, -0.2, 2.0, 1.5])
    loss, g = M.loss_and_grad(p, norm, xn, t, w)
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 8), rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 12),
                          rng.choice(np.arange(M.OFF_W3, M.OFF_B3), 12), np.arange(M.OFF_W4, M.NPARAM),
                          rng.choice(np.arange(M.OFF_B1, M.OFF_W2), 4)])
    for i in idx:
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (M.loss_and_grad(pp, norm, xn, t, w)[0] - M.loss_and_grad(pm, norm, xn, t, w)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_exact_coefficients_against_scipy():
    _, k2, (l0, dl), _, _ = bh.build_scene(shot("c2_radio_1024.ini"))
    rng = np.random.default_rng(1)
    # region carrying the frame's flux (emission-weighted 0.1..99.9 % quantiles) and the full box
    for lo, hi, th_lo, th_hi, tol in ((3.9, 4.9, 10.0, 42.0, 2e-4), (-4.0, 6.0, 0.5, 100.0, 1e-3)):
        u0 = rng.uniform(lo, hi, 20000)
        th = np.exp(rng.uniform(np.log(th_lo), np.log(th_hi), 20000))
        tb = rng.uniform(0.05, np.pi - 0.05, 20000)
        fj, fa = M.exact_lnf(u0, th, tb, k2, l0, dl)
        ref = M.leung_lnf_scipy(u0, th, tb)
        keep = ref > -60
        assert np.max(np.abs(fj[keep] - ref[keep])) < tol
        # Kirchhoff (Rayleigh-Jeans): alpha B_nu = j with B_nu = 2 nu^2 k T_e / c^2
        assert np.allclose(fa - fj, -np.log(2.0) - np.log(th) - 2 * u0 * np.log(10), atol=1e-12)


def test_k2_table_against_scipy():
    from scipy.special import kve
    _, k2, (l0, dl), _, _ = bh.build_scene(shot("c2_radio_1024.ini"))
    th = np.exp(l0 + dl * np.arange(512, dtype=np.float64))
    ref = np.log(kve(2, 1.0 / th)) - 1.0 / th
    assert np.max(np.abs(k2 - ref) / np.maximum(1.0, np.abs(ref))) < 2e-6
