"""Pins of the oracle's swept angle psi and photon-ring order channels (blackholes/src/trace.cuh:97-102,
134-140) that do not go through the restatement's own formulas.

- psi against closed forms. For a ~ 0 the orbit is planar and psi is the angle swept within the
  orbital plane, which the Binet equation of Schwarzschild null geodesics gives as a quadrature:
  dpsi / dr = b / (r^2 sqrt(1 - b^2 (1 - 2/r) / r^2)), b^2 = lambda^2 + eta, integrated from the
  camera through the periapsis (escaped rays) to where the ray stopped. For an equatorial ray
  psi = |phi_end - phi_cam|, and cos(theta) stays 0.
- the order split against truncation. Tracing the same ray with max_steps = k stops it after k
  attempts and returns the partial sums, so consecutive k isolate one step's contribution dI and the
  psi after it: dI must land in channel floor(psi / pi) (1, 2, or 3 for >= 3; channel 0 only in the
  total) and in no other."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from helpers import D2R, setup, shot
from oracle_ctypes import OrScene


def _schwarzschild_setup(oracle, incl_deg):
    cfg = setup(shot("c1_vacuum_256.ini"), with_grid=False)
    S = OrScene()
    for name, _ in OrScene._fields_:
        setattr(S, name, getattr(cfg["OS"], name))
    # a = 0 exactly (the scene builder's floor of 1e-6, scene_build.cpp:86, would leave a frame
    # dragging of order a ln(1 / (r - r_+)) in captured rays' psi)
    a = 0.0
    S.a = a
    S.rh = 1.0 + math.sqrt(1.0 - a * a)
    S.r_stop = S.rh + 1e-3
    S.mode = 1                    # radio: psi is tracked
    S.flow_on = 0
    F = oracle.make_frame(a, 1000.0, incl_deg * D2R, fov=30.0, W=256)
    return S, F


def _direction(F, x, y):
    d = np.array(F.F) + np.array(F.R) * (F.cx + x) + np.array(F.U) * (F.cy + y)
    return d / np.linalg.norm(d)


def _binet_psi(b, r_from, r_to, escaped):
    """In-plane angle swept by a Schwarzschild null geodesic of impact parameter b from r_from
    inward, through its periapsis if escaped, to r_to."""
    f = lambda r: b / (r * r * math.sqrt(max(1.0 - b * b * (1.0 - 2.0 / r) / (r * r), 0.0)))
    if not escaped:
        return quad(f, r_to, r_from, epsabs=0, epsrel=1e-13, limit=400)[0]
    # periapsis: largest root of r^3 - b^2 r + 2 b^2 = 0; substitute r = rp + s^2 at the root
    roots = np.roots([1.0, 0.0, -b * b, 2.0 * b * b])
    rp = max(z.real for z in roots if abs(z.imag) < 1e-9)
    for _ in range(3):
        rp -= (rp ** 3 - b * b * rp + 2 * b * b) / (3 * rp * rp - b * b)
    h = lambda s: 2.0 * b / ((rp + s * s) ** 2 * math.sqrt((1.0 - b * b * (1.0 - 2.0 / (rp + s * s)) / (rp + s * s) ** 2)
                                                           / (s * s)))
    inn = quad(h, 0.0, math.sqrt(r_from - rp), epsabs=0, epsrel=1e-13, limit=400)[0]
    out = quad(h, 0.0, math.sqrt(r_to - rp), epsabs=0, epsrel=1e-13, limit=400)[0]
    return inn + out


@pytest.mark.parametrize("kind", ["libm", "bhm"])
def test_psi_equatorial(kind):
    from oracle_ctypes import Oracle
    O = Oracle(kind)
    S, F = _schwarzschild_setup(O, 90.0)
    n = 0
    # b = |x|; psi > pi only within a few tenths of b = 3 sqrt(3) = 5.196
    for x in np.concatenate([np.linspace(-14.0, 14.0, 15), [-5.4, -5.21, 5.2, 5.205, 5.25]]):
        # image-plane point (x, 0) in M (offsets are tangents: M / r_cam), in the equatorial plane
        o = O.trace_ray(S, F, float(x) / F.r - F.cx, -F.cy)
        assert o.fate in (1, -1)
        assert abs(o.y[2]) < 1e-14     # cos(theta) of the camera at 90 deg is cos(pi/2) = 6e-17
        assert abs(o.psi - abs(o.y[4] - F.ph)) <= 1e-12 * max(1.0, o.psi)
        n += o.psi > math.pi
    assert n >= 2   # some rays pass behind the hole


@pytest.mark.parametrize("kind", ["libm", "bhm"])
def test_psi_binet_quadrature(kind):
    from oracle_ctypes import Oracle
    O = Oracle(kind)
    rng = np.random.default_rng(4)
    for incl in (17.0, 60.0):
        S, F = _schwarzschild_setup(O, incl)
        pts = [(rng.uniform(-14, 14), rng.uniform(-14, 14)) for _ in range(10)]
        pts += [(5.3, 0.4), (0.3, -5.25), (-3.0, 4.0), (2.0, 1.0)]   # near b = 3 sqrt(3) and inside it
        for x, y in pts:
            x, y = x / F.r, y / F.r                            # M -> image-plane tangents
            o = O.trace_ray(S, F, x, y)
            assert o.fate in (1, -1)
            ph = O.camera_photon(S.a, np.array(F.kp), F.ph, F.t, _direction(F, x, y), np.array(F.beta))
            b = math.sqrt(ph[0] ** 2 + ph[1])
            want = _binet_psi(b, F.r, o.y[0], o.fate == 1)
            assert abs(o.psi - want) < 2e-8 * max(1.0, want), (incl, x, y, b, o.psi, want)


def _trace(O, S, F, x, y, k):
    S.max_steps = k
    return O.trace_ray(S, F, x, y)


def test_order_channels_by_truncation(oracle_bhm):
    """A ray of the 230 GHz frame that crosses the flow at orders 0..3: at every accepted step the
    intensity added goes to the order channel of the psi after that step."""
    cfg = setup(shot("c2_radio_1024.ini"), with_grid=False)
    S = OrScene()
    for name, _ in OrScene._fields_:
        setattr(S, name, getattr(cfg["OS"], name))
    F = cfg["OF"]
    O = oracle_bhm
    # a ray grazing the critical curve: bisect on the fate along the image x axis
    lo, hi = 2.0 / F.r, 8.0 / F.r                              # M -> image-plane tangents
    S.max_steps = 20000
    f_lo = O.trace_ray(S, F, lo - F.cx, -F.cy).fate
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if O.trace_ray(S, F, mid - F.cx, -F.cy).fate == f_lo:
            lo = mid
        else:
            hi = mid
    x, y = hi + 1e-12 - F.cx, -F.cy
    full = _trace(O, S, F, x, y, 20000)
    assert full.order >= 3 and full.steps < 20000
    prev = _trace(O, S, F, x, y, 1)
    seen = set()
    for k in range(2, full.steps + 1):
        cur = _trace(O, S, F, x, y, k)
        d = [cur.I[c] - prev.I[c] for c in range(4)]
        n = int(cur.psi / math.pi)
        ch = min(n, 3)
        if d[0] > 1e-9 * full.I[0]:
            seen.add(ch)
        for c in (1, 2, 3):
            # (a contribution below I[0]'s last bit can still change a smaller channel)
            want = d[0] if c == ch else 0.0
            assert abs(d[c] - want) <= 1e-15 * abs(full.I[0]), (k, n, c, d)
            if c != ch:
                assert d[c] == 0.0
        prev = cur
    assert seen >= {0, 1, 2, 3}
    assert np.allclose([prev.I[c] for c in range(4)], [full.I[c] for c in range(4)], rtol=0, atol=0)
