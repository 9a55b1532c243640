"""The celestial sphere of the visible mode: the black-hole-frame -> galactic rotation of
build_scene (scene_build.cpp:181-204) against the reference's own build_scene, and the oracle's
sky_uv (trace.cuh:25-33) and checkerboard against what the geometry fixes."""
import math
import os

import numpy as np
import pytest

import bhrt_amd as bh
from helpers import ROOT, oracle_frame, pixel_grid, scene_dict, setup, shot
from oracle_ctypes import Oracle, make_scene

GOLDEN = os.path.join(ROOT, "tests", "golden", "sky_vectors.npz")
GAIA = "/root/reference/blackholes/shots/03_gaia_bh3.ini"


@pytest.fixture(scope="module", params=["libm", "bhm"])
def oracle(request):
    return Oracle(request.param)


def test_sky_rot_matches_reference():
    g = np.load(GOLDEN)
    base = shot("c1_vacuum_256.ini")
    for sets, want, scale in zip(g["sky_sets"], g["sky_rot"], g["sky_scale"]):
        S = bh.build_scene(base, str(sets))[0]
        got = np.array(S.sky_rot[:])
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), (str(sets), got - want)
        assert S.sky_scale == np.float32(scale)
        assert S.sky_on == 1 and (S.sky_nu, S.sky_nv) == (24, 12)
    if os.path.exists(GAIA) and "gaia_rot" in g:
        S = bh.build_scene(GAIA)[0]
        assert np.array_equal(np.array(S.sky_rot[:]), g["gaia_rot"])


def test_sky_keys():
    base = shot("c1_vacuum_256.ini")
    S = bh.build_scene(base, "sky.checker=36, 18")[0]
    assert (S.sky_on, S.sky_nu, S.sky_nv) == (1, 36, 18)
    S = bh.build_scene(base, "sky.checker=40")[0]
    assert (S.sky_nu, S.sky_nv) == (40, 20)
    assert bh.build_scene(base, "sky.checker=0")[0].sky_on == 0
    assert bh.build_scene(shot("c2_radio_1024.ini"))[0].sky_on == 0   # radio frames have no sky
    with pytest.raises(bh.BhError, match="sky.checker"):
        bh.build_scene(base, "sky.checker=-3, 2")


def test_sky_rot_is_a_rotation():
    g = np.load(GOLDEN)
    for m in g["sky_rot"]:
        M = m.reshape(3, 3)
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-13) and abs(np.linalg.det(M) - 1) < 1e-13


def test_sky_uv_axes(oracle):
    """Identity rotation: the galactic axes land on their texture coordinates."""
    I = np.eye(3)
    pi = math.pi
    cases = [((pi / 2, 0.0), (0.5, 0.5)), ((pi / 2, pi / 2), (0.25, 0.5)), ((pi / 2, -pi / 2), (0.75, 0.5)),
             ((pi / 2, pi), (0.0, 0.5)), ((1e-9, 0.3), (None, 0.0)), ((pi - 1e-9, 0.3), (None, 1.0)),
             ((pi / 4, 0.0), (0.5, 0.25)), ((3 * pi / 4, pi / 2), (0.25, 0.75))]
    for (th, ph), (u0, v0) in cases:
        u, v = oracle.sky_uv(I, th, ph)
        if u0 is not None:
            assert min(abs(u - u0), 1 - abs(u - u0)) < 1e-6, (th, ph, u)
        assert abs(v - v0) < 1e-6, (th, ph, v)
        assert 0.0 <= u < 1.0


def test_sky_behind_the_hole_is_l_b(oracle):
    """The escape direction straight behind the hole, seen from the camera at (incl, phi), is
    (pi - incl, phi + pi); the rotation sends it to the galactic direction (l, b)."""
    g = np.load(GOLDEN)
    for sets, m in zip(g["sky_sets"], g["sky_rot"]):
        kv = dict(item.split("=") for item in str(sets).split(";"))
        incl, phi = math.radians(float(kv["camera.incl"])), math.radians(float(kv["camera.phi"]))
        l, b = math.radians(float(kv["sky.l_deg"])), math.radians(float(kv["sky.b_deg"]))
        if incl < 1e-6 or incl > math.pi - 1e-6:
            continue   # on the pole the camera's up falls back to the x axis
        u, v = oracle.sky_uv(m, math.pi - incl, phi + math.pi)
        ue = (0.5 - l / (2 * math.pi)) % 1.0
        assert min(abs(u - ue), 1 - abs(u - ue)) < 2e-6 and abs(v - (0.5 - b / math.pi)) < 2e-6, (str(sets), u, v)


def test_checkerboard(oracle):
    nu, nv, s = 24, 12, 1.5
    vals = np.array([[oracle.sky_checker(nu, nv, s, (i + 0.5) / nu, (j + 0.5) / nv) for i in range(nu)]
                     for j in range(nv)], dtype=np.float32)
    par = (np.add.outer(np.arange(nv), np.arange(nu)) & 1).astype(bool)
    assert np.all(vals[par] == np.float32(s)) and np.all(vals[~par] == np.float32(0.1) * np.float32(s))
    # edges: u -> 1 and v = 1 stay in the last square
    assert oracle.sky_checker(nu, nv, s, 0.99999994, 1.0) == vals[nv - 1, nu - 1]


def test_visible_frame_mirror_symmetry(oracle):
    """Face-on to the spin axis at 90 degrees with a ~ 0 and the galactic centre behind the hole: the
    image is mirror-symmetric about the vertical axis, which reflects the galactic longitude
    (u -> 1 - u) and keeps the latitude; captured rays stay dark and escaped rays take the checker."""
    sets = "bh.spin=1e-6;camera.incl=90;camera.width=64;camera.height=64;sky.l_deg=0;sky.b_deg=0;sky.roll_deg=0"
    cfg = setup(shot("c1_vacuum_256.ini"), sets, with_grid=False)
    px, py = pixel_grid(64, 64)
    out, rays, _ = oracle.trace_pixels(cfg["OS"], cfg["OF"], px, py, spp=1, jitter=False)
    R = rays.reshape(64, 64)
    I = out.reshape(64, 64, 4)
    esc = R["fate"] == 1
    assert esc.mean() > 0.5 and (R["fate"] == -1).any()
    m = esc & esc[:, ::-1]
    du = (R["sky_u"] + R["sky_u"][:, ::-1]) % 1.0
    assert np.all(np.minimum(du, 1 - du)[m] < 1e-5)
    assert np.all(np.abs(R["sky_v"] - R["sky_v"][:, ::-1])[m] < 1e-5)
    assert np.all(I[~esc][:, :3] == 0) and np.all(I[..., 3] == 0)
    s = cfg["S"]
    for y, x in zip(*np.nonzero(esc)):
        want = oracle.sky_checker(s.sky_nu, s.sky_nv, s.sky_scale, float(R["sky_u"][y, x]), float(R["sky_v"][y, x]))
        assert np.all(I[y, x, :3] == np.float64(want))
