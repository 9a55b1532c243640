"""The oracle's restatement of the reference's flow turbulence (oracle/sheared.cpp), pinned by what
the mathematics fixes. The reference's noise is device code that cannot be compiled here, so its
own values are not available: parity with them is unpinned, and these checks make the plausible
mistakes (a wrong gradient table, fade, octave scaling, cross-fade normalisation, advection
direction or rate, texel-centre or wrap convention, filter-weight rounding) fail.

- gradient noise (device_util.cuh:31-59) vanishes at lattice points, where its gradient is one of
  the 12 cube-edge directions, each chosen about equally often by the hash;
- fBm (device_util.cuh:61-71): one octave is 2.2 x the gradient noise; four octaves have zero mean
  and 2.2^2 times its variance;
- sheared noise (emission.cuh:34-51): continuous across the birth of a generation, of the same
  variance mid-way between generations (the 1 / sqrt((1-f)^2 + f^2) normalisation), and carried
  by the Keplerian rotation Omega = 1 / (R^1.5 + a) (R >= r_isco) within a generation;
- the slice filter: exact at texel centres, 1/256-rounded weights between them, wrap in phi, clamp
  in the other coordinates;
- the bake (render.cu:449-460): ln of the density factor has mean -a^2/2 and spread ~a.
"""
import math

import numpy as np
import pytest

from oracle_ctypes import OrScene, Oracle


@pytest.fixture(scope="module")
def orc():
    return Oracle("bhm")


def scene(a=0.94, amp=0.7, scale=0.45, tcorr=0.5, seed=2654435761):
    S = OrScene()
    S.a, S.rh, S.r_isco = a, 1.0 + math.sqrt(1 - a * a), 2.0
    S.h_r, S.r_flow_max = 0.3, 60.0
    S.flow_turb_amp, S.flow_turb_scale, S.flow_tcorr, S.seed = amp, scale, tcorr, seed
    return S


def test_gnoise_zero_at_lattice_points_with_edge_gradients(orc):
    rng = np.random.default_rng(3)
    dirs = {}
    eps = 1e-3
    for _ in range(2400):
        ix, iy, iz = (int(v) for v in rng.integers(-500, 500, 3))
        seed = int(rng.integers(0, 2 ** 32))
        assert orc.gnoise(ix, iy, iz, seed) == 0.0
        g = np.array([orc.gnoise(ix + eps, iy, iz, seed), orc.gnoise(ix, iy + eps, iz, seed),
                      orc.gnoise(ix, iy, iz + eps, seed)]) / eps
        r = np.round(g)
        assert np.all(np.abs(g - r) < 1e-2), g
        assert np.sum(np.abs(r)) == 2 and np.max(np.abs(r)) == 1, r   # a cube edge: two of (+-1, +-1, 0)
        dirs[tuple(r)] = dirs.get(tuple(r), 0) + 1
    assert len(dirs) == 12
    assert min(dirs.values()) > 0.7 * 200 and max(dirs.values()) < 1.3 * 200


def test_fbm_octave_scaling_and_moments(orc):
    rng = np.random.default_rng(4)
    pts = rng.uniform(-50, 50, (4000, 3)).astype(np.float32)
    for p in pts[:200]:
        assert orc.fbm(*p, 1, 77) == np.float32(orc.gnoise(*p, 77) * np.float32(2.2))
    # independent octaves (seeds 1013 apart) of weight 0.55^i, rescaled by 2.2 / sqrt(sum 0.55^2i):
    # the variance is 2.2^2 = 4.84 times the gradient noise's
    v = np.array([orc.fbm(*p, 4, 12345) for p in pts])
    g = np.array([orc.gnoise(*p, 999) for p in pts])
    assert abs(v.mean()) < 0.05
    assert 0.9 * 4.84 < v.var() / g.var() < 1.1 * 4.84


def sheared_at(orc, B, R, phi, z, t):
    return orc.sheared_noise(B, R, phi, z, t)


def test_sheared_noise_crossfade_and_advection(orc):
    S = scene()
    B = orc.sheared_params(S, 0.0, 1000.0)
    rng = np.random.default_rng(5)
    cont, adv, f0, fh = [], [], [], []
    for _ in range(600):
        R = float(rng.uniform(3, 50))
        phi = float(rng.uniform(0, 2 * math.pi))
        z = float(rng.uniform(-0.3, 0.3) * R)
        Om = 1.0 / (max(R, S.r_isco) ** 1.5 + S.a)
        P = 2 * math.pi / Om * S.flow_tcorr
        k = float(rng.integers(-20, 20))
        tb = k * P
        # continuity across the birth of generation k
        cont.append(abs(sheared_at(orc, B, R, phi, z, tb + 1e-4 * P) - sheared_at(orc, B, R, phi, z, tb - 1e-4 * P)))
        # within a generation the pattern moves with the flow: phi + Omega dt at t + dt
        d = 0.004 * P
        adv.append(abs(sheared_at(orc, B, R, phi + Om * d, z, tb + 0.01 * P + d)
                       - sheared_at(orc, B, R, phi, z, tb + 0.01 * P)))
        f0.append(sheared_at(orc, B, R, phi, z, tb + 1e-6 * P))
        fh.append(sheared_at(orc, B, R, phi, z, tb + 0.5 * P))
    assert np.median(cont) < 5e-3 and max(cont) < 0.05
    # the cross-fade moves the blend by 0.4 % of a generation: the change is a small fraction of the
    # field, far below what a wrong rotation rate or direction gives (the pattern moves by
    # Omega * dt * R / (scale R) ~ 0.05 feature sizes)
    wrong = [abs(sheared_at(orc, B, R, 0.7 - Om * 0.004 * P, 0.1 * R, 0.5 * P + 0.004 * P)
                 - sheared_at(orc, B, R, 0.7, 0.1 * R, 0.5 * P))
             for R in np.linspace(3, 50, 200) for Om, P in [(1.0 / (max(R, 2.0) ** 1.5 + 0.94),
                                                             2 * math.pi * (max(R, 2.0) ** 1.5 + 0.94) * 0.5)]]
    assert np.median(adv) < 0.3 * np.median(wrong)
    # mid-way between generations the normalisation keeps the variance
    assert 0.75 < np.var(fh) / np.var(f0) < 1.33


def test_slice_filter_conventions(orc):
    nz, ny, nx = 4, 5, 8
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    T = (i + 10 * j + 100 * k).astype(np.float32)
    for (ii, jj, kk) in [(0, 0, 0), (3, 2, 1), (7, 4, 3)]:
        x, y, z = (ii + 0.5) / nx, (jj + 0.5) / ny, (kk + 0.5) / nz
        assert orc.sheared_tex(T, x, y, z) == T[kk, jj, ii]
    # half-way in y and z: averages; a third of a texel in x: weight rounded to 85/256
    assert orc.sheared_tex(T, 3.5 / nx, 2.0 / ny, 2.0 / nz) == np.float32(3 + 15 + 150)
    got = orc.sheared_tex(T, (3.5 + 1 / 3) / nx, 2.5 / ny, 1.5 / nz)
    assert got == np.float32(3 + 85 / 256 + 20 + 100)
    # wrap in x: the first texel's left half blends the last and first texels
    assert orc.sheared_tex(T, 0.0, 0.5 / ny, 0.5 / nz) == np.float32(3.5)
    assert orc.sheared_tex(T, 1.0 + 1.5 / nx, 0.5 / ny, 0.5 / nz) == T[0, 0, 1]
    # clamp in y and z
    assert orc.sheared_tex(T, 0.5 / nx, 7.0, -3.0) == T[0, ny - 1, 0]


def test_bake_statistics(orc):
    S = scene(amp=0.5)
    B = orc.sheared_params(S, 0.0, 1000.0, dims=(64, 24, 8, 3))
    v = orc.sheared_bake(B)
    assert v.shape == (3, 8, 24, 64) and np.all(v > 0)
    lv = np.log(v.astype(np.float64))
    assert abs(lv.mean() + 0.5 * 0.5 ** 2) < 0.12
    assert 0.15 < lv.std() < 1.0
    # the bake is the cell function
    for (i, j, l, k) in [(0, 0, 0, 0), (63, 23, 7, 2), (17, 5, 3, 1)]:
        assert v[k, l, j, i] == orc.sheared_cell(B, i, j, l, k)
