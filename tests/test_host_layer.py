"""The product's host layer (libbhrt.so, no GPU needed) against the reference's own host code:
golden vectors from tests/golden/ref_vectors.npz (tests/golden/make_golden.py). Also the turbulence
grid's statistics and the exported C ABI."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import bhrt_amd as bh
from helpers import ROOT

REF_SHOTS = "/root/reference/blackholes/shots"
SCENE_KEYS = ["a", "rh", "r_isco", "tol", "max_steps", "r_escape", "r_stop", "step_frac", "mode", "flow_on", "n0",
              "n_index", "theta_index", "theta0", "h_r", "beta_p", "r_flow_max", "nu_obs", "rg_cm", "flow_kep",
              "flow_turb_amp", "k2_n", "k2_l0", "k2_dl", "seed", "distance_cm", "uas_per_m", "flux_target_jy",
              "mass_msun", "flow_turb_scale", "flow_tcorr"]


def scene_vector(path, sets=""):
    S, k2, (l0, dl), info, _ = bh.build_scene(path, sets)
    d = {name: getattr(S, name) for name, _ in S._fields_}
    d.update(k2_n=len(k2), k2_l0=l0, k2_dl=dl, distance_cm=info["distance_cm"], uas_per_m=info["uas_per_m"],
             flux_target_jy=info["flux_target_jy"], mass_msun=info["mass_msun"])
    return np.array([float(d[k]) for k in SCENE_KEYS]), k2


def resolve(name):
    if name.startswith("REF:"):
        p = os.path.join("/root/reference", name[4:])
        return p if os.path.exists(p) else None
    return os.path.join(ROOT, name)


def test_build_scene_matches_reference(golden):
    for name, want in zip(golden["scene_names"], golden["scene_out"]):
        path = resolve(str(name))
        if path is None:
            continue   # reference shot files are only present where the reference is
        got, k2 = scene_vector(path)
        assert np.array_equal(got, want), (name, [(k, a, b) for k, a, b in zip(SCENE_KEYS, got, want) if a != b])
        assert np.array_equal(k2, golden["k2_table"])


def test_build_scene_overrides(golden):
    p = os.path.join(ROOT, "spacetime-optics_amd", "shots", "c2_radio_1024.ini")
    got, _ = scene_vector(p, "bh.spin=1.5;render.seed=7;render.horizon_margin=0.02;flow.n0=3e5")
    assert np.array_equal(got, golden["scene_set_out"])   # spin clamped to 0.9999, seed hash, margin
    got, _ = scene_vector(p, "bh.spin=-0.3;flow.turbulence_scale=0.3;flow.correlation_orbits=0.8")
    assert np.array_equal(got, golden["scene_set2_out"])   # a <= 0 -> 1e-6; sheared-noise parameters


def test_config_grammar(golden):
    path = os.path.join(ROOT, "tests", "golden", "grammar_test.ini")
    for k, v in zip(golden["cfg_keys"], golden["cfg_vals"]):
        got = bh.config_str(path, str(k))
        assert (got if got is not None else "<absent>") == str(v), k


def test_config_errors(tmp_path):
    bad = str(tmp_path / "bad_section.ini")
    with open(bad, "w") as f:
        f.write("[render\nmode = radio\n")
    with pytest.raises(bh.BhError, match="bad section header"):
        bh.config_str(bad, "render.mode")
    with open(bad, "w") as f:
        f.write("[render]\njust a line\n")
    with pytest.raises(bh.BhError, match="expected key = value"):
        bh.config_str(bad, "render.mode")
    with pytest.raises(bh.BhError, match="cannot open"):
        bh.config_str(bad + ".missing", "render.mode")


def test_build_shot_tracks(golden):
    got = bh.build_shot(os.path.join(ROOT, "tests", "golden", "track_test.ini"))
    assert np.array_equal(got, golden["shot_track"])
    for key, f in (("shot_ref02", "02_photon_ring_zoom.ini"), ("shot_ref01", "01_m87_eht.ini")):
        p = os.path.join(REF_SHOTS, f)
        if key in golden and os.path.exists(p):
            assert np.array_equal(bh.build_shot(p), golden[key])


def test_make_frame(golden):
    for row, want in zip(golden["mf_in"], golden["mf_out"]):
        F = bh.make_frame(*row[:16], orbit=bool(row[16]), W=int(row[17]))
        got = np.concatenate([[F.r, F.th, F.ph, F.t], list(F.kp), list(F.F), list(F.R), list(F.U), list(F.beta),
                              [F.cx, F.cy, F.pix, 0.0]])
        assert np.array_equal(got.view(np.int64), want.view(np.int64))


def test_turbulence_grid_statistics():
    """Seeded log-normal field: mean 1, std(ln) = sigma, Gaussian correlation of length 5 M."""
    n, half, sigma, corr = 128, 60.0, 0.3, 5.0
    g = bh.turbulence(n, 1, half, sigma, corr, 30.0, 10.0, 1)[0].astype(np.float64)
    assert np.array_equal(g, bh.turbulence(n, 1, half, sigma, corr, 30.0, 10.0, 1)[0])   # deterministic
    assert not np.array_equal(g, bh.turbulence(n, 1, half, sigma, corr, 30.0, 10.0, 2)[0])
    lg = np.log(g) + 0.5 * sigma ** 2
    assert abs(g.mean() - 1.0) < 0.02
    assert abs(lg.std() - sigma) < 0.02
    dx = 2 * half / n
    lag = int(round(corr / dx))
    c = np.mean(lg * np.roll(lg, lag, axis=2)) / lg.var()
    want = math.exp(-0.5 * (lag * dx / corr) ** 2)
    assert abs(c - want) < 0.05


def test_c_abi_exports_every_declared_symbol():
    L = bh.load()
    declared = []
    for h in ("bhrt.h", "bhrt_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        declared += re.findall(r"^(?:int|void|const char\*|double)\s+\**(bh\w+)\(", text, flags=re.M)
    assert len(declared) >= 25
    assert sorted(declared) == sorted(bh.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


def test_struct_layouts_match_the_library():
    """Every ctypes mirror of a public struct has the C compiler's size, and every field sits at the
    C compiler's offset (bhh_abi_offsets, struct by struct, fields in declaration order): a missing,
    added, reordered or wrongly sized field on either side fails."""
    L = bh.load()
    mirrors = [bh.Scene, bh.Tables, bh.Frame, bh.Settings, bh.TraceOpts, bh.Stats, bh.RAY_DTYPE, bh.BatchFrame,
               bh.BatchCfg, bh.MlpTrainCfg, bh.MlpReport, bh.MlpBakeReport, bh.PostParams, bh.ToneParams,
               bh.HudGlyph, bh.HudRect, bh.Font]
    n = L.bhh_abi_sizes(None, 0)
    assert n == len(mirrors)
    sizes = (C.c_size_t * n)()
    assert L.bhh_abi_sizes(sizes, n) == n
    for m, got in zip(mirrors, sizes):
        want = m.itemsize if isinstance(m, np.dtype) else C.sizeof(m)
        assert got == want, (getattr(m, "__name__", "RAY_DTYPE"), got, want)
    want = []
    for m in mirrors:
        if isinstance(m, np.dtype):
            want += [(m.names[i], m.fields[f][1]) for i, f in enumerate(m.names)]
        else:
            want += [(m.__name__ + "." + name, getattr(m, name).offset) for name, _ in m._fields_]
    n = L.bhh_abi_offsets(None, 0)
    offs = (C.c_size_t * n)()
    assert L.bhh_abi_offsets(offs, n) == n
    assert n == len(want), (n, len(want))
    for (name, w), got in zip(want, offs):
        assert got == w, (name, got, w)


def test_error_paths_without_gpu():
    S = bh.Scene()
    ctx = C.c_void_p()
    err = C.create_string_buffer(256)
    rc = bh.load().bh_create(0, None, None, None, C.byref(ctx), err, 256)
    assert rc == bh.BH_ERR_ARG
    assert bh.load().bh_update_scene(None, C.byref(S)) == bh.BH_ERR_ARG
    assert bh.load().bh_math_eval(99, None, None, None, 1) == bh.BH_ERR_ARG
