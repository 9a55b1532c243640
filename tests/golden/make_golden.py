"""Writes tests/golden/ref_vectors.npz: input/output vectors of the reference's own host code.

Runs in the build container only (it needs oracle/_ref/libref.so, compiled from the reference
sources by oracle/Makefile). Every output below is produced by the reference's functions:
kerr_point, camera_photon, geodesic_rhs, dp54_step, project_radial, initial_step, vacuum_fate,
critical_curve_point (blackholes/src/kerr.hpp), make_frame / find_critical_point / build_shot
(blackholes/src/shot.cpp), build_scene + K2 table (blackholes/src/scene_build.cpp) and Config
(common/config.cpp). The inputs are seeded and stored beside the outputs.

    python tests/golden/make_golden.py        (ref_vectors.npz)
    python tests/golden/make_golden.py sky    (sky_vectors.npz: sky_rot of build_scene)
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from oracle_ctypes import Ref  # noqa: E402

REF_ROOT = "/root/reference"
SHOTS = os.path.join(ROOT, "spacetime-optics_amd", "shots")
D2R = math.pi / 180


def main():
    ref = Ref()
    rng = np.random.default_rng(20261202)
    out = {}

    # kerr_point: (a, r, th)
    kin = np.column_stack([rng.uniform(0.0, 0.999, 64), 10 ** rng.uniform(0.3, 7, 64), rng.uniform(0.05, 3.1, 64)])
    out["kp_in"] = kin
    out["kp_out"] = np.array([ref.kerr_point(*row) for row in kin])
    out["horizon_in"] = kin[:, 0].copy()
    out["horizon_out"] = np.array([ref.lib.ref_horizon_radius(a) for a in kin[:, 0]])
    out["isco_out"] = np.array([ref.lib.ref_isco_radius(a) for a in kin[:, 0]])

    # camera_photon: (a, kp, phi0, t0, d, beta) -> (lam, eta, E, y[6])
    rows_in, rows_out = [], []
    for i in range(64):
        a = rng.uniform(0.0, 0.998)
        r = 10 ** rng.uniform(1, 7)
        th = rng.uniform(0.1, 3.0)
        kp = ref.kerr_point(a, r, th)
        d = rng.normal(size=3)
        d[0] = -abs(d[0]) - 2.0
        d /= np.linalg.norm(d)
        beta = np.zeros(3) if i % 2 == 0 else np.array([0.0, 0.0, rng.uniform(-0.5, 0.5)])
        phi0, t0 = rng.uniform(-3, 3), rng.uniform(0, 100)
        rows_in.append(np.concatenate([[a], kp, [phi0, t0], d, beta]))
        rows_out.append(ref.camera_photon(a, kp, phi0, t0, d, beta))
    out["photon_in"] = np.array(rows_in)
    out["photon_out"] = np.array(rows_out)

    # dp54_step / geodesic_rhs / project_radial along real trajectories: integrate rays with the
    # reference's own step and record every (state, step) pair
    rhs_in, rhs_out, dp_in, dp_out, pr_in, pr_out = [], [], [], [], [], []
    for i in range(24):
        a = [0.94, 0.5, 1e-6, 0.99][i % 4]
        F = ref.make_frame(a, 1000.0, 17 * D2R, fov=30.0, W=256)
        kp, Fv, Rv, Uv = F[4:15], F[15:18], F[18:21], F[21:24]
        x, y = rng.uniform(-8, 8, 2) / 1000.0
        d = Fv + Rv * x + Uv * y
        d /= math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        ph = ref.camera_photon(a, kp, F[2], F[3], d)
        lam, eta = ph[0], ph[1]
        yv = ph[3:].copy()
        k1 = ref.geodesic_rhs(a, lam, eta, yv)
        h = ref.lib.ref_initial_step(yv.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_double)))
        tol = [1e-10, 1e-9, 1e-12][i % 3]
        for it in range(400):
            err, yo, k7 = ref.dp54_step(a, lam, eta, yv, k1, h, tol)
            if it % 7 == 0:
                dp_in.append(np.concatenate([[a, lam, eta], yv, k1, [h, tol]]))
                dp_out.append(np.concatenate([[err], yo, k7]))
                rhs_in.append(np.concatenate([[a, lam, eta], yv]))
                rhs_out.append(k1)
            if err > 1.0:
                h *= max(0.2, 0.9 * err ** -0.2)
                continue
            yv, k1 = yo, k7
            pr_in.append(np.concatenate([[a, lam, eta], yv, k1]))
            yv, k1 = ref.project_radial(a, lam, eta, yv, k1)
            pr_out.append(np.concatenate([yv, k1]))
            h *= min(5.0, max(0.2, 0.9 * max(err, 1e-30) ** -0.2))
            if yv[0] < 1.5 or (yv[0] > 1100 and yv[1] > 0):
                break
    out["rhs_in"], out["rhs_out"] = np.array(rhs_in), np.array(rhs_out)
    out["dp54_in"], out["dp54_out"] = np.array(dp_in), np.array(dp_out)
    out["proj_in"], out["proj_out"] = np.array(pr_in[::5]), np.array(pr_out[::5])

    # vacuum_fate for the 256^2, +-15 M, a = 0.94, r = 1000, i = 17 deg geometry: 24 x 24 pixel
    # centres plus rays straddling the shadow edge
    F = ref.make_frame(0.94, 1000.0, 17 * D2R, fov=30.0, W=256)
    kp, Fv, Rv, Uv = F[4:15], F[15:18], F[18:21], F[21:24]
    photons, offsets = [], []
    for py in range(4, 256, 11):
        for px in range(4, 256, 11):
            offsets.append(((px + 0.5 - 128) * F[29], (128 - (py + 0.5)) * F[29]))
    for psi in np.linspace(0, 2 * math.pi, 32, endpoint=False):
        for rho in (4.0, 4.3, 4.6, 4.9, 5.2, 5.5, 5.8):
            offsets.append((rho * math.cos(psi) / 1000.0, rho * math.sin(psi) / 1000.0))
    for (x, y) in offsets:
        d = Fv + Rv * x + Uv * y
        d = d * (1.0 / math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))
        photons.append(ref.camera_photon(0.94, kp, F[2], F[3], d))
    photons = np.array(photons)
    out["fate_photons"] = photons
    out["fate_tol1e-10"] = np.array([ref.vacuum_fate(0.94, p, 1e-10, 1100.0, 20000) for p in photons], dtype=np.int32)

    # critical_curve_point
    cc_in = np.column_stack([rng.uniform(0.05, 0.99, 200), rng.uniform(0.1, 3.0, 200), rng.uniform(1.0, 4.0, 200)])
    cc = []
    import ctypes
    for a, th, rp in cc_in:
        al, be = ctypes.c_double(), ctypes.c_double()
        ok = ref.lib.ref_critical_curve_point(a, th, rp, ctypes.byref(al), ctypes.byref(be))
        cc.append((ok, al.value if ok else 0.0, be.value if ok else 0.0))
    out["ccp_in"], out["ccp_out"] = cc_in, np.array(cc)

    # find_critical_point: self-test geometry (r = 1e7, fov 30 M), the c3 zoom window geometry and
    # the reference's shot-02 camera (r = 1e7, incl 163, roll 108)
    fcp_in = []
    for a in (1e-6, 0.5, 0.94, 0.99):
        for incl in (90.0, 17.0):
            for psi in (0.0, math.pi):
                fcp_in.append((a, 1e7, incl * D2R, 0.0, 0.0, psi))
    fcp_in.append((0.94, 1000.0, 17 * D2R, 0.0, 0.0, -100 * D2R))
    fcp_in.append((0.94, 1e7, 163 * D2R, 0.0, 108 * D2R, -100 * D2R))
    out["fcp_in"] = np.array(fcp_in)
    out["fcp_out"] = np.array([ref.find_critical_point(a, r, incl, phi, psi, roll=roll)
                               for (a, r, incl, phi, roll, psi) in fcp_in])

    # make_frame over camera variants (a, r, incl, phi, roll, yaw, pitch, fov, fov_in_m, zoom, target,
    # tx, ty, offx, offy, time, orbit, W)
    mf_in = []
    for i in range(24):
        mf_in.append((rng.uniform(0, 0.99), 10 ** rng.uniform(1.5, 7), rng.uniform(0.1, 3.0), rng.uniform(-1, 1),
                      rng.uniform(-2, 2) * (i % 2), rng.uniform(-0.3, 0.3) * (i % 3 == 0),
                      rng.uniform(-0.3, 0.3) * (i % 4 == 0), rng.uniform(10, 60) if i % 5 else rng.uniform(0.2, 1.0),
                      1 if i % 5 else 0, 10 ** rng.uniform(0, 4), rng.uniform(0, 1), rng.uniform(-5, 5),
                      rng.uniform(-5, 5), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0, 500), i % 6 == 0,
                      int(rng.integers(64, 4096))))
    out["mf_in"] = np.array(mf_in, dtype=np.float64)
    out["mf_out"] = np.array([ref.make_frame(*row[:16], orbit=bool(row[16]), W=int(row[17])) for row in mf_in])

    # build_scene of the in-repo shot files (+ the reference's M87 base and zoom shot when present)
    names, scenes, k2s = [], [], []
    files = [os.path.join(SHOTS, f) for f in ("c1_vacuum_256.ini", "c2_radio_1024.ini", "c3_zoom_2048.ini",
                                             "c4_mlp_4k.ini")]
    files += [os.path.join(HERE, "grammar_test.ini")]
    ref_shots = [os.path.join(REF_ROOT, "blackholes/shots", f) for f in ("base_m87.ini", "02_photon_ring_zoom.ini",
                                                                        "01_m87_eht.ini")]
    for f in files + ref_shots:
        if not os.path.exists(f):
            continue
        s, k2 = ref.build_scene(f)
        names.append(os.path.relpath(f, ROOT) if f.startswith(ROOT) else "REF:" + os.path.relpath(f, REF_ROOT))
        scenes.append(s)
        k2s.append(k2)
    out["scene_names"] = np.array(names)
    out["scene_out"] = np.array(scenes)
    out["k2_table"] = k2s[0]
    # overrides through "--set"
    s, _ = ref.build_scene(os.path.join(SHOTS, "c2_radio_1024.ini"),
                           "bh.spin=1.5;render.seed=7;render.horizon_margin=0.02;flow.n0=3e5")
    out["scene_set_out"] = s
    s, _ = ref.build_scene(os.path.join(SHOTS, "c2_radio_1024.ini"),
                           "bh.spin=-0.3;flow.turbulence_scale=0.3;flow.correlation_orbits=0.8")
    out["scene_set2_out"] = s

    # build_shot: the track fixture and the reference's zoom and EHT shots
    out["shot_track"] = ref.build_shot(os.path.join(HERE, "track_test.ini"))
    for key, f in (("shot_ref02", "02_photon_ring_zoom.ini"), ("shot_ref01", "01_m87_eht.ini")):
        p = os.path.join(REF_ROOT, "blackholes/shots", f)
        if os.path.exists(p):
            out[key] = ref.build_shot(p)

    # Config grammar: raw values of the grammar fixture
    keys = ["top_level", "bh.spin", "bh.mass_msun", "render.tolerance", "render.mode", "render.label", "render.empty",
            "render.flag_yes", "render.flag_off", "render.flag_bad", "render.list", "camera.width", "camera.height",
            "camera.equals", "render.missing"]
    vals = [ref.config_str(os.path.join(HERE, "grammar_test.ini"), k) for k in keys]
    out["cfg_keys"] = np.array(keys)
    out["cfg_vals"] = np.array(["<absent>" if v is None else v for v in vals])

    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"), {k: np.shape(v) for k, v in out.items()})


def sky():
    """tests/golden/sky_vectors.npz: sky_rot of the reference's build_scene for camera and [sky]
    orientations (seeded, plus the pole and the reference's own sky shot)."""
    ref = Ref()
    rng = np.random.default_rng(20261203)
    base = os.path.join(SHOTS, "c1_vacuum_256.ini")
    rows = [(0.0, 0.0, 0.0, 0.0, 0.0), (180.0, 30.0, 10.0, -5.0, 20.0), (90.0, 0.0, 0.0, 0.0, 0.0),
            (17.0, 0.0, 2.0, -2.0, 0.0)]
    for _ in range(60):
        rows.append((rng.uniform(0, 180), rng.uniform(-180, 360), rng.uniform(-180, 360), rng.uniform(-85, 85),
                     rng.uniform(-180, 180)))
    sets, rot, scale = [], [], []
    for k, (incl, phi, l, b, roll) in enumerate(rows):
        st = (f"camera.incl={incl!r};camera.phi={phi!r};sky.l_deg={l!r};sky.b_deg={b!r};sky.roll_deg={roll!r};"
              f"sky.brightness={0.5 + k / 16!r}")
        m, sc = ref.sky_rot(base, st)
        sets.append(st)
        rot.append(m)
        scale.append(sc)
    out = {"sky_sets": np.array(sets), "sky_rot": np.array(rot), "sky_scale": np.array(scale)}
    gaia = os.path.join(REF_ROOT, "blackholes", "shots", "03_gaia_bh3.ini")
    if os.path.exists(gaia):
        m, sc = ref.sky_rot(gaia)
        out["gaia_rot"] = m
    np.savez_compressed(os.path.join(HERE, "sky_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "sky_vectors.npz"), {k: np.shape(v) for k, v in out.items()})


if __name__ == "__main__":
    if sys.argv[1:] == ["sky"]:
        sky()
    else:
        main()
