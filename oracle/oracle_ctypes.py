"""ctypes binding of the CPU oracle (oracle.h) and of the reference driver (_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker or the timed CPU baseline, never as part of the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class OrSheared(C.Structure):
    _fields_ = [
        ("a", C.c_double), ("r_isco", C.c_double), ("cam_t", C.c_double), ("cam_r", C.c_double),
        ("lr0", C.c_float), ("lr1", C.c_float), ("zmax", C.c_float), ("d0", C.c_float), ("dd", C.c_float),
        ("amp", C.c_float), ("scale", C.c_float), ("tcorr", C.c_float), ("seed", C.c_uint32),
        ("nphi", C.c_int), ("nlr", C.c_int), ("nz", C.c_int), ("nt", C.c_int),
    ]


class OrScene(C.Structure):
    _fields_ = [
        ("a", C.c_double), ("rh", C.c_double), ("r_isco", C.c_double), ("tol", C.c_double),
        ("max_steps", C.c_int), ("r_escape", C.c_double), ("r_stop", C.c_double), ("step_frac", C.c_double),
        ("mode", C.c_int), ("flow_on", C.c_int),
        ("n0", C.c_double), ("n_index", C.c_double), ("theta_index", C.c_double), ("theta0", C.c_double),
        ("h_r", C.c_double), ("beta_p", C.c_double), ("r_flow_max", C.c_double), ("nu_obs", C.c_double),
        ("rg_cm", C.c_double), ("flow_kep", C.c_double), ("flow_turb_amp", C.c_float),
        ("turb_grid", C.POINTER(C.c_float)), ("turb_n", C.c_int), ("turb_half", C.c_double),
        ("turb_inv_dx", C.c_double), ("turb_nt", C.c_int), ("turb_t0", C.c_double), ("turb_inv_dt", C.c_double),
        ("k2_lntab", C.POINTER(C.c_float)), ("k2_n", C.c_int), ("k2_l0", C.c_float), ("k2_dl", C.c_float),
        ("seed", C.c_uint32), ("flow_turb_scale", C.c_float), ("flow_tcorr", C.c_float),
        ("turb_model", C.c_int), ("ts_slices", C.POINTER(C.c_float)), ("ts", OrSheared),
        ("sky_on", C.c_int), ("sky_nu", C.c_int), ("sky_nv", C.c_int), ("sky_scale", C.c_float),
        ("sky_rot", C.c_double * 9),
    ]


class OrFrame(C.Structure):
    _fields_ = [
        ("r", C.c_double), ("th", C.c_double), ("ph", C.c_double), ("t", C.c_double),
        ("kp", C.c_double * 11), ("F", C.c_double * 3), ("R", C.c_double * 3), ("U", C.c_double * 3),
        ("beta", C.c_double * 3), ("cx", C.c_double), ("cy", C.c_double), ("pix", C.c_double),
        ("W", C.c_int), ("H", C.c_int), ("spp", C.c_int), ("seed", C.c_uint32),
    ]


class OrRay(C.Structure):
    _fields_ = [
        ("y", C.c_double * 6), ("th_inf", C.c_double), ("ph_inf", C.c_double), ("I", C.c_double * 4),
        ("trans", C.c_double), ("psi", C.c_double),
        ("fate", C.c_int), ("ncross", C.c_int), ("order", C.c_int), ("steps", C.c_int),
        ("accepted", C.c_int), ("flow_steps", C.c_int), ("sky_u", C.c_float), ("sky_v", C.c_float),
    ]


RAY_DTYPE = np.dtype([
    ("y", "<f8", (6,)), ("th_inf", "<f8"), ("ph_inf", "<f8"), ("I", "<f8", (4,)), ("trans", "<f8"),
    ("psi", "<f8"), ("fate", "<i4"), ("ncross", "<i4"), ("order", "<i4"), ("steps", "<i4"),
    ("accepted", "<i4"), ("flow_steps", "<i4"), ("sky_u", "<f4"), ("sky_v", "<f4"),
])
assert RAY_DTYPE.itemsize == C.sizeof(OrRay)

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)


def _dp(a):
    return a.ctypes.data_as(_D)


class Oracle:
    """One build of the oracle: kind 'bhm' (GPU-parity elementary functions) or 'libm' (glibc)."""

    def __init__(self, kind="bhm"):
        path = os.path.join(HERE, f"liboracle_{kind}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
        self.kind = kind
        L = self.lib = C.CDLL(path)
        L.or_kerr_point.argtypes = [C.c_double, C.c_double, C.c_double, _D]
        L.or_horizon_radius.restype = C.c_double
        L.or_horizon_radius.argtypes = [C.c_double]
        L.or_isco_radius.restype = C.c_double
        L.or_isco_radius.argtypes = [C.c_double]
        L.or_camera_photon.argtypes = [C.c_double, _D, C.c_double, C.c_double, _D, _D, _D]
        L.or_radial_potential.restype = C.c_double
        L.or_radial_potential.argtypes = [C.c_double] * 4
        L.or_polar_potential.restype = C.c_double
        L.or_polar_potential.argtypes = [C.c_double] * 4
        L.or_geodesic_rhs.argtypes = [C.c_double, C.c_double, C.c_double, _D, _D]
        L.or_dp54_step.restype = C.c_double
        L.or_dp54_step.argtypes = [C.c_double, C.c_double, C.c_double, _D, _D, C.c_double, C.c_double, _D, _D]
        L.or_project_radial.argtypes = [C.c_double, C.c_double, C.c_double, _D, _D]
        L.or_initial_step.restype = C.c_double
        L.or_initial_step.argtypes = [_D]
        L.or_vacuum_fate.restype = C.c_int
        L.or_vacuum_fate.argtypes = [C.c_double, _D, C.c_double, C.c_double, C.c_int]
        L.or_critical_curve_point.restype = C.c_int
        L.or_critical_curve_point.argtypes = [C.c_double] * 3 + [_D, _D]
        L.or_make_frame.argtypes = ([C.c_double] * 8 + [C.c_int] + [C.c_double] * 7 + [C.c_int, C.c_int, C.c_int,
                                    C.c_uint32, C.POINTER(OrFrame)])
        L.or_find_critical_point.restype = C.c_double
        L.or_find_critical_point.argtypes = [C.c_double, C.POINTER(OrFrame), C.c_double]
        L.or_pcg.restype = C.c_uint32
        L.or_pcg.argtypes = [C.c_uint32]
        L.or_pixel_seed.argtypes = [C.c_uint32, C.c_int, _F, _F]
        L.or_r2_sample.argtypes = [C.c_uint32, C.c_float, C.c_float, _F, _F]
        L.or_table_lerp.restype = C.c_float
        L.or_table_lerp.argtypes = [_F, C.c_int, C.c_float, C.c_float, C.c_float]
        L.or_sample_offsets.argtypes = [C.POINTER(OrFrame), C.c_int, C.c_int, C.c_int, C.c_int, _D, _D]
        L.or_k2_table.argtypes = [C.c_int, _F, _F, _F]
        L.or_render_adaptive.restype = C.c_longlong
        L.or_render_adaptive.argtypes = [C.POINTER(OrScene), C.POINTER(OrFrame), C.c_int, C.c_int, C.c_float, C.c_int,
                                         C.c_int, _D, C.POINTER(C.c_uint32)]
        L.or_turb_factor.restype = C.c_double
        L.or_turb_factor.argtypes = [C.POINTER(OrScene), C.c_double, C.c_double, C.c_double, C.c_double]
        L.or_flow_step.argtypes = [C.POINTER(OrScene)] + [C.c_double] * 7 + [_D, _D]
        _SH = C.POINTER(OrSheared)
        L.or_sheared_params.argtypes = [C.POINTER(OrScene), C.c_double, C.c_double, C.POINTER(C.c_int), _SH]
        L.or_sheared_gnoise.restype = C.c_float
        L.or_sheared_gnoise.argtypes = [C.c_float, C.c_float, C.c_float, C.c_uint32]
        L.or_sheared_fbm.restype = C.c_float
        L.or_sheared_fbm.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, C.c_uint32]
        L.or_sheared_noise.restype = C.c_float
        L.or_sheared_noise.argtypes = [_SH, C.c_double, C.c_double, C.c_double, C.c_double]
        L.or_sheared_tex.restype = C.c_float
        L.or_sheared_tex.argtypes = [_F, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float]
        L.or_sheared_cell.restype = C.c_float
        L.or_sheared_cell.argtypes = [_SH, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_sheared_bake.argtypes = [_SH, _F, C.c_int]
        L.or_sheared_factor.restype = C.c_float
        L.or_sheared_factor.argtypes = [C.POINTER(OrScene)] + [C.c_double] * 5
        L.or_post_blur.argtypes = [_F, C.c_int, C.c_int, C.c_double]
        L.or_frame_stats.argtypes = [_F, C.c_int, C.c_int, C.c_float, C.c_double, _D, _D]
        L.or_tonemap_radio.argtypes = ([_F, C.c_int, C.c_int, C.c_int] + [C.c_float] * 7 +
                                       [C.c_uint32, C.POINTER(C.c_uint8)])
        L.or_hud_draw.argtypes = [C.POINTER(C.c_uint8), C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                  C.POINTER(C.c_uint8), C.c_int, C.c_int, C.c_float]
        L.or_escape_angles.argtypes = [_D, _D, _D, _D]
        L.or_trace_ray.argtypes = [C.POINTER(OrScene), C.POINTER(OrFrame), C.c_double, C.c_double, C.POINTER(OrRay)]
        L.or_trace_pixels.restype = C.c_int
        L.or_trace_pixels.argtypes = [C.POINTER(OrScene), C.POINTER(OrFrame), _I, _I, C.c_int, C.c_int, C.c_int,
                                      C.c_int, _D, C.c_void_p]
        L.or_selftest.argtypes = [_D, _D, _I]
        L.or_math_eval.argtypes = [C.c_int, _D, _D, _D, C.c_int]
        L.or_sky_uv.argtypes = [_D, C.c_double, C.c_double, _F, _F]
        L.or_trace_sample.argtypes = [C.POINTER(OrScene), C.POINTER(OrFrame), C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(OrRay)]
        L.or_sky_checker.restype = C.c_float
        L.or_sky_checker.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float]

    MATH_FN = {"exp": 0, "log": 1, "expm1": 2, "cbrt": 3, "pow_m02": 4, "sin": 5, "cos": 6, "atan2": 7,
               "acos": 8, "sqrt": 9, "div": 10, "asin": 11}

    def sky_uv(self, rot, th, ph):
        """Galactic (u, v) of escape directions (th, ph) through the row-major 3 x 3 rotation rot."""
        r = np.ascontiguousarray(rot, dtype=np.float64).ravel()
        th, ph = np.broadcast_arrays(np.asarray(th, dtype=np.float64), np.asarray(ph, dtype=np.float64))
        u, v = np.zeros(th.shape, np.float32), np.zeros(th.shape, np.float32)
        uu, vv = C.c_float(), C.c_float()
        for i, (a, b) in enumerate(zip(th.ravel(), ph.ravel())):
            self.lib.or_sky_uv(_dp(r), float(a), float(b), C.byref(uu), C.byref(vv))
            u.flat[i], v.flat[i] = uu.value, vv.value
        return u, v

    def trace_sample(self, S, F, px, py, s, jitter=True):
        """Per-ray record of sample s of pixel (px, py)."""
        o = OrRay()
        self.lib.or_trace_sample(C.byref(S), C.byref(F), px, py, s, int(jitter), C.byref(o))
        return o

    def sky_checker(self, nu, nv, scale, u, v):
        return self.lib.or_sky_checker(nu, nv, scale, float(u), float(v))

    def math_eval(self, fn, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, dtype=np.float64)
        out = np.zeros_like(x)
        self.lib.or_math_eval(self.MATH_FN[fn], _dp(x), _dp(y), _dp(out), len(x))
        return out

    # --- kerr.hpp -------------------------------------------------------------------------------
    def kerr_point(self, a, r, th):
        o = np.zeros(11)
        self.lib.or_kerr_point(a, r, th, _dp(o))
        return o

    def camera_photon(self, a, kp, phi0, t0, d, beta=(0.0, 0.0, 0.0)):
        kp = np.ascontiguousarray(kp, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        b = np.ascontiguousarray(beta, dtype=np.float64)
        o = np.zeros(9)
        self.lib.or_camera_photon(a, _dp(kp), phi0, t0, _dp(d), _dp(b), _dp(o))
        return o

    def geodesic_rhs(self, a, lam, eta, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        d = np.zeros(6)
        self.lib.or_geodesic_rhs(a, lam, eta, _dp(y), _dp(d))
        return d

    def dp54_step(self, a, lam, eta, y, k1, h, tol):
        y = np.ascontiguousarray(y, dtype=np.float64)
        k1 = np.ascontiguousarray(k1, dtype=np.float64)
        yo = np.zeros(6)
        k7 = np.zeros(6)
        e = self.lib.or_dp54_step(a, lam, eta, _dp(y), _dp(k1), h, tol, _dp(yo), _dp(k7))
        return e, yo, k7

    def project_radial(self, a, lam, eta, y, k):
        y = np.array(y, dtype=np.float64)
        k = np.array(k, dtype=np.float64)
        self.lib.or_project_radial(a, lam, eta, _dp(y), _dp(k))
        return y, k

    def vacuum_fate(self, a, photon, tol, r_escape, max_steps):
        p = np.ascontiguousarray(photon, dtype=np.float64)
        return self.lib.or_vacuum_fate(a, _dp(p), tol, r_escape, max_steps)

    def critical_curve_point(self, a, theta_o, rp):
        al = C.c_double()
        be = C.c_double()
        ok = self.lib.or_critical_curve_point(a, theta_o, rp, C.byref(al), C.byref(be))
        return (al.value, be.value) if ok else None

    # --- shot.cpp -------------------------------------------------------------------------------
    def make_frame(self, a, r, incl, phi=0.0, roll=0.0, yaw=0.0, pitch=0.0, fov=30.0, fov_in_m=True, zoom=1.0,
                   target=0.0, target_x=0.0, target_y=0.0, off_x=0.0, off_y=0.0, time_m=0.0, orbit=False,
                   W=1, H=None, seed=0):
        F = OrFrame()
        self.lib.or_make_frame(a, r, incl, phi, roll, yaw, pitch, fov, int(fov_in_m), zoom, target, target_x,
                               target_y, off_x, off_y, time_m, int(orbit), W, W if H is None else H, seed,
                               C.byref(F))
        return F

    def find_critical_point(self, a, F, psi):
        return self.lib.or_find_critical_point(a, C.byref(F), psi)

    def selftest(self):
        rows = np.zeros((8, 6))
        drift = C.c_double()
        steps = C.c_int()
        self.lib.or_selftest(_dp(rows), C.byref(drift), C.byref(steps))
        return rows, drift.value, steps.value

    # --- sampling / tables ----------------------------------------------------------------------
    def k2_table(self, n=512):
        t = np.zeros(n, dtype=np.float32)
        l0 = C.c_float()
        dl = C.c_float()
        self.lib.or_k2_table(n, t.ctypes.data_as(_F), C.byref(l0), C.byref(dl))
        return t, l0.value, dl.value

    # --- tracing --------------------------------------------------------------------------------
    def trace_ray(self, S, F, x, y):
        o = OrRay()
        self.lib.or_trace_ray(C.byref(S), C.byref(F), x, y, C.byref(o))
        return o

    def trace_pixels(self, S, F, px, py, spp=1, jitter=False, nthreads=0, want_rays=True):
        px = np.ascontiguousarray(px, dtype=np.int32)
        py = np.ascontiguousarray(py, dtype=np.int32)
        n = len(px)
        out_px = np.zeros((n, 4))
        rays = np.zeros(n, dtype=RAY_DTYPE) if (want_rays and spp == 1) else None
        used = self.lib.or_trace_pixels(C.byref(S), C.byref(F), px.ctypes.data_as(_I), py.ctypes.data_as(_I), n,
                                        spp, int(jitter), nthreads, _dp(out_px),
                                        rays.ctypes.data if rays is not None else None)
        return out_px, rays, used

    def turb_factor(self, S, r, u, phi, t=0.0):
        return self.lib.or_turb_factor(C.byref(S), r, u, phi, t)

    def render_adaptive(self, S, F, spp, spp_max, target_err, jitter=True, nthreads=0):
        """(image H x W x 4 float64, samples per pixel H x W, total samples)."""
        W, H = F.W, F.H
        img = np.zeros((H, W, 4), dtype=np.float64)
        ns = np.zeros((H, W), dtype=np.uint32)
        total = self.lib.or_render_adaptive(C.byref(S), C.byref(F), spp, spp_max, C.c_float(target_err), int(jitter),
                                            nthreads, img.ctypes.data_as(_D),
                                            ns.ctypes.data_as(C.POINTER(C.c_uint32)))
        return img, ns, total

    # --- sheared turbulence (sheared.cpp) ------------------------------------------------------
    def sheared_params(self, S, cam_t, cam_r, dims=(0, 0, 0, 0)):
        B = OrSheared()
        d = (C.c_int * 4)(*[int(v) for v in dims])
        self.lib.or_sheared_params(C.byref(S), cam_t, cam_r, d, C.byref(B))
        return B

    def sheared_bake(self, B, nthreads=0):
        """nt x nz x nlr x nphi float32 slices."""
        out = np.zeros((B.nt, B.nz, B.nlr, B.nphi), dtype=np.float32)
        self.lib.or_sheared_bake(C.byref(B), out.ctypes.data_as(_F), nthreads)
        return out

    def attach_sheared(self, S, B, slices):
        """Makes S read the baked slices (turb_model 1)."""
        slices = np.ascontiguousarray(slices, dtype=np.float32)
        S.turb_model = 1
        S.ts = B
        S.ts_slices = slices.ctypes.data_as(_F)
        S._ts_keep = slices

    def gnoise(self, x, y, z, seed):
        return self.lib.or_sheared_gnoise(x, y, z, seed)

    def fbm(self, x, y, z, octaves, seed):
        return self.lib.or_sheared_fbm(x, y, z, octaves, seed)

    def sheared_noise(self, B, R, phi, z, t):
        return self.lib.or_sheared_noise(C.byref(B), R, phi, z, t)

    def sheared_tex(self, T, x, y, z):
        T = np.ascontiguousarray(T, dtype=np.float32)
        nz, ny, nx = T.shape
        return self.lib.or_sheared_tex(T.ctypes.data_as(_F), nx, ny, nz, x, y, z)

    def sheared_cell(self, B, i, j, l, k):
        return self.lib.or_sheared_cell(C.byref(B), i, j, l, k)

    def sheared_factor(self, S, r, R, z, phi, t):
        return self.lib.or_sheared_factor(C.byref(S), r, R, z, phi, t)

    # --- radio post step and tone map (post.cpp) -----------------------------------------------
    def post_blur(self, img, sigma):
        """Beam blur of an H x W x 4 float32 image (a new array)."""
        out = np.ascontiguousarray(img, dtype=np.float32).copy()
        H, W = out.shape[:2]
        self.lib.or_post_blur(out.ctypes.data_as(_F), W, H, sigma)
        return out

    def frame_stats(self, img, hide_direct=0.0, percentile=0.995):
        px = np.ascontiguousarray(img, dtype=np.float32)
        H, W = px.shape[:2]
        lp, lm = C.c_double(), C.c_double()
        self.lib.or_frame_stats(px.ctypes.data_as(_F), W, H, hide_direct, percentile, C.byref(lp), C.byref(lm))
        return lp.value, lm.value

    def tonemap_radio(self, img, colormap=0, exposure=1.0, gamma=1.0, overlay=0.0, hide_direct=0.0, fade=1.0,
                      vignette=0.0, grain=0.0, seed=0):
        px = np.ascontiguousarray(img, dtype=np.float32)
        H, W = px.shape[:2]
        rgb = np.zeros((H, W, 3), dtype=np.uint8)
        self.lib.or_tonemap_radio(px.ctypes.data_as(_F), W, H, colormap, exposure, gamma, overlay, hide_direct, fade,
                                  vignette, grain, seed, rgb.ctypes.data_as(C.POINTER(C.c_uint8)))
        return rgb

    def hud_draw(self, rgb, glyphs, rects, atlas, off):
        """HUD over an H x W x 3 uint8 image (copied): glyphs / rects as sequences of (x, y, scale, glyph,
        rgba) / (x0, y0, x1, y1, rgba, layer); atlas glyphs x cell_h x cell_w uint8."""
        out = np.ascontiguousarray(rgb, dtype=np.uint8).copy()
        H, W = out.shape[:2]
        g = np.array([tuple(v) for v in glyphs], dtype=[("x", "<f4"), ("y", "<f4"), ("s", "<f4"), ("g", "<i4"), ("c", "<u4")])
        r = np.array([tuple(v) for v in rects], dtype=[("x0", "<i4"), ("y0", "<i4"), ("x1", "<i4"), ("y1", "<i4"),
                                                        ("c", "<u4"), ("l", "<i4")])
        at = np.ascontiguousarray(atlas, dtype=np.uint8)
        self.lib.or_hud_draw(out.ctypes.data_as(C.POINTER(C.c_uint8)), W, H, g.ctypes.data_as(C.c_void_p), len(g),
                             r.ctypes.data_as(C.c_void_p), len(r), at.ctypes.data_as(C.POINTER(C.c_uint8)),
                             at.shape[2], at.shape[1], C.c_float(off))
        return out

    def flow_step(self, S, r, u, phi, t, lam, Ecam, h):
        dI = C.c_double()
        dt = C.c_double()
        self.lib.or_flow_step(C.byref(S), r, u, phi, t, lam, Ecam, h, C.byref(dI), C.byref(dt))
        return dI.value, dt.value


def make_scene(p, k2=None, grid=None, turb_half=60.0, turb_t0=0.0, turb_dt=0.0):
    """OrScene from a dict of scene scalars (keys as in the struct). grid: n^3, or nt x n^3 time
    slices spaced turb_dt from turb_t0. Keeps references to the arrays on the returned object so
    they outlive the call."""
    S = OrScene()
    for name, _ in OrScene._fields_:
        if name in p:
            setattr(S, name, p[name])
    keep = []
    if k2 is not None:
        k2 = np.ascontiguousarray(k2, dtype=np.float32)
        keep.append(k2)
        S.k2_lntab = k2.ctypes.data_as(_F)
        S.k2_n = len(k2)
    if grid is not None:
        grid = np.ascontiguousarray(grid, dtype=np.float32)
        keep.append(grid)
        n = grid.shape[-1]
        S.turb_grid = grid.ctypes.data_as(_F)
        S.turb_n = n
        S.turb_half = turb_half
        S.turb_inv_dx = n / (2.0 * turb_half)
        S.turb_nt = grid.shape[0] if grid.ndim == 4 else 1
        S.turb_t0 = turb_t0
        S.turb_inv_dt = 1.0 / turb_dt if turb_dt > 0 else 0.0
    S._keep = keep
    return S


class Ref:
    """The reference's own host code (oracle/_ref/libref.so, built from the reference sources)."""

    def __init__(self):
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        L.ref_kerr_point.argtypes = [C.c_double, C.c_double, C.c_double, _D]
        L.ref_horizon_radius.restype = C.c_double
        L.ref_horizon_radius.argtypes = [C.c_double]
        L.ref_isco_radius.restype = C.c_double
        L.ref_isco_radius.argtypes = [C.c_double]
        L.ref_camera_photon.argtypes = [C.c_double, _D, C.c_double, C.c_double, _D, _D, _D]
        L.ref_radial_potential.restype = C.c_double
        L.ref_radial_potential.argtypes = [C.c_double] * 4
        L.ref_geodesic_rhs.argtypes = [C.c_double, C.c_double, C.c_double, _D, _D]
        L.ref_dp54_step.restype = C.c_double
        L.ref_dp54_step.argtypes = [C.c_double, C.c_double, C.c_double, _D, _D, C.c_double, C.c_double, _D, _D]
        L.ref_project_radial.argtypes = [C.c_double, C.c_double, C.c_double, _D, _D]
        L.ref_initial_step.restype = C.c_double
        L.ref_initial_step.argtypes = [_D]
        L.ref_vacuum_fate.restype = C.c_int
        L.ref_vacuum_fate.argtypes = [C.c_double, _D, C.c_double, C.c_double, C.c_int]
        L.ref_critical_curve_point.restype = C.c_int
        L.ref_critical_curve_point.argtypes = [C.c_double] * 3 + [_D, _D]
        L.ref_build_scene.restype = C.c_int
        L.ref_build_scene.argtypes = [C.c_char_p, C.c_char_p, _D, _F]
        L.ref_build_shot.restype = C.c_int
        L.ref_build_shot.argtypes = [C.c_char_p, C.c_char_p, _D, C.c_int]
        L.ref_config_str.restype = C.c_int
        L.ref_config_str.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
        L.ref_make_frame.argtypes = [C.c_double] * 8 + [C.c_int] + [C.c_double] * 7 + [C.c_int, C.c_int, _D]
        L.ref_find_critical_point.restype = C.c_double
        L.ref_find_critical_point.argtypes = [C.c_double] * 8

    def kerr_point(self, a, r, th):
        o = np.zeros(11)
        self.lib.ref_kerr_point(a, r, th, _dp(o))
        return o

    def camera_photon(self, a, kp, phi0, t0, d, beta=(0.0, 0.0, 0.0)):
        kp = np.ascontiguousarray(kp, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        b = np.ascontiguousarray(beta, dtype=np.float64)
        o = np.zeros(9)
        self.lib.ref_camera_photon(a, _dp(kp), phi0, t0, _dp(d), _dp(b), _dp(o))
        return o

    def geodesic_rhs(self, a, lam, eta, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        d = np.zeros(6)
        self.lib.ref_geodesic_rhs(a, lam, eta, _dp(y), _dp(d))
        return d

    def dp54_step(self, a, lam, eta, y, k1, h, tol):
        y = np.ascontiguousarray(y, dtype=np.float64)
        k1 = np.ascontiguousarray(k1, dtype=np.float64)
        yo = np.zeros(6)
        k7 = np.zeros(6)
        e = self.lib.ref_dp54_step(a, lam, eta, _dp(y), _dp(k1), h, tol, _dp(yo), _dp(k7))
        return e, yo, k7

    def project_radial(self, a, lam, eta, y, k):
        y = np.array(y, dtype=np.float64)
        k = np.array(k, dtype=np.float64)
        self.lib.ref_project_radial(a, lam, eta, _dp(y), _dp(k))
        return y, k

    def vacuum_fate(self, a, photon, tol, r_escape, max_steps):
        p = np.ascontiguousarray(photon, dtype=np.float64)
        return self.lib.ref_vacuum_fate(a, _dp(p), tol, r_escape, max_steps)

    def build_scene(self, path, sets=""):
        out = np.zeros(31)
        k2 = np.zeros(512, dtype=np.float32)
        ok = self.lib.ref_build_scene(path.encode(), sets.encode(), _dp(out), k2.ctypes.data_as(_F))
        if not ok:
            raise RuntimeError(f"reference build_scene failed for {path}")
        return out, k2

    def sky_rot(self, path, sets=""):
        """(sky_rot row-major 9 values, sky_scale) of the reference's build_scene."""
        out = np.zeros(10)
        if not self.lib.ref_sky_rot(path.encode(), sets.encode(), _dp(out)):
            raise RuntimeError(f"reference build_scene failed for {path}")
        return out[:9], out[9]

    def build_shot(self, path, sets="", maxn=4096):
        out = np.zeros((maxn, 20))
        n = self.lib.ref_build_shot(path.encode(), sets.encode(), _dp(out), maxn)
        return out[:max(n, 0)]

    def config_str(self, path, key, sets=""):
        buf = C.create_string_buffer(4096)
        r = self.lib.ref_config_str(path.encode(), sets.encode(), key.encode(), buf, 4096)
        return buf.value.decode() if r == 1 else None

    def make_frame(self, a, r, incl, phi=0.0, roll=0.0, yaw=0.0, pitch=0.0, fov=30.0, fov_in_m=True, zoom=1.0,
                   target=0.0, target_x=0.0, target_y=0.0, off_x=0.0, off_y=0.0, time_m=0.0, orbit=False, W=1):
        o = np.zeros(31)
        self.lib.ref_make_frame(a, r, incl, phi, roll, yaw, pitch, fov, int(fov_in_m), zoom, target, target_x,
                                target_y, off_x, off_y, time_m, int(orbit), W, _dp(o))
        return o

    def find_critical_point(self, a, r, incl, phi, psi, roll=0.0, yaw=0.0, pitch=0.0):
        return self.lib.ref_find_critical_point(a, r, incl, phi, roll, yaw, pitch, psi)
