"""ctypes binding of libbhrt.so (include/bhrt.h and include/bhrt_host.h).

Plumbing for tests and the benchmark: every call goes to the HIP library; there is no Python or CPU
fallback, and loading fails loudly when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbhrt.so")

BH_OK, BH_ERR_ARG, BH_ERR_HIP, BH_ERR_NODEVICE, BH_ERR_UNSUPPORTED = 0, -1, -2, -3, -4
MODE_VISIBLE, MODE_RADIO = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
ORDER_ROWS, ORDER_IMPACT, ORDER_CRITICAL = 1, 2, 3
TURB_GRID, TURB_SHEARED = 0, 1
EMIS_EXACT, EMIS_MLP = 0, 1
ENGINE_AUTO, ENGINE_FRAME = 0, 1     # bh_render: persistent engine for one sample per pixel / frame kernel

# Every symbol declared in include/bhrt.h and include/bhrt_host.h
EXPORTS = [
    "bh_device_count", "bh_create", "bh_update_scene", "bh_get_scene", "bh_render", "bh_render_resident",
    "bh_trace_tile", "bh_radio_flux", "bh_vacuum_fates", "bh_math_eval", "bh_destroy", "bh_last_error",
    "bhh_config_str", "bhh_build_scene", "bhh_turbulence", "bhh_build_shot", "bhh_make_frame",
    "bhh_create_from_ini", "bhh_find_critical_point", "bhh_selftest", "bhh_write_png", "bhh_smooth_exposure",
    "bh_mlp_train", "bh_mlp_get", "bh_mlp_set", "bh_mlp_infer", "bh_mlp_exact", "bh_mlp_grad", "bh_mlp_bake",
    "bh_trace_batch", "bh_autotune", "bh_render_adaptive", "bh_post_radio", "bh_tonemap", "bh_sheared_slices",
    "bh_tonemap_hud", "bhh_make_hud", "bhh_save_font", "bhh_abi_sizes", "bh_render_noalloc", "bhh_abi_offsets",
]
MLP_NPARAM = 8706
SHOT_VALUES = 20   # values per frame of bhh_build_shot
MLP_NNORM = 10   # input centres[3], half-widths[3], output means[2], scales[2]


class Scene(C.Structure):
    _fields_ = [
        ("a", C.c_double), ("rh", C.c_double), ("r_isco", C.c_double), ("tol", C.c_double),
        ("max_steps", C.c_int), ("r_escape", C.c_double), ("r_stop", C.c_double), ("step_frac", C.c_double),
        ("mode", C.c_int), ("flow_on", C.c_int),
        ("n0", C.c_double), ("n_index", C.c_double), ("theta_index", C.c_double), ("theta0", C.c_double),
        ("h_r", C.c_double), ("beta_p", C.c_double), ("r_flow_max", C.c_double), ("nu_obs", C.c_double),
        ("rg_cm", C.c_double), ("flow_kep", C.c_double), ("flow_turb_amp", C.c_float), ("seed", C.c_uint32),
        ("flow_turb_scale", C.c_float), ("flow_tcorr", C.c_float),
        ("sky_on", C.c_int), ("sky_nu", C.c_int), ("sky_nv", C.c_int), ("sky_scale", C.c_float),
        ("sky_rot", C.c_double * 9),
    ]


class Tables(C.Structure):
    _fields_ = [
        ("k2_ln", C.POINTER(C.c_float)), ("k2_n", C.c_int), ("k2_l0", C.c_float), ("k2_dl", C.c_float),
        ("turb", C.POINTER(C.c_float)), ("turb_n", C.c_int), ("turb_nt", C.c_int),
        ("turb_half", C.c_double), ("turb_t0", C.c_double), ("turb_dt", C.c_double),
        ("turb_model", C.c_int), ("ts_dims", C.c_int * 4),
    ]


class Frame(C.Structure):
    _fields_ = [
        ("r", C.c_double), ("th", C.c_double), ("ph", C.c_double), ("t", C.c_double),
        ("kp", C.c_double * 11), ("F", C.c_double * 3), ("R", C.c_double * 3), ("U", C.c_double * 3),
        ("beta", C.c_double * 3), ("cx", C.c_double), ("cy", C.c_double), ("pix", C.c_double),
        ("W", C.c_int), ("H", C.c_int), ("spp", C.c_int), ("seed", C.c_uint32),
    ]


class Settings(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("spp", C.c_int), ("spp_max", C.c_int), ("target_err", C.c_float)]


class TraceOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("emissivity", C.c_int), ("jitter", C.c_int), ("spp", C.c_int),
                ("tol_far", C.c_double), ("r_near", C.c_double), ("r_emit", C.c_double), ("far_fp32", C.c_int),
                ("full_flow", C.c_int), ("engine", C.c_int), ("tol_t", C.c_double), ("relax_after", C.c_int),
                ("tol_relax", C.c_double), ("relax_cap", C.c_double)]


class Stats(C.Structure):
    _fields_ = [
        ("ms_trace", C.c_double), ("ms_kernel", C.c_double), ("samples", C.c_longlong),
        ("step_limited", C.c_longlong), ("step_attempts", C.c_longlong), ("steps_accepted", C.c_longlong),
        ("flow_steps", C.c_longlong), ("fp64_steps", C.c_longlong), ("flux_sum", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchFrame(C.Structure):
    _fields_ = [("scene", Scene), ("frame", Frame)]


class BatchCfg(C.Structure):
    _fields_ = [("block", C.c_int), ("grid_per_cu", C.c_int), ("interval", C.c_int), ("refill_min", C.c_int),
                ("use_graph", C.c_int), ("waves", C.c_int), ("order", C.c_int), ("prio_age", C.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class HudGlyph(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("scale", C.c_float), ("glyph", C.c_int), ("rgba", C.c_uint32)]


class HudRect(C.Structure):
    _fields_ = [("x0", C.c_int), ("y0", C.c_int), ("x1", C.c_int), ("y1", C.c_int), ("rgba", C.c_uint32),
                ("layer", C.c_int)]


class Font(C.Structure):
    _fields_ = [("atlas", C.POINTER(C.c_uint8)), ("cell_w", C.c_int), ("cell_h", C.c_int), ("glyphs", C.c_int),
                ("advance", C.c_float), ("baseline", C.c_int)]


def hud_arrays(glyphs, rects):
    """ctypes arrays of HudGlyph / HudRect from sequences of tuples (x, y, scale, glyph, rgba) and
    (x0, y0, x1, y1, rgba, layer)."""
    ga = (HudGlyph * max(len(glyphs), 1))(*[HudGlyph(*g) for g in glyphs])
    ra = (HudRect * max(len(rects), 1))(*[HudRect(*r) for r in rects])
    return ga, ra


class PostParams(C.Structure):
    _fields_ = [("blur_sigma_px", C.c_double), ("hide_direct", C.c_float)]


class ToneParams(C.Structure):
    _fields_ = [("colormap", C.c_int), ("exposure", C.c_float), ("gamma", C.c_float), ("overlay", C.c_float),
                ("hide_direct", C.c_float), ("fade", C.c_float), ("vignette", C.c_float), ("grain", C.c_float),
                ("seed", C.c_uint32)]


class MlpTrainCfg(C.Structure):
    _fields_ = [("n_samples", C.c_int), ("batch", C.c_int), ("epochs", C.c_int), ("lr", C.c_float),
                ("lr_min", C.c_float), ("hot_weight", C.c_float), ("seed", C.c_uint64)]


class MlpReport(C.Structure):
    _fields_ = [("ms_sample", C.c_double), ("ms_train", C.c_double), ("steps", C.c_int), ("loss_first", C.c_float),
                ("loss_last", C.c_float), ("train_flop", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class MlpBakeReport(C.Structure):
    _fields_ = [("ms_bake", C.c_double), ("ms_judge", C.c_double), ("cells", C.c_longlong),
                ("hot_cells", C.c_longlong), ("failed_hot_cells", C.c_longlong)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


RAY_DTYPE = np.dtype([
    ("y", "<f8", (6,)), ("th_inf", "<f8"), ("ph_inf", "<f8"), ("I", "<f8", (4,)), ("trans", "<f8"),
    ("psi", "<f8"), ("fate", "<i4"), ("ncross", "<i4"), ("order", "<i4"), ("steps", "<i4"),
    ("accepted", "<i4"), ("flow_steps", "<i4"), ("sky_u", "<f4"), ("sky_v", "<f4"),
])

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_lib = None


def load():
    """Loads libbhrt.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() or make -C spacetime-optics_amd")
    L = C.CDLL(LIB_PATH)
    L.bh_device_count.restype = C.c_int
    L.bh_create.restype = C.c_int
    L.bh_create.argtypes = [C.c_int, C.POINTER(Scene), C.POINTER(Tables), C.POINTER(Settings), C.POINTER(C.c_void_p),
                            C.c_char_p, C.c_size_t]
    L.bh_update_scene.restype = C.c_int
    L.bh_update_scene.argtypes = [C.c_void_p, C.POINTER(Scene)]
    L.bh_get_scene.restype = C.c_int
    L.bh_get_scene.argtypes = [C.c_void_p, C.POINTER(Scene)]
    L.bh_render.restype = C.c_int
    L.bh_render.argtypes = [C.c_void_p, C.POINTER(Frame), C.POINTER(TraceOpts), _F, C.POINTER(Stats)]
    L.bh_render_resident.restype = C.c_int
    L.bh_render_resident.argtypes = [C.c_void_p, C.POINTER(Frame), C.POINTER(TraceOpts), C.POINTER(C.c_void_p),
                                     C.POINTER(Stats)]
    L.bh_render_noalloc.restype = C.c_int
    L.bh_render_noalloc.argtypes = [C.c_void_p, C.POINTER(Frame), C.POINTER(TraceOpts), C.c_void_p, C.POINTER(Stats)]
    L.bh_trace_tile.restype = C.c_int
    L.bh_trace_tile.argtypes = [C.c_void_p, C.POINTER(Frame), C.c_int, C.c_int, C.c_int, C.c_int,
                                C.POINTER(TraceOpts), C.c_void_p]
    L.bh_radio_flux.restype = C.c_int
    L.bh_radio_flux.argtypes = [C.c_void_p, C.POINTER(Frame), C.c_int, C.c_double, _D]
    L.bh_vacuum_fates.restype = C.c_int
    L.bh_vacuum_fates.argtypes = [C.c_void_p, _D, C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_int)]
    L.bh_mlp_train.restype = C.c_int
    L.bh_mlp_train.argtypes = [C.c_void_p, C.POINTER(MlpTrainCfg), C.POINTER(MlpReport)]
    L.bh_mlp_get.restype = C.c_int
    L.bh_mlp_get.argtypes = [C.c_void_p, _F, _F]
    L.bh_mlp_set.restype = C.c_int
    L.bh_mlp_set.argtypes = [C.c_void_p, _F, _F]
    L.bh_mlp_infer.restype = C.c_int
    L.bh_mlp_infer.argtypes = [C.c_void_p, _F, C.c_int, _F]
    L.bh_mlp_exact.restype = C.c_int
    L.bh_mlp_exact.argtypes = [C.c_void_p, _F, C.c_int, _D]
    L.bh_mlp_grad.restype = C.c_int
    L.bh_mlp_grad.argtypes = [C.c_void_p, _F, _F, _F, _F, _F, C.c_int, _F, _F]
    L.bh_mlp_bake.restype = C.c_int
    L.bh_mlp_bake.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.POINTER(MlpBakeReport)]
    L.bh_render_adaptive.restype = C.c_int
    L.bh_render_adaptive.argtypes = [C.c_void_p, C.POINTER(Frame), C.POINTER(TraceOpts), _F, C.POINTER(C.c_uint32),
                                     C.POINTER(Stats)]
    L.bh_trace_batch.restype = C.c_int
    L.bh_trace_batch.argtypes = [C.c_void_p, C.POINTER(BatchFrame), C.c_int, C.POINTER(TraceOpts),
                                 C.POINTER(BatchCfg), _F, C.POINTER(Stats)]
    L.bh_autotune.restype = C.c_int
    L.bh_autotune.argtypes = [C.c_void_p, C.POINTER(BatchFrame), C.c_int, C.POINTER(TraceOpts), C.POINTER(BatchCfg),
                              _D, _D, C.c_int, C.POINTER(C.c_int)]
    L.bh_sheared_slices.restype = C.c_int
    L.bh_sheared_slices.argtypes = [C.c_void_p, C.POINTER(Frame), _F, C.c_size_t]
    L.bh_post_radio.restype = C.c_int
    L.bh_post_radio.argtypes = [C.c_void_p, _F, C.c_int, C.c_int, C.POINTER(PostParams), C.c_double, _D, _D]
    L.bh_tonemap.restype = C.c_int
    L.bh_tonemap.argtypes = [C.c_void_p, _F, C.c_int, C.c_int, C.POINTER(ToneParams), C.POINTER(C.c_uint8)]
    L.bh_tonemap_hud.restype = C.c_int
    L.bh_tonemap_hud.argtypes = [C.c_void_p, _F, C.c_int, C.c_int, C.POINTER(ToneParams), C.POINTER(HudGlyph), C.c_int,
                                 C.POINTER(HudRect), C.c_int, C.POINTER(Font), C.POINTER(C.c_uint8)]
    L.bh_math_eval.restype = C.c_int
    L.bh_math_eval.argtypes = [C.c_int, _D, _D, _D, C.c_int]
    L.bh_destroy.argtypes = [C.c_void_p]
    L.bh_last_error.restype = C.c_char_p
    L.bh_last_error.argtypes = [C.c_void_p]
    L.bhh_config_str.restype = C.c_int
    L.bhh_config_str.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
    L.bhh_build_scene.restype = C.c_int
    L.bhh_build_scene.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(Scene), _F, _F, _F, _D, C.POINTER(TraceOpts),
                                  C.c_char_p, C.c_size_t]
    L.bhh_turbulence.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_uint32, _F]
    L.bhh_build_shot.restype = C.c_int
    L.bhh_build_shot.argtypes = [C.c_char_p, C.c_char_p, _D, C.c_int]
    L.bhh_make_frame.argtypes = ([C.c_double] * 8 + [C.c_int] + [C.c_double] * 7 + [C.c_int] * 4 + [C.c_uint32,
                                 C.POINTER(Frame)])
    L.bhh_create_from_ini.restype = C.c_int
    L.bhh_create_from_ini.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p),
                                      C.POINTER(Scene), C.POINTER(Frame), C.POINTER(TraceOpts), _D, C.c_char_p,
                                      C.c_size_t]
    L.bhh_find_critical_point.restype = C.c_int
    L.bhh_find_critical_point.argtypes = [C.c_void_p, C.POINTER(Frame), C.c_double, _D]
    L.bhh_write_png.restype = C.c_int
    L.bhh_write_png.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint8), C.c_int]
    L.bhh_smooth_exposure.argtypes = [_D, C.c_int, C.c_double, C.c_double, _D]
    L.bhh_make_hud.restype = C.c_int
    L.bhh_make_hud.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                               C.POINTER(HudGlyph), C.c_int, C.POINTER(C.c_int), C.POINTER(HudRect), C.c_int,
                               C.POINTER(C.c_int), C.c_char_p, C.c_size_t]
    L.bhh_save_font.restype = C.c_int
    L.bhh_save_font.argtypes = [C.c_char_p, C.POINTER(Font)]
    L.bhh_abi_sizes.restype = C.c_int
    L.bhh_abi_sizes.argtypes = [C.POINTER(C.c_size_t), C.c_int]
    L.bhh_abi_offsets.restype = C.c_int
    L.bhh_abi_offsets.argtypes = [C.POINTER(C.c_size_t), C.c_int]
    L.bhh_selftest.restype = C.c_int
    L.bhh_selftest.argtypes = [C.c_int, _D, C.c_char_p, C.c_size_t]
    _lib = L
    return L


class BhError(RuntimeError):
    pass


def device_count():
    return load().bh_device_count()


# --- host layer (no GPU needed) ------------------------------------------------------------------
def config_str(ini, key, sets=""):
    buf = C.create_string_buffer(4096)
    r = load().bhh_config_str(ini.encode(), sets.encode(), key.encode(), buf, 4096)
    if r < 0:
        raise BhError(buf.value.decode())
    return buf.value.decode() if r == 1 else None


def build_scene(ini, sets=""):
    """(Scene, k2 table, (l0, dl), info dict, TraceOpts) of a shot .ini."""
    S = Scene()
    k2 = np.zeros(512, dtype=np.float32)
    l0, dl = C.c_float(), C.c_float()
    info = np.zeros(6)
    opts = TraceOpts()
    err = C.create_string_buffer(512)
    rc = load().bhh_build_scene(ini.encode(), sets.encode(), C.byref(S), k2.ctypes.data_as(_F), C.byref(l0),
                                C.byref(dl), info.ctypes.data_as(_D), C.byref(opts), err, 512)
    if rc != BH_OK:
        raise BhError(err.value.decode())
    keys = ["mass_msun", "distance_cm", "rg_cm", "uas_per_m", "flux_target_jy", "normalize_fov_m"]
    return S, k2, (l0.value, dl.value), dict(zip(keys, info)), opts


def turbulence(n=128, nt=1, half=60.0, sigma=0.3, corr_m=5.0, corr_t=30.0, dt=10.0, seed=1):
    out = np.zeros((nt, n, n, n), dtype=np.float32)
    load().bhh_turbulence(n, nt, half, sigma, corr_m, corr_t, dt, seed, out.ctypes.data_as(_F))
    return out


def build_shot(ini, sets="", maxn=4096):
    out = np.zeros((maxn, SHOT_VALUES))
    n = load().bhh_build_shot(ini.encode(), sets.encode(), out.ctypes.data_as(_D), maxn)
    if n < 0:
        raise BhError(f"build_shot failed for {ini}")
    return out[:n]


def write_png(path, img, level=4):
    """8-bit PNG of an H x W (grey), H x W x 3 or H x W x 4 uint8 array."""
    a = np.ascontiguousarray(img, dtype=np.uint8)
    ch = 1 if a.ndim == 2 else a.shape[2]
    if load().bhh_write_png(str(path).encode(), a.shape[1], a.shape[0], ch, a.ctypes.data_as(C.POINTER(C.c_uint8)),
                            level) != 0:
        raise BhError(f"cannot write {path}")


def make_hud(ini, sets="", frame=0, W=1920, H=1080, cell_h=44, advance=21.6):
    """HUD layout of one frame of a shot (bhh_make_hud): (glyphs, rects) as lists of (x, y, scale,
    glyph, rgba) and (x0, y0, x1, y1, rgba, layer)."""
    L = load()
    ng, nr = C.c_int(), C.c_int()
    err = C.create_string_buffer(512)
    cap = 4096
    ga, ra = (HudGlyph * cap)(), (HudRect * 1024)()
    if L.bhh_make_hud(str(ini).encode(), sets.encode(), frame, W, H, cell_h, advance, ga, cap, C.byref(ng), ra, 1024,
                      C.byref(nr), err, 512) != 0:
        raise BhError(f"bhh_make_hud: {err.value.decode()}")
    glyphs = [(g.x, g.y, g.scale, g.glyph, g.rgba) for g in ga[:min(ng.value, cap)]]
    rects = [(r.x0, r.y0, r.x1, r.y1, r.rgba, r.layer) for r in ra[:min(nr.value, 1024)]]
    return glyphs, rects


def save_font(path, atlas, advance=None, baseline=0):
    """Writes a HUD font file from a glyphs x cell_h x cell_w uint8 coverage atlas (glyph g = ASCII
    32 + g for g < 95, then U+00B5, U+00B0, U+00D7, U+00B1); advance defaults to 0.9 cell widths."""
    at = np.ascontiguousarray(atlas, dtype=np.uint8)
    n, ch, cw = at.shape
    adv = 0.9 * cw if advance is None else advance
    f = Font(at.ctypes.data_as(C.POINTER(C.c_uint8)), cw, ch, n, adv, baseline)
    if load().bhh_save_font(str(path).encode(), C.byref(f)) != 0:
        raise BhError(f"cannot write {path}")


def font_from_header(path):
    """(atlas, baseline) from a C header in the layout of bhrt's common/font_atlas.h: FONT_CELL_W,
    FONT_CELL_H, FONT_GLYPHS, FONT_BASELINE constants and the FONT_ATLAS byte list. For converting a
    bhrt install's own font (bhrt lays its text out with 21.6 px per character at scale 1)."""
    import re
    text = open(path).read()
    dim = {k: int(re.search(rf"{k}\s*=\s*(\d+)", text).group(1))
           for k in ("FONT_CELL_W", "FONT_CELL_H", "FONT_GLYPHS", "FONT_BASELINE")}
    body = text[text.index("FONT_ATLAS"):]
    body = body[body.index("{") + 1:body.index("}")]
    vals = np.array([int(v) for v in re.findall(r"\d+", body)], dtype=np.uint8)
    n = dim["FONT_GLYPHS"] * dim["FONT_CELL_H"] * dim["FONT_CELL_W"]
    if vals.size != n:
        raise BhError(f"{path}: {vals.size} atlas bytes, expected {n}")
    return vals.reshape(dim["FONT_GLYPHS"], dim["FONT_CELL_H"], dim["FONT_CELL_W"]), dim["FONT_BASELINE"]


def smooth_exposure(lp, key, sigma):
    lp = np.ascontiguousarray(lp, dtype=np.float64)
    out = np.zeros_like(lp)
    load().bhh_smooth_exposure(lp.ctypes.data_as(_D), len(lp), key, sigma, out.ctypes.data_as(_D))
    return out


def make_frame(a, r, incl, phi=0.0, roll=0.0, yaw=0.0, pitch=0.0, fov=30.0, fov_in_m=True, zoom=1.0, target=0.0,
               target_x=0.0, target_y=0.0, off_x=0.0, off_y=0.0, time_m=0.0, orbit=False, W=1, H=None, spp=1, seed=0):
    F = Frame()
    load().bhh_make_frame(a, r, incl, phi, roll, yaw, pitch, fov, int(fov_in_m), zoom, target, target_x, target_y,
                          off_x, off_y, time_m, int(orbit), W, W if H is None else H, spp, seed, C.byref(F))
    return F


def k2_from_scene(ini, sets=""):
    _, k2, (l0, dl), _, _ = build_scene(ini, sets)
    return k2, l0, dl


MATH_FN = {"exp": 0, "log": 1, "expm1": 2, "cbrt": 3, "pow_m02": 4, "sin": 5, "cos": 6, "atan2": 7, "acos": 8,
           "sqrt": 9, "div": 10, "asin": 11}


def math_eval(fn, x, y=None):
    """Device evaluation of one bh_math function (diagnostic entry bh_math_eval)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, dtype=np.float64)
    out = np.zeros_like(x)
    rc = load().bh_math_eval(MATH_FN[fn], x.ctypes.data_as(_D), y.ctypes.data_as(_D), out.ctypes.data_as(_D), len(x))
    if rc != BH_OK:
        raise BhError(f"bh_math_eval({fn}) failed ({rc})")
    return out


def selftest(device=0):
    rows = np.zeros((8, 6))
    err = C.create_string_buffer(512)
    rc = load().bhh_selftest(device, rows.ctypes.data_as(_D), err, 512)
    if rc != BH_OK:
        raise BhError(err.value.decode())
    return rows


# --- device context ---------------------------------------------------------------------------------
def batch_from_ini(ini, sets="", n0=None):
    """(Scene, Frame) list of a [batch] shot: frame i has spin spins[i % len(spins)], size
    sizes[(i // len(spins)) % len(sizes)] squared and time i * frame_dt_m; camera from [camera].
    n0 (the normalised density, e.g. Tracer.scene().n0) replaces every scene's n0 when given."""
    import math
    def floats(key, default):
        v = config_str(ini, key, sets)
        return [float(x) for x in v.split(",")] if v else default
    nf = int(config_str(ini, "batch.frames", sets) or 1)
    spins = floats("batch.spins", [float(config_str(ini, "bh.spin", sets) or 0.94)])
    sizes = [int(x) for x in floats("batch.sizes", [float(config_str(ini, "camera.width", sets))])]
    dt = float(config_str(ini, "batch.frame_dt_m", sets) or 0.0)
    r = float(config_str(ini, "camera.r", sets))
    incl = float(config_str(ini, "camera.incl", sets)) * math.pi / 180
    fov = float(config_str(ini, "camera.fov_m", sets))
    scenes = {}
    out = []
    for i in range(nf):
        a = spins[i % len(spins)]
        if a not in scenes:
            extra = f"bh.spin={a!r}"
            scenes[a] = build_scene(ini, f"{sets};{extra}" if sets else extra)[0]
            if n0 is not None:
                scenes[a].n0 = n0
        S = scenes[a]
        n = sizes[(i // len(spins)) % len(sizes)]
        F = make_frame(S.a, r, incl, fov=fov, W=n, H=n, spp=1, seed=(0x9E3779B9 * (i + 1)) & 0xFFFFFFFF)
        F.t = i * dt
        out.append((S, F))
    return out


class Tracer:
    """A device context (bh_ctx)."""

    def __init__(self, scene, k2=None, k2_l0=0.0, k2_dl=0.0, grid=None, turb_half=60.0, turb_t0=0.0, turb_dt=0.0,
                 device=0, settings=None, turb_model=TURB_GRID, ts_dims=(0, 0, 0, 0)):
        L = load()
        T = Tables()
        self._keep = []
        if k2 is not None:
            k2 = np.ascontiguousarray(k2, dtype=np.float32)
            self._keep.append(k2)
            T.k2_ln = k2.ctypes.data_as(_F)
            T.k2_n = len(k2)
            T.k2_l0 = k2_l0
            T.k2_dl = k2_dl
        if grid is not None:
            g = np.ascontiguousarray(grid, dtype=np.float32)
            if g.ndim == 3:
                g = g[None]
            self._keep.append(g)
            T.turb = g.ctypes.data_as(_F)
            T.turb_n = g.shape[-1]
            T.turb_nt = g.shape[0]
            T.turb_half = turb_half
            T.turb_t0 = turb_t0
            T.turb_dt = turb_dt
        T.turb_model = turb_model
        for i in range(4):
            T.ts_dims[i] = int(ts_dims[i])
        rs = settings or Settings()
        ctx = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = L.bh_create(device, C.byref(scene), C.byref(T), C.byref(rs), C.byref(ctx), err, 512)
        if rc != BH_OK:
            raise BhError(f"bh_create: {err.value.decode()} ({rc})")
        self.ctx = ctx
        self.L = L

    @classmethod
    def from_ini(cls, ini, sets="", device=0, normalize=True, frame_index=0):
        """(Tracer, Scene, Frame, TraceOpts, flux Jy) for a shot .ini."""
        L = load()
        self = cls.__new__(cls)
        self._keep = []
        self.L = L
        ctx = C.c_void_p()
        S, F, O = Scene(), Frame(), TraceOpts()
        jy = C.c_double()
        err = C.create_string_buffer(512)
        rc = L.bhh_create_from_ini(ini.encode(), sets.encode(), device, int(normalize), frame_index, C.byref(ctx),
                                   C.byref(S), C.byref(F), C.byref(O), C.byref(jy), err, 512)
        if rc != BH_OK:
            raise BhError(f"bhh_create_from_ini: {err.value.decode()} ({rc})")
        self.ctx = ctx
        return self, S, F, O, jy.value

    def _check(self, rc, what):
        if rc != BH_OK:
            raise BhError(f"{what}: {self.L.bh_last_error(self.ctx).decode()} ({rc})")

    def close(self):
        if getattr(self, "ctx", None):
            self.L.bh_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update_scene(self, S):
        self._check(self.L.bh_update_scene(self.ctx, C.byref(S)), "bh_update_scene")

    def scene(self):
        S = Scene()
        self._check(self.L.bh_get_scene(self.ctx, C.byref(S)), "bh_get_scene")
        return S

    @staticmethod
    def opts(precision=PREC_FP64, emissivity=EMIS_EXACT, jitter=0, spp=1, tol_far=0.0, r_near=0.0, r_emit=0.0,
             far_fp32=0, full_flow=0, engine=ENGINE_AUTO, tol_t=0.0, relax_after=0, tol_relax=0.0, relax_cap=0.0):
        return TraceOpts(precision, emissivity, jitter, spp, tol_far, r_near, r_emit, far_fp32, full_flow, engine,
                         tol_t, relax_after, tol_relax, relax_cap)

    def render(self, F, opts=None):
        px = np.zeros((F.H, F.W, 4), dtype=np.float32)
        st = Stats()
        o = opts or self.opts()
        self._check(self.L.bh_render(self.ctx, C.byref(F), C.byref(o), px.ctypes.data_as(_F), C.byref(st)),
                    "bh_render")
        return px, st

    def render_adaptive(self, F, opts=None):
        """(image H x W x 4, samples per pixel H x W, Stats) with the context's spp / spp_max / target_err."""
        px = np.zeros((F.H, F.W, 4), dtype=np.float32)
        ns = np.zeros((F.H, F.W), dtype=np.uint32)
        st = Stats()
        o = opts or self.opts()
        self._check(self.L.bh_render_adaptive(self.ctx, C.byref(F), C.byref(o), px.ctypes.data_as(_F),
                                              ns.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(st)),
                    "bh_render_adaptive")
        return px, ns, st

    def render_resident(self, F, opts=None):
        st = Stats()
        o = opts or self.opts()
        ptr = C.c_void_p()
        self._check(self.L.bh_render_resident(self.ctx, C.byref(F), C.byref(o), C.byref(ptr), C.byref(st)),
                    "bh_render_resident")
        return ptr.value, st

    def render_noalloc(self, F, d_px, opts=None):
        """bh_render_noalloc: the frame into the caller's device buffer d_px (an address of W*H*4
        floats); returns the Stats."""
        st = Stats()
        o = opts or self.opts()
        self._check(self.L.bh_render_noalloc(self.ctx, C.byref(F), C.byref(o), C.c_void_p(d_px), C.byref(st)),
                    "bh_render_noalloc")
        return st

    def trace_tile(self, F, x0, y0, w, h, opts=None):
        out = np.zeros(w * h, dtype=RAY_DTYPE)
        o = opts or self.opts()
        self._check(self.L.bh_trace_tile(self.ctx, C.byref(F), x0, y0, w, h, C.byref(o), out.ctypes.data),
                    "bh_trace_tile")
        return out.reshape(h, w)

    def radio_flux(self, F, n, pix_sr):
        jy = C.c_double()
        self._check(self.L.bh_radio_flux(self.ctx, C.byref(F), n, pix_sr, C.byref(jy)), "bh_radio_flux")
        return jy.value

    def vacuum_fates(self, photons, tol, r_escape, max_steps):
        ph = np.ascontiguousarray(photons, dtype=np.float64).reshape(-1, 9)
        fates = np.zeros(len(ph), dtype=np.int32)
        self._check(self.L.bh_vacuum_fates(self.ctx, ph.ctypes.data_as(_D), len(ph), tol, r_escape, max_steps,
                                           fates.ctypes.data_as(C.POINTER(C.c_int))), "bh_vacuum_fates")
        return fates

    def sheared_slices(self, F, dims):
        """The baked sheared-turbulence slices for frame F: (nt, nz, nlr, nphi) float32; dims as
        given to the constructor (nphi, nlr, nz, nt)."""
        nphi, nlr, nz, nt = dims
        out = np.zeros((nt, nz, nlr, nphi), dtype=np.float32)
        self._check(self.L.bh_sheared_slices(self.ctx, C.byref(F), out.ctypes.data_as(_F), out.size),
                    "bh_sheared_slices")
        return out

    # --- radio post step and display ---------------------------------------------------------
    def post_radio(self, img, blur_sigma_px=0.0, hide_direct=0.0, percentile=0.995):
        """Beam blur of an H x W x 4 float32 frame (returns a new array) and its display statistics
        (lum_percentile, lum_max)."""
        out = np.ascontiguousarray(img, dtype=np.float32).copy()
        H, W = out.shape[:2]
        pp = PostParams(blur_sigma_px, hide_direct)
        lp, lm = C.c_double(), C.c_double()
        self._check(self.L.bh_post_radio(self.ctx, out.ctypes.data_as(_F), W, H, C.byref(pp), percentile, C.byref(lp),
                                         C.byref(lm)), "bh_post_radio")
        return out, lp.value, lm.value

    def tonemap(self, img, colormap=0, exposure=1.0, gamma=1.0, overlay=0.0, hide_direct=0.0, fade=1.0, vignette=0.0,
                grain=0.0, seed=0, glyphs=None, rects=None, atlas=None, advance=0.0, baseline=0):
        """Radio tone map of an H x W x 4 float32 frame to H x W x 3 bytes; with glyphs / rects (sequences
        of (x, y, scale, glyph, rgba) / (x0, y0, x1, y1, rgba, layer)) and atlas (glyphs x cell_h x
        cell_w uint8), the HUD is drawn on top (bh_tonemap_hud)."""
        px = np.ascontiguousarray(img, dtype=np.float32)
        H, W = px.shape[:2]
        rgb = np.zeros((H, W, 3), dtype=np.uint8)
        tp = ToneParams(colormap, exposure, gamma, overlay, hide_direct, fade, vignette, grain, seed)
        if glyphs is None and rects is None:
            self._check(self.L.bh_tonemap(self.ctx, px.ctypes.data_as(_F), W, H, C.byref(tp),
                                          rgb.ctypes.data_as(C.POINTER(C.c_uint8))), "bh_tonemap")
            return rgb
        glyphs, rects = list(glyphs or []), list(rects or [])
        ga, ra = hud_arrays(glyphs, rects)
        font = None
        if atlas is not None:
            at = np.ascontiguousarray(atlas, dtype=np.uint8)
            font = Font(at.ctypes.data_as(C.POINTER(C.c_uint8)), at.shape[2], at.shape[1], at.shape[0], advance, baseline)
        self._check(self.L.bh_tonemap_hud(self.ctx, px.ctypes.data_as(_F), W, H, C.byref(tp), ga, len(glyphs), ra,
                                          len(rects), C.byref(font) if font is not None else None,
                                          rgb.ctypes.data_as(C.POINTER(C.c_uint8))), "bh_tonemap_hud")
        return rgb

    # --- persistent batches -------------------------------------------------------------------
    @staticmethod
    def _batch(frames):
        arr = (BatchFrame * len(frames))()
        for i, (S, F) in enumerate(frames):
            arr[i].scene = S
            arr[i].frame = F
        return arr

    def trace_batch(self, frames, opts=None, cfg=None, block=0, grid_per_cu=0, interval=0, refill_min=0,
                    use_graph=1, waves=0, order=0, prio_age=0, resident=False):
        """frames: list of (Scene, Frame). Returns ([H x W x 4 images], Stats), or (None, Stats) with
        resident=True (the images stay in device memory). cfg None and all launch arguments 0 -> the
        last autotuned configuration, else the defaults."""
        arr = self._batch(frames)
        npix = sum(F.W * F.H for _, F in frames)
        px = np.zeros((0 if resident else npix, 4), dtype=np.float32)
        st = Stats()
        o = opts or self.opts()
        if cfg is None and (block or grid_per_cu or interval or refill_min or waves or order or prio_age or
                            use_graph != 1):
            cfg = BatchCfg(block, grid_per_cu, interval, refill_min, use_graph, waves, order, prio_age)
        pc = C.byref(cfg) if cfg is not None else None
        dst = None if resident else px.ctypes.data_as(_F)
        self._check(self.L.bh_trace_batch(self.ctx, arr, len(frames), C.byref(o), pc, dst, C.byref(st)),
                    "bh_trace_batch")
        if resident:
            return None, st
        out, off = [], 0
        for _, F in frames:
            out.append(px[off:off + F.W * F.H].reshape(F.H, F.W, 4))
            off += F.W * F.H
        return out, st

    def autotune(self, frames, opts=None, max_tried=256):
        arr = self._batch(frames)
        best = BatchCfg()
        ms = C.c_double()
        tried = np.zeros((max_tried, 8), dtype=np.float64)
        nt = C.c_int()
        o = opts or self.opts()
        self._check(self.L.bh_autotune(self.ctx, arr, len(frames), C.byref(o), C.byref(best), C.byref(ms),
                                       tried.ctypes.data_as(_D), max_tried, C.byref(nt)), "bh_autotune")
        return best, ms.value, tried[:min(nt.value, max_tried)]

    # --- surrogate ----------------------------------------------------------------------------
    def mlp_train(self, n_samples=0, batch=0, epochs=0, lr=0.0, lr_min=0.0, hot_weight=-1.0, seed=0):
        cfg = MlpTrainCfg(n_samples, batch, epochs, lr, lr_min, hot_weight, seed)
        rep = MlpReport()
        self._check(self.L.bh_mlp_train(self.ctx, C.byref(cfg), C.byref(rep)), "bh_mlp_train")
        return rep.as_dict()

    def mlp_get(self):
        p = np.zeros(MLP_NPARAM, dtype=np.float32)
        nrm = np.zeros(MLP_NNORM, dtype=np.float32)
        self._check(self.L.bh_mlp_get(self.ctx, p.ctypes.data_as(_F), nrm.ctypes.data_as(_F)), "bh_mlp_get")
        return p, nrm

    def mlp_set(self, params, norm):
        p = np.ascontiguousarray(params, dtype=np.float32)
        nrm = np.ascontiguousarray(norm, dtype=np.float32)
        self._check(self.L.bh_mlp_set(self.ctx, p.ctypes.data_as(_F), nrm.ctypes.data_as(_F)), "bh_mlp_set")

    def mlp_infer(self, u0, theta, theta_b):
        x = np.ascontiguousarray(np.column_stack([u0, theta, theta_b]), dtype=np.float32)
        out = np.zeros((len(x), 2), dtype=np.float32)
        self._check(self.L.bh_mlp_infer(self.ctx, x.ctypes.data_as(_F), len(x), out.ctypes.data_as(_F)),
                    "bh_mlp_infer")
        return out

    def mlp_exact(self, u0, theta, theta_b):
        x = np.ascontiguousarray(np.column_stack([u0, theta, theta_b]), dtype=np.float32)
        out = np.zeros((len(x), 2), dtype=np.float64)
        self._check(self.L.bh_mlp_exact(self.ctx, x.ctypes.data_as(_F), len(x), out.ctypes.data_as(_D)),
                    "bh_mlp_exact")
        return out

    def mlp_grad(self, params, norm, xn, tn, w):
        p = np.ascontiguousarray(params, dtype=np.float32)
        nrm = np.ascontiguousarray(norm, dtype=np.float32)
        xn = np.ascontiguousarray(xn, dtype=np.float32)
        tn = np.ascontiguousarray(tn, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        g = np.zeros(MLP_NPARAM, dtype=np.float32)
        loss = C.c_float()
        self._check(self.L.bh_mlp_grad(self.ctx, p.ctypes.data_as(_F), nrm.ctypes.data_as(_F), xn.ctypes.data_as(_F),
                                       tn.ctypes.data_as(_F), w.ctypes.data_as(_F), len(xn), C.byref(loss),
                                       g.ctypes.data_as(_F)), "bh_mlp_grad")
        return loss.value, g

    def mlp_bake(self, n0=0, n1=0, n2=0, judge_tol=0.0):
        rep = MlpBakeReport()
        self._check(self.L.bh_mlp_bake(self.ctx, n0, n1, n2, judge_tol, C.byref(rep)), "bh_mlp_bake")
        return rep.as_dict()

    def find_critical_point(self, F, psi):
        rho = C.c_double()
        self._check(self.L.bhh_find_critical_point(self.ctx, C.byref(F), psi, C.byref(rho)),
                    "bhh_find_critical_point")
        return rho.value
