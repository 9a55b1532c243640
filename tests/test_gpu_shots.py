"""The photon-ring zoom (2048^2 window 0.0005 M wide on the critical curve, tolerance 1e-12) and the
4K frame with the emissivity surrogate, through the product's shot loader (bhh_create_from_ini).

Bars: zoom tile bit-exact against the oracle (FP64); the persistent kernel bit-exact against the
frame kernel on the zoom; mixed precision keeps fates and winding counts on the zoom (FP64 near the
photon sphere); the 4K frame with the trained surrogate within 0.1% flux and 1% median per-pixel
error of the FP64 frame."""
import ctypes as C

import numpy as np
import pytest

import bhrt_amd as bh
from helpers import pixel_grid, scene_dict, shot
from oracle_ctypes import OrFrame, make_scene
from test_gpu_parity import assert_rays_equal

pytestmark = pytest.mark.gpu


def oracle_frame(F):
    O = OrFrame()
    C.memmove(C.addressof(O), C.addressof(F), C.sizeof(OrFrame))
    return O


@pytest.fixture(scope="module")
def zoom():
    t, S, F, O, _ = bh.Tracer.from_ini(shot("c3_zoom_2048.ini"), normalize=False)
    yield t, S, F
    t.close()


def test_zoom_tile_bit_exact(zoom, oracle_bhm):
    t, S, F = zoom
    x0, y0 = 1024 - 24, 1024 - 24
    gpu = t.trace_tile(F, x0, y0, 48, 48)
    px, py = pixel_grid(48, 48, x0, y0)
    _, ref, _ = oracle_bhm.trace_pixels(make_scene(scene_dict(S)), oracle_frame(F), px, py, spp=1, jitter=False)
    assert_rays_equal(gpu, ref, "zoom tile")
    steps, n = gpu["steps"].ravel(), gpu["ncross"].ravel()
    print(f"zoom tile: steps {steps.min()}..{steps.max()} (mean {steps.mean():.0f}), windings {sorted(set(n))}")
    assert S.tol == 1e-12


def test_zoom_frame_persistent_and_mixed(zoom):
    """A 512^2 sampling of the zoom window (same 0.0005 M field): every ray orbits (4 to 9+ equatorial
    crossings) and the step counts spread about 2x; the persistent kernel equals the frame kernel
    bit for bit; the mixed tracer keeps the fates and most winding counts."""
    t, S, F = zoom
    G = bh.Frame.from_buffer_copy(F)
    G.pix = F.pix * 4
    G.W = G.H = 512   # same field: pixel offsets scale with pix, the window offset (cx, cy) stays
    img, st = t.render(G)
    rays = t.trace_tile(G, 0, 0, 512, 512).ravel()
    steps = rays["steps"]
    print(f"zoom 512^2: steps {steps.min()}..{steps.max()}, windings {np.bincount(rays['ncross'])}")
    assert steps.max() > 1.5 * steps.min() and len(np.unique(rays['ncross'])) >= 4
    got, stb = t.trace_batch([(S, G)], t.opts(precision=bh.PREC_FP64))
    assert np.array_equal(got[0].view(np.int32), img.view(np.int32))
    assert stb.step_attempts == st.step_attempts
    mix = t.trace_tile(G, 0, 0, 512, 512, t.opts(precision=bh.PREC_MIXED)).ravel()
    same_fate = (mix["fate"] == rays["fate"]).mean()
    same_n = (mix["ncross"] == rays["ncross"]).mean()
    print(f"zoom mixed: fate agreement {same_fate:.5f}, winding agreement {same_n:.5f}, "
          f"steps {mix['steps'].mean():.0f} vs {steps.mean():.0f}")
    assert same_fate > 0.99 and same_n > 0.9


def test_4k_frame_with_surrogate():
    sets = "camera.width=1920;camera.height=1080"
    t, S, F, O, jy = bh.Tracer.from_ini(shot("c4_mlp_4k.ini"), sets, normalize=True)
    assert O.precision == bh.PREC_MIXED and O.emissivity == bh.EMIS_MLP
    t.mlp_train()
    bake = t.mlp_bake()
    ref, s64 = t.render(F, t.opts(precision=bh.PREC_FP64))
    img, smx = t.render(F, O)
    Ia, Ib = img[..., 0].astype(np.float64), ref[..., 0].astype(np.float64)
    mask = Ib > 1e-3 * Ib.max()
    flux = abs(Ia.sum() - Ib.sum()) / Ib.sum()
    med = np.median(np.abs(Ia[mask] - Ib[mask]) / Ib[mask])
    print(f"1920x1080 c4: flux err {flux:.2e}, pixel median {med:.2e}; kernel {smx.ms_kernel:.1f} vs "
          f"{s64.ms_kernel:.1f} ms; bake {bake}")
    assert flux < 1e-3 and med < 1e-2
    t.close()
