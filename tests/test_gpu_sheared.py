"""The reference's flow turbulence (baked sheared noise) on the GPU against the oracle's restatement:
the bake bit for bit (small slices in full, the reference's 512 x 192 x 48 x 24 slices on a
sample of cells), the FP64 tracer reading the slices bit for bit (per-ray outputs on a block across
the photon ring), the mixed tracer within the accuracy bar, re-baking when the camera time moves,
and the persistent batch path bit-identical to the frame path."""
import numpy as np
import pytest

import bhrt_amd as bh
from helpers import pixel_grid, setup, shot
from test_gpu_parity import assert_rays_equal, oracle_rays

pytestmark = pytest.mark.gpu

SMALL = (96, 32, 12, 6)


def sheared_setup(oracle, dims, W=256, H=256, bake=True):
    cfg = setup(shot("c2_radio_1024.ini"), W=W, H=H, with_grid=False)
    cfg["S"].n0 = 291464.1043528089
    cfg["OS"].n0 = cfg["S"].n0
    t = bh.Tracer(cfg["S"], k2=cfg["k2"], k2_l0=cfg["l0"], k2_dl=cfg["dl"], turb_model=bh.TURB_SHEARED, ts_dims=dims)
    F = cfg["F"]
    if not bake:   # GPU-only comparisons: no oracle slices
        return cfg, t, None, None
    B = oracle.sheared_params(cfg["OS"], F.t, F.r, dims)
    slices = oracle.sheared_bake(B)
    oracle.attach_sheared(cfg["OS"], B, slices)
    return cfg, t, B, slices


def test_bake_matches_oracle(oracle_bhm):
    cfg, t, B, slices = sheared_setup(oracle_bhm, SMALL)
    got = t.sheared_slices(cfg["F"], SMALL)
    assert got.shape == slices.shape
    bad = np.argwhere(got.view(np.int32) != slices.view(np.int32))
    assert len(bad) == 0, f"{len(bad)} cells differ, first {bad[:3].tolist()}"
    assert 0.05 < slices.std() < 3.0   # flow.turbulence = 0.3: a log-normal field of spread ~0.17
    t.close()


def test_bake_full_size_sample(oracle_bhm):
    dims = (512, 192, 48, 24)   # the reference's slices (render.cu:448)
    cfg = setup(shot("c2_radio_1024.ini"), W=64, H=64, with_grid=False)
    t = bh.Tracer(cfg["S"], k2=cfg["k2"], k2_l0=cfg["l0"], k2_dl=cfg["dl"], turb_model=bh.TURB_SHEARED)
    got = t.sheared_slices(cfg["F"], dims)
    B = oracle_bhm.sheared_params(cfg["OS"], cfg["F"].t, cfg["F"].r, dims)
    rng = np.random.default_rng(11)
    idx = [tuple(int(rng.integers(0, n)) for n in dims) for _ in range(3000)]
    for (i, j, l, k) in idx:
        want = np.float32(oracle_bhm.sheared_cell(B, i, j, l, k))
        assert got[k, l, j, i].view(np.int32) == want.view(np.int32), (i, j, l, k, got[k, l, j, i], want)
    t.close()


def test_fp64_tracer_reads_slices_bit_exact(oracle_bhm):
    cfg, t, B, slices = sheared_setup(oracle_bhm, SMALL)
    x0, y0 = 128 - 24 + 24, 128 - 24 - 10
    gpu = t.trace_tile(cfg["F"], x0, y0, 48, 48)
    px, py = pixel_grid(48, 48, x0, y0)
    ref = oracle_rays(oracle_bhm, cfg, px, py)
    assert_rays_equal(gpu, ref, "sheared block")
    assert (gpu["I"][..., 0] > 0).mean() > 0.5
    # the turbulence matters: the same frame without it differs
    S0 = cfg["S"]
    S0.flow_turb_amp = 0.0
    t.update_scene(S0)
    flat = t.trace_tile(cfg["F"], x0, y0, 48, 48)
    assert np.abs(flat["I"][..., 0] - gpu["I"][..., 0]).max() > 1e-3 * gpu["I"][..., 0].max()
    t.close()


@pytest.mark.parametrize("dims", [SMALL, (512, 192, 48, 24)])
def test_mixed_tracer_with_sheared_turbulence(oracle_bhm, dims):
    """Within the accuracy bar at the small slices and at the reference's 512 x 192 x 48 x 24 (there the
    1/256-quantised filter weights make the density step at 1/256 of a texel, so the mixed tracer's
    slightly different sample points give ~4e-3 median differences)."""
    cfg, t, _, _ = sheared_setup(oracle_bhm, dims, bake=False)
    ref, _ = t.render(cfg["F"], t.opts(precision=bh.PREC_FP64))
    mix, _ = t.render(cfg["F"], t.opts(precision=bh.PREC_MIXED))
    Ia, Ib = mix[..., 0].astype(np.float64), ref[..., 0].astype(np.float64)
    flux = abs(Ia.sum() - Ib.sum()) / Ib.sum()
    m = Ib > 1e-3 * Ib.max()
    med = np.median(np.abs(Ia[m] - Ib[m]) / Ib[m])
    print(f"sheared turbulence {dims}, mixed vs FP64: flux {flux:.2e}, pixel median {med:.2e}")
    assert flux < 1e-3 and med < 1e-2
    t.close()


def test_rebake_on_camera_time_and_batch(oracle_bhm):
    cfg, t, _, slices = sheared_setup(oracle_bhm, SMALL, W=64, H=64)
    F2 = cfg["F"].__class__.from_buffer_copy(cfg["F"])
    F2.t = cfg["F"].t + 40.0
    later = t.sheared_slices(F2, SMALL)
    B2 = oracle_bhm.sheared_params(cfg["OS"], F2.t, F2.r, SMALL)
    assert np.array_equal(later, oracle_bhm.sheared_bake(B2))
    assert not np.array_equal(later, slices)
    # batches: one bake per launch; the persistent path equals the frame path
    o = t.opts(precision=bh.PREC_FP64)
    want, _ = t.render(cfg["F"], o)
    got, _ = t.trace_batch([(cfg["S"], cfg["F"])] * 2, o)
    for g in got:
        assert np.array_equal(g, want)
    with pytest.raises(bh.BhError, match="share the camera time"):
        t.trace_batch([(cfg["S"], cfg["F"]), (cfg["S"], F2)], o)
    t.close()
