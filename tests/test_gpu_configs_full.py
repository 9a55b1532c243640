"""The surrogate frame, the batched animation and the photon-ring zoom at their stated sizes.

- 4K frame (shots/c4_mlp_4k.ini): 3840 x 2160 rays of the 230 GHz flow, mixed precision + trained
  surrogate, against the FP64 frame: flux within 0.1 % and per-pixel median within 1 %; one FP64
  tile of the same frame bit for bit against the oracle.
- Batched animation (shots/c5_batch_anim.ini): 64 frames of 512^2 to 2048^2 pixels (125.8 M rays),
  spins 0 / 0.5 / 0.94 / 0.998, 32 time slices of a 64^3 turbulence grid read at each sample's
  emission time, mixed precision + surrogate, in one persistent launch: bit-identical to rendering the
  frames one by one; the 512^2 frames within the accuracy bars of their FP64 frames; two FP64 tiles
  (different spins and sizes) bit for bit against the oracle.
- Photon-ring zoom (shots/c3_zoom_2048.ini): 2048^2 rays at tolerance 1e-12, FP64 on the persistent
  engine bit-identical to the frame kernel, the mixed tracer capturing the same rays."""
import numpy as np
import pytest

import bhrt_amd as bh
from helpers import pixel_grid, scene_dict, setup, shot, tracer_for, oracle_frame
from oracle_ctypes import make_scene
from test_gpu_batch import assert_same, render_each, stats_vec
from test_gpu_parity import assert_rays_equal

pytestmark = pytest.mark.gpu

FLUX_BAR, MEDIAN_BAR = 1e-3, 1e-2


def frame_errors(img, ref):
    Ia, Ib = img[..., 0].astype(np.float64), ref[..., 0].astype(np.float64)
    mask = Ib > 1e-3 * Ib.max()
    return abs(Ia.sum() - Ib.sum()) / Ib.sum(), float(np.median(np.abs(Ia[mask] - Ib[mask]) / Ib[mask]))


def test_c4_full_size_with_surrogate(oracle_bhm):
    cfg = setup(shot("c4_mlp_4k.ini"))
    F = cfg["F"]
    assert (F.W, F.H) == (3840, 2160)
    t = tracer_for(cfg)
    try:
        t.mlp_train()
        bake = t.mlp_bake()
        print(f"c4 3840x2160: bake {bake}", flush=True)
        ref, s64 = t.render(F, t.opts(precision=bh.PREC_FP64))
        img, smx = t.render(F, t.opts(precision=bh.PREC_MIXED, emissivity=bh.EMIS_MLP))
        flux, med = frame_errors(img, ref)
        print(f"c4 3840x2160: flux err {flux:.2e}, pixel median {med:.2e}; kernel {smx.ms_kernel:.1f} ms "
              f"(FP64 {s64.ms_kernel:.1f} ms), {smx.samples / smx.ms_kernel * 1e-3:.0f} Mrays/s", flush=True)
        assert smx.samples == s64.samples == 3840 * 2160
        assert flux < FLUX_BAR and med < MEDIAN_BAR
        # an FP64 tile across the photon ring of the same frame, bit for bit against the oracle
        x0, y0 = 3840 // 2 + 150, 2160 // 2 - 80
        gpu = t.trace_tile(F, x0, y0, 24, 24)
        px, py = pixel_grid(24, 24, x0, y0)
        _, want, _ = oracle_bhm.trace_pixels(cfg["OS"], cfg["OF"], px, py, spp=1, jitter=False)
        assert_rays_equal(gpu, want, "c4 tile")
        assert (gpu["I"][..., 0] > 0).mean() > 0.5
    finally:
        t.close()


def c5_grid(ini, S):
    """The time-sliced grid bhh_create_from_ini builds for the shot (same keys and defaults)."""
    num = lambda k, d: float(bh.config_str(ini, k) or d)
    n, nt = int(num("turbulence.grid", 128)), int(num("turbulence.slices", 1))
    half = num("turbulence.half_m", S.r_flow_max)
    t0, dt = num("turbulence.t0_m", 0.0), num("turbulence.dt_m", 10.0)
    g = bh.turbulence(n, nt, half, S.flow_turb_amp, num("turbulence.corr_m", 5.0), num("turbulence.corr_t_m", 30.0),
                      dt, int(num("turbulence.seed", 1)))
    return g, half, t0, dt


def test_c5_full_size_one_launch(oracle_bhm):
    ini = shot("c5_batch_anim.ini")
    t, S0, F0, O, _ = bh.Tracer.from_ini(ini, normalize=False)
    try:
        assert O.precision == bh.PREC_MIXED and O.emissivity == bh.EMIS_MLP
        frames = bh.batch_from_ini(ini)
        sizes = sorted({F.W for _, F in frames})
        assert len(frames) == 64 and sizes == [512, 1024, 1536, 2048]
        assert sorted({round(S.a, 6) for S, _ in frames}) == [1e-06, 0.5, 0.94, 0.998]   # a = 0 is clamped to 1e-6
        assert sum(F.W * F.H for _, F in frames) == 16 * (512 ** 2 + 1024 ** 2 + 1536 ** 2 + 2048 ** 2)
        t.mlp_train()
        t.mlp_bake()
        o = t.opts(precision=bh.PREC_MIXED, emissivity=bh.EMIS_MLP)
        got, st = t.trace_batch(frames, o)
        print(f"c5 one launch: {st.samples / 1e6:.1f} M rays in {st.ms_kernel:.1f} ms "
              f"({st.samples / st.ms_kernel * 1e-3:.0f} Mrays/s)", flush=True)
        assert st.samples == sum(F.W * F.H for _, F in frames)
        want, flux, cnt = render_each(t, frames, o)
        assert_same(got, want, "c5 one launch vs frame by frame")
        assert np.array_equal(stats_vec(st), cnt)
        assert abs(st.flux_sum - flux) <= 1e-12 * abs(flux)
        del want
        # accuracy of the mixed + surrogate frames against FP64, on the 512^2 frames (all four spins)
        small = [i for i, (_, F) in enumerate(frames) if F.W == 512]
        ref, _ = t.trace_batch([frames[i] for i in small], t.opts(precision=bh.PREC_FP64))
        worst = (0.0, 0.0)
        for k, i in enumerate(small):
            flux, med = frame_errors(got[i], ref[k])
            worst = (max(worst[0], flux), max(worst[1], med))
            assert flux < FLUX_BAR and med < MEDIAN_BAR, (i, frames[i][0].a, flux, med)
        print(f"c5 512^2 frames: worst flux err {worst[0]:.2e}, worst pixel median {worst[1]:.2e}", flush=True)
        # two FP64 tiles against the oracle: different spins, sizes and emission times
        g, half, t0, dt = c5_grid(ini, S0)
        for i in (2, 63):
            S, F = frames[i]
            t.update_scene(S)
            OS = make_scene(scene_dict(S), k2=bh.k2_from_scene(ini)[0], grid=g, turb_half=half, turb_t0=t0,
                            turb_dt=dt)
            _, _, (l0, dl), _, _ = bh.build_scene(ini)
            OS.k2_l0, OS.k2_dl = l0, dl
            x0, y0 = F.W // 2 + F.W // 10, F.H // 2 - 8
            gpu = t.trace_tile(F, x0, y0, 16, 16)
            px, py = pixel_grid(16, 16, x0, y0)
            _, rays, _ = oracle_bhm.trace_pixels(OS, oracle_frame(F), px, py, spp=1, jitter=False)
            assert_rays_equal(gpu, rays, f"c5 frame {i} tile")
            assert (gpu["I"][..., 0] > 0).any()
    finally:
        t.close()


def test_c3_full_size_zoom():
    """The photon-ring zoom at its stated size: 2048^2 rays in the 0.0005 M window on the critical
    curve, tolerance 1e-12 (every ray orbits; step counts spread over several times the mean). The
    FP64 frame on the persistent engine equals the frame kernel's bit for bit, counters included, and
    no ray reaches the step limit; the mixed tracer captures the same rays on all but 0.1 % of the
    pixels (FP64 near the photon sphere)."""
    t, S, F, O, _ = bh.Tracer.from_ini(shot("c3_zoom_2048.ini"), normalize=False)
    try:
        assert (F.W, F.H) == (2048, 2048) and S.tol == 1e-12
        img, st = t.render(F, t.opts(precision=bh.PREC_FP64))
        n = F.W * F.H
        print(f"c3 2048^2 FP64, persistent: {st.ms_kernel:.1f} ms ({n / st.ms_kernel * 1e-3:.1f} Mrays/s), "
              f"{st.step_attempts / n:.0f} attempts per ray", flush=True)
        ref, sf = t.render(F, t.opts(precision=bh.PREC_FP64, engine=bh.ENGINE_FRAME))
        print(f"c3 2048^2 FP64, frame kernel: {sf.ms_kernel:.1f} ms", flush=True)
        assert np.array_equal(img.view(np.int32), ref.view(np.int32))
        assert (st.step_attempts, st.steps_accepted, st.step_limited) == \
               (sf.step_attempts, sf.steps_accepted, sf.step_limited)
        assert st.step_limited == 0
        mix, sm = t.render(F, t.opts(precision=bh.PREC_MIXED))
        cap64, capm = np.all(img[..., :3] == 0, axis=-1), np.all(mix[..., :3] == 0, axis=-1)
        agree = float((cap64 == capm).mean())
        print(f"c3 2048^2 mixed: {sm.ms_kernel:.1f} ms ({n / sm.ms_kernel * 1e-3:.1f} Mrays/s), "
              f"{sm.step_attempts / n:.0f} attempts per ray; captured-pixel agreement {agree:.5f} "
              f"(captured {cap64.mean():.3f})", flush=True)
        assert agree > 0.999 and 0.01 < cap64.mean() < 0.99
    finally:
        t.close()
