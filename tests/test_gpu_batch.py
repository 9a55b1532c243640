"""Persistent batched tracer (bh_trace_batch, bh_autotune) against the one-frame-per-launch path.

The persistent kernel runs the same per-ray integrator (init / step / result) as the frame kernel
(bh_render with engine BH_ENGINE_FRAME), only scheduled differently (work queues, lane retirement and refill, graph replay), so every pixel must
be bit-identical to the frame kernel's for both precisions, for every launch configuration, and for
batches that mix frame sizes, spins and turbulence time slices. Frame totals (flux, step counters)
must agree: integer counters exactly, the flux sum to 1e-12 relative (summation order).
"""
import numpy as np
import pytest

import bhrt_amd as bh
from helpers import D2R, setup, shot, tracer_for

pytestmark = pytest.mark.gpu


def frame_for(S, ini, W, H, t=0.0, sets=""):
    r = float(bh.config_str(ini, "camera.r", sets))
    incl = float(bh.config_str(ini, "camera.incl", sets)) * D2R
    fov = float(bh.config_str(ini, "camera.fov_m", sets))
    F = bh.make_frame(S.a, r, incl, fov=fov, W=W, H=H, spp=1, seed=0x9E3779B9)
    F.t = t
    return F


def render_each(t, frames, opts):
    """bh_render of each (Scene, Frame) in turn on the same context, on the frame kernel."""
    imgs, flux, cnt = [], 0.0, np.zeros(5, dtype=np.int64)
    o = bh.TraceOpts.from_buffer_copy(opts)
    o.engine = bh.ENGINE_FRAME
    for S, F in frames:
        t.update_scene(S)
        img, st = t.render(F, o)
        imgs.append(img)
        flux += st.flux_sum
        cnt += [st.step_attempts, st.steps_accepted, st.flow_steps, st.step_limited, st.fp64_steps]
    return imgs, flux, cnt


def assert_same(imgs_a, imgs_b, what):
    assert len(imgs_a) == len(imgs_b)
    for i, (a, b) in enumerate(zip(imgs_a, imgs_b)):
        assert a.shape == b.shape, (what, i)
        bad = np.argwhere((a.view(np.int32) != b.view(np.int32)).any(axis=-1))
        assert len(bad) == 0, f"{what}: frame {i} differs at {len(bad)} pixels, first {bad[:3].tolist()}"


def stats_vec(st):
    return np.array([st.step_attempts, st.steps_accepted, st.flow_steps, st.step_limited, st.fp64_steps])


@pytest.fixture(scope="module")
def c2():
    cfg = setup(shot("c2_radio_1024.ini"), W=256, H=256)
    t = tracer_for(cfg)
    yield cfg, t
    t.close()


@pytest.mark.parametrize("prec", [bh.PREC_FP64, bh.PREC_MIXED])
def test_batch_single_frame_every_config(c2, prec):
    cfg, t = c2
    o = t.opts(precision=prec)
    want, flux, cnt = render_each(t, [(cfg["S"], cfg["F"])], o)
    for block, gpc, interval, refill, graph, order, age in [
            (256, 2, 8, 16, 1, bh.ORDER_ROWS, -1), (256, 2, 8, 16, 0, bh.ORDER_IMPACT, -1),
            (64, 1, 1, 1, 0, bh.ORDER_ROWS, 64), (128, 4, 32, 48, 1, bh.ORDER_IMPACT, 128),
            (256, 8, 3, 64, 1, bh.ORDER_IMPACT, 1), (256, 4, 32, 32, 1, 0, 0),
            (256, 4, 32, 32, 1, bh.ORDER_CRITICAL, 256), (128, 3, 16, 32, 0, bh.ORDER_CRITICAL, -1),
            (64, 32, 16, 32, 1, bh.ORDER_CRITICAL, 128)]:   # 32 per CU: more than a CU holds, capped
        got, st = t.trace_batch([(cfg["S"], cfg["F"])], o, block=block, grid_per_cu=gpc, interval=interval,
                                refill_min=refill, use_graph=graph, order=order, prio_age=age)
        assert_same(got, want, f"prec {prec} cfg {(block, gpc, interval, refill, graph, order, age)}")
        assert np.array_equal(stats_vec(st), cnt)
        assert st.samples == 256 * 256
        assert abs(st.flux_sum - flux) <= 1e-12 * abs(flux)


@pytest.mark.parametrize("prec", [bh.PREC_FP64, bh.PREC_MIXED])
def test_batch_mixed_sizes_and_spins(c2, prec):
    """One launch over frames of different size (ragged, not multiples of the 8 x 8 queue tile) and spin."""
    cfg, t = c2
    ini = shot("c2_radio_1024.ini")
    frames = []
    for spin, W, H in [(0.94, 256, 256), (0.5, 130, 97), (0.1, 64, 200), (0.99, 8, 8), (0.94, 1, 3)]:
        S, *_ = bh.build_scene(ini, f"bh.spin={spin}")
        frames.append((S, frame_for(S, ini, W, H)))
    o = t.opts(precision=prec)
    want, flux, cnt = render_each(t, frames, o)
    t.update_scene(cfg["S"])
    for order, age in [(0, 0), (bh.ORDER_ROWS, -1), (bh.ORDER_IMPACT, 32), (bh.ORDER_CRITICAL, 256)]:
        got, st = t.trace_batch(frames, o, order=order, prio_age=age)
        assert_same(got, want, f"batch prec {prec} order {order} prio_age {age}")
        assert np.array_equal(stats_vec(st), cnt)
        assert abs(st.flux_sum - flux) <= 1e-12 * abs(flux)


def test_batch_turbulence_time_slices():
    """Frames at different times read different turbulence slices (linear in time between slices)."""
    cfg = setup(shot("c2_radio_1024.ini"), W=96, H=96)
    grid = bh.turbulence(64, 4, 60.0, 0.3, 5.0, 30.0, 10.0, 3)
    t = bh.Tracer(cfg["S"], k2=cfg["k2"], k2_l0=cfg["l0"], k2_dl=cfg["dl"], grid=grid, turb_half=60.0,
                  turb_t0=-1060.0, turb_dt=10.0)   # the flow is crossed ~940-1070 M before the frame time
    ini = shot("c2_radio_1024.ini")
    frames = [(cfg["S"], frame_for(cfg["S"], ini, 96, 96, t=tt)) for tt in (0.0, 4.0, 15.0, 29.0)]
    o = t.opts(precision=bh.PREC_FP64)
    want, _, _ = render_each(t, frames, o)
    got, _ = t.trace_batch(frames, o)
    assert_same(got, want, "time slices")
    assert not np.array_equal(want[0], want[2])   # the slices differ
    t.close()


@pytest.mark.parametrize("prec", [bh.PREC_FP64, bh.PREC_MIXED])
def test_batch_vacuum(prec):
    cfg = setup(shot("c1_vacuum_256.ini"))
    t = tracer_for(cfg)
    o = t.opts(precision=prec)
    want, _, cnt = render_each(t, [(cfg["S"], cfg["F"])], o)
    got, st = t.trace_batch([(cfg["S"], cfg["F"])] * 3, o)
    assert_same(got, want * 3, f"vacuum prec {prec}")
    assert np.array_equal(stats_vec(st), 3 * cnt)
    t.close()


def test_graph_replay_with_eager_work_between(c2):
    cfg, t = c2
    o = t.opts(precision=bh.PREC_MIXED)
    frames = [(cfg["S"], cfg["F"])] * 2
    first, st0 = t.trace_batch(frames, o)
    for _ in range(4):
        t.render(cfg["F"], t.opts(engine=bh.ENGINE_FRAME))   # eager launches between replays
        again, st = t.trace_batch(frames, o)
        assert_same(again, first, "replay")
        assert np.array_equal(stats_vec(st), stats_vec(st0))
        # which rays a workgroup retires is decided at run time, so the per-workgroup flux partials
        # (and the order of the frame sum) vary between launches: equal to rounding
        assert abs(st.flux_sum - st0.flux_sum) <= 1e-12 * abs(st0.flux_sum)


def test_autotune(c2):
    cfg = setup(shot("c2_radio_1024.ini"), W=128, H=128)
    t = tracer_for(cfg)
    frames = [(cfg["S"], cfg["F"])]
    o = t.opts(precision=bh.PREC_MIXED)
    want, _ = t.trace_batch(frames, o, block=256, grid_per_cu=2, interval=8, refill_min=16)
    best, ms, tried = t.autotune(frames, o)
    assert len(tried) >= 10 and ms > 0
    assert ms <= tried[:, 4].min() + 1e-9
    assert best.block in (64, 128, 256) and 1 <= best.interval <= 64
    got, _ = t.trace_batch(frames, o)   # no configuration given: the tuned one
    assert_same(got, want, "tuned")
    t.close()


def test_batch_errors(c2):
    cfg, t = c2
    v = setup(shot("c1_vacuum_256.ini"))
    with pytest.raises(bh.BhError, match="share the render mode"):
        t.trace_batch([(cfg["S"], cfg["F"]), (v["S"], v["F"])])
    with pytest.raises(bh.BhError, match="block must be"):
        t.trace_batch([(cfg["S"], cfg["F"])], block=96)
    with pytest.raises(bh.BhError, match="empty batch"):
        t.trace_batch([])
    with pytest.raises(bh.BhError, match="MLP emissivity"):
        t.trace_batch([(cfg["S"], cfg["F"])], t.opts(precision=bh.PREC_FP64, emissivity=bh.EMIS_MLP))
    with pytest.raises(bh.BhError, match="order must be"):
        t.trace_batch([(cfg["S"], cfg["F"])], order=4)
    with pytest.raises(bh.BhError, match="prio_age must be"):
        t.trace_batch([(cfg["S"], cfg["F"])], prio_age=(1 << 24) + 1)


@pytest.mark.parametrize("prec", [bh.PREC_FP64, bh.PREC_MIXED])
def test_render_engines_agree(c2, prec):
    """bh_render runs on the persistent engine by default (one sample per pixel: a batch of one
    frame; more: the adaptive render's first pass alone): the same pixels and counters as on the
    frame kernel, centre and jittered."""
    cfg, t = c2
    t.update_scene(cfg["S"])
    for jitter in (0, 1):
        want, fst = t.render(cfg["F"], t.opts(precision=prec, jitter=jitter, engine=bh.ENGINE_FRAME))
        got, ast = t.render(cfg["F"], t.opts(precision=prec, jitter=jitter))
        assert_same([got], [want], f"engines prec {prec} jitter {jitter}")
        assert np.array_equal(stats_vec(ast), stats_vec(fst)) and ast.samples == fst.samples == 256 * 256
        assert abs(ast.flux_sum - fst.flux_sum) <= 1e-12 * abs(fst.flux_sum)
    want, fst = t.render(cfg["F"], t.opts(precision=prec, jitter=1, spp=2, engine=bh.ENGINE_FRAME))
    got, ast = t.render(cfg["F"], t.opts(precision=prec, jitter=1, spp=2))
    assert_same([got], [want], f"spp 2 prec {prec}")
    assert np.array_equal(stats_vec(ast), stats_vec(fst)) and ast.samples == fst.samples == 2 * 256 * 256
    assert abs(ast.flux_sum - fst.flux_sum) <= 1e-12 * abs(fst.flux_sum)
    with pytest.raises(bh.BhError, match="engine must be"):
        t.render(cfg["F"], t.opts(engine=2))


@pytest.mark.parametrize("prec", [bh.PREC_FP64, bh.PREC_MIXED])
def test_batched_animation_shot_reduced(prec):
    """The batched-animation shot (time-sliced turbulence, four spins, mixed sizes) at reduced size:
    one persistent launch equals frame-by-frame rendering."""
    ini = shot("c5_batch_anim.ini")
    sets = "batch.frames=8;batch.sizes=40,64;turbulence.grid=32;turbulence.slices=8;turbulence.dt_m=40"
    t, S, F, _, _ = bh.Tracer.from_ini(ini, sets, normalize=False)
    frames = bh.batch_from_ini(ini, sets)
    assert len({round(s.a, 6) for s, _ in frames}) == 4 and len({f.W for _, f in frames}) == 2
    o = t.opts(precision=prec)
    want, flux, cnt = render_each(t, frames, o)
    got, st = t.trace_batch(frames, o)
    assert_same(got, want, "c5 reduced")
    assert np.array_equal(stats_vec(st), cnt)
    assert abs(st.flux_sum - flux) <= 1e-12 * abs(flux)
    t.close()


def test_render_many_samples_beyond_the_sample_budget():
    """bh_render keeps every sample of a frame with spp > 1 on the persistent engine only while they
    fit 1 GiB (32 bytes each); 2048^2 pixels x 9 samples (1.2 GB) run on the frame kernel instead,
    with the same pixels as engine BH_ENGINE_FRAME."""
    cfg = setup(shot("c2_radio_1024.ini"), W=2048, H=2048)
    t = tracer_for(cfg)
    o = t.opts(precision=bh.PREC_MIXED, jitter=1, spp=9)
    assert 2048 * 2048 * 9 * 32 > 1 << 30
    got, st = t.render(cfg["F"], o)
    want, sf = t.render(cfg["F"], t.opts(precision=bh.PREC_MIXED, jitter=1, spp=9, engine=bh.ENGINE_FRAME))
    assert_same([got], [want], "spp 9 at 2048^2")
    assert np.array_equal(stats_vec(st), stats_vec(sf)) and st.samples == 9 * 2048 * 2048
    t.close()


def assert_records_same(a, b, what):
    assert a.shape == b.shape, what
    for f in a.dtype.names:
        x, y = np.ascontiguousarray(a[f]), np.ascontiguousarray(b[f])
        neq = (x.view(np.uint8) != y.view(np.uint8)).reshape(x.shape[0], x.shape[1], -1).any(axis=-1)
        bad = np.argwhere(neq)
        assert len(bad) == 0, f"{what}: field {f} differs at {len(bad)} rays, first {bad[:3].tolist()}"


@pytest.mark.parametrize("prec", [bh.PREC_FP64, bh.PREC_MIXED])
def test_trace_tile_engines_agree(c2, prec):
    """bh_trace_tile runs on the persistent engine by default (the region's 8 x 8 tiles as queue
    entries, in the launch's order: rows for FP64, critical for mixed); the frame kernel (engine
    BH_ENGINE_FRAME) must give the same records bit for bit: the whole frame and a ragged region
    away from the origin, centre and jittered rays, the radio frame, the vacuum frame and
    time-sliced turbulence (where the mixed tracer bounds the coordinate-time error)."""
    cfg, t = c2
    regions = ((0, 0, 256, 256), (77, 101, 37, 29))
    for jitter in (0, 1):
        for x0, y0, w, h in regions:
            got = t.trace_tile(cfg["F"], x0, y0, w, h, t.opts(precision=prec, jitter=jitter))
            want = t.trace_tile(cfg["F"], x0, y0, w, h, t.opts(precision=prec, jitter=jitter, engine=bh.ENGINE_FRAME))
            assert_records_same(got, want, f"radio prec {prec} jitter {jitter} region {(x0, y0, w, h)}")
            assert (got["flow_steps"] > 0).any()
    vac = setup(shot("c1_vacuum_256.ini"))
    tv = tracer_for(vac)
    try:
        for x0, y0, w, h in regions:
            got = tv.trace_tile(vac["F"], x0, y0, w, h, tv.opts(precision=prec))
            want = tv.trace_tile(vac["F"], x0, y0, w, h, tv.opts(precision=prec, engine=bh.ENGINE_FRAME))
            assert_records_same(got, want, f"vacuum prec {prec} region {(x0, y0, w, h)}")
    finally:
        tv.close()
    ts = setup(shot("c2_radio_1024.ini"), W=96, H=96)
    grid = bh.turbulence(64, 4, 60.0, 0.3, 5.0, 30.0, 10.0, 3)
    tt = bh.Tracer(ts["S"], k2=ts["k2"], k2_l0=ts["l0"], k2_dl=ts["dl"], grid=grid, turb_half=60.0,
                   turb_t0=-1060.0, turb_dt=10.0)
    try:
        F = frame_for(ts["S"], shot("c2_radio_1024.ini"), 96, 96, t=15.0)
        got = tt.trace_tile(F, 5, 9, 83, 80, tt.opts(precision=prec))
        want = tt.trace_tile(F, 5, 9, 83, 80, tt.opts(precision=prec, engine=bh.ENGINE_FRAME))
        assert_records_same(got, want, f"time slices prec {prec}")
    finally:
        tt.close()


def test_trace_tile_errors(c2):
    """Both engines reject an empty region with the same error."""
    cfg, t = c2
    for engine in (bh.ENGINE_AUTO, bh.ENGINE_FRAME):
        for w, h in ((0, 16), (16, 0)):
            with pytest.raises(bh.BhError, match="empty tile"):
                t.trace_tile(cfg["F"], 0, 0, w, h, t.opts(engine=engine))
