"""The oracle's adaptive sampling (oracle or_render_adaptive, restating render.cu:89-242 and
677-721 with deterministic sums), pinned without a GPU:
- with centre rays (no jitter) every sample of a pixel is the same ray, the luminance variance is
  zero and no pixel gets extra samples: the image is the plain spp mean;
- with an unreachable target (target_err huge) no pixel gets extra samples either;
- every pixel's value is the plain mean over samples 0..nsamp-1 of that pixel: pass 2 continues the
  pixel's R2 + PCG sequence and adds in sample order (checked against or_trace_pixels at each nsamp);
- extra samples come in chunks of 8, never exceed spp_max, and a smaller target never gives fewer.
"""
import numpy as np

from helpers import pixel_grid, setup, shot


def small_c2(W=16, H=16):
    cfg = setup(shot("c2_radio_1024.ini"), W=W, H=H)
    OF = cfg["OF"]
    OF.seed = 0x1234567
    return cfg, OF


def test_no_extra_samples_without_variance(oracle_bhm):
    cfg, OF = small_c2(12, 12)
    img, ns, total = oracle_bhm.render_adaptive(cfg["OS"], OF, 4, 36, 0.03, jitter=False)
    assert (ns == 4).all() and total == 4 * 144
    px, py = pixel_grid(12, 12)
    want, _, _ = oracle_bhm.trace_pixels(cfg["OS"], OF, px, py, spp=4, jitter=False)
    assert np.array_equal(img.reshape(-1, 4), want)


def test_adaptive_pixels_are_plain_means_over_their_samples(oracle_bhm):
    cfg, OF = small_c2()
    img, ns, total = oracle_bhm.render_adaptive(cfg["OS"], OF, 4, 36, 0.03, jitter=True)
    assert total == int(ns.sum())
    assert ((ns - 4) % 8 == 0).all() and ns.max() <= 4 + 32 and ns.min() >= 4
    assert (ns > 4).any() and (ns == 4).any()   # the frame has both
    px, py = pixel_grid(16, 16)
    flat = ns.ravel()
    for m in np.unique(flat):
        sel = flat == m
        want, _, _ = oracle_bhm.trace_pixels(cfg["OS"], OF, px[sel], py[sel], spp=int(m), jitter=True)
        assert np.array_equal(img.reshape(-1, 4)[sel], want), m
    # unreachable target: no extra samples; a tighter target: never fewer samples
    _, ns_loose, _ = oracle_bhm.render_adaptive(cfg["OS"], OF, 4, 36, 1e30, jitter=True)
    assert (ns_loose == 4).all()
    _, ns_tight, _ = oracle_bhm.render_adaptive(cfg["OS"], OF, 4, 36, 0.01, jitter=True)
    assert (ns_tight >= ns).all() and (ns_tight > ns).any()


def test_extra_samples_reach_the_target_error(oracle_bhm):
    """What the sample counts mean: a pixel asks for the number of samples at which the standard
    error of its mean luminance falls to target x (its mean + 3 % of the frame's mean), capped at
    spp_max, and the request is spread over its 3 x 3 neighbourhood in chunks of 8. The statistics
    here are computed in FP64 from the individual samples (not the float running sums), so a wrong
    factor, a missing 1/n or a wrong reference level fails by whole chunks; rounding differences of
    the float sums may move a request across a chunk boundary in a few pixels."""
    W = H = 16
    spp, spp_max, target = 4, 36, 0.03
    cfg, OF = small_c2(W, H)
    _, ns, _ = oracle_bhm.render_adaptive(cfg["OS"], OF, spp, spp_max, target, jitter=True)
    Y = np.zeros((H, W, spp))
    for y in range(H):
        for x in range(W):
            for s in range(spp):
                Y[y, x, s] = np.float32(oracle_bhm.trace_sample(cfg["OS"], OF, x, y, s).I[0])
    mean = Y.mean(axis=2)
    se = np.sqrt(Y.var(axis=2) / spp)
    yref = 0.03 * mean.mean()
    need = spp * (se / (target * (mean + yref))) ** 2   # samples at which se * sqrt(spp / m) = target level
    extra = np.clip(np.ceil(need - 1e-9) - spp, 0, spp_max - spp)
    pad = np.pad(extra, 1, mode="edge")
    neigh = np.max([pad[dy:dy + H, dx:dx + W] for dy in range(3) for dx in range(3)], axis=0)
    want = spp + 8 * np.ceil(neigh / 8)
    got = ns.astype(float)
    assert (got > spp).sum() >= 10 and (got == spp).sum() >= 10
    assert np.abs(got - want).max() <= 8 and (got != want).mean() < 0.03, np.argwhere(got != want)[:5]
