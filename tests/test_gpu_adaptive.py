"""Adaptive sampling on the GPU (bh_render_adaptive) against the oracle's restatement
(or_render_adaptive) and against the plain renderer.

Bar: per-pixel sample counts equal and pixels bit-identical to the oracle (FP64 tracer, same
jittered samples, same deterministic sums); every pixel equal to the plain renderer's mean over the
same number of samples (both precisions), since pass 2 continues each pixel's sample sequence.
"""
import numpy as np
import pytest

import bhrt_amd as bh
from helpers import pixel_grid, setup, shot

pytestmark = pytest.mark.gpu


def tracer(cfg, spp, spp_max, target):
    st = bh.Settings(cfg["F"].W, cfg["F"].H, spp, spp_max, target)
    return bh.Tracer(cfg["S"], k2=cfg["k2"], k2_l0=cfg["l0"], k2_dl=cfg["dl"], grid=cfg["grid"],
                     turb_half=cfg["half"], settings=st)


def test_adaptive_matches_oracle(oracle_bhm):
    cfg = setup(shot("c2_radio_1024.ini"), W=24, H=24)
    cfg["F"].seed = cfg["OF"].seed = 0x1234567
    t = tracer(cfg, 4, 36, 0.03)
    img, ns, st = t.render_adaptive(cfg["F"], t.opts(jitter=1, spp=4))
    ref, ns_ref, total = oracle_bhm.render_adaptive(cfg["OS"], cfg["OF"], 4, 36, 0.03, jitter=True)
    print(f"samples per pixel: {np.bincount(ns.ravel())[4::8]} (4, 12, 20, ...); total {st.samples}")
    assert np.array_equal(ns, ns_ref)
    assert st.samples == total == int(ns.sum())
    assert (ns > 4).any()
    assert np.array_equal(img, ref.astype(np.float32))
    t.close()


@pytest.mark.parametrize("prec", [bh.PREC_FP64, bh.PREC_MIXED])
def test_adaptive_equals_plain_mean_over_same_samples(prec):
    cfg = setup(shot("c2_radio_1024.ini"), W=128, H=128)
    cfg["F"].seed = 0xBEEF
    t = tracer(cfg, 8, 64, 0.03)
    img, ns, st = t.render_adaptive(cfg["F"], t.opts(precision=prec, jitter=1, spp=8))
    assert st.samples == int(ns.sum())
    counts = np.unique(ns)
    print(f"prec {prec}: sample counts {dict(zip(counts.tolist(), np.bincount(ns.ravel())[counts].tolist()))}, "
          f"kernel {st.ms_kernel:.2f} ms")
    assert len(counts) > 1
    for m in counts:
        plain, _ = t.render(cfg["F"], t.opts(precision=prec, jitter=1, spp=int(m)))
        sel = ns == m
        assert np.array_equal(img[sel], plain[sel]), m
    t.close()


def test_adaptive_vacuum_takes_no_extra_samples():
    """Visible mode drives the decision with channel 1, which the vacuum tracer leaves at zero when
    the sky is off."""
    cfg = setup(shot("c1_vacuum_256.ini"), "sky.checker=0", W=64, H=64)
    t = tracer(cfg, 2, 32, 0.03)
    img, ns, st = t.render_adaptive(cfg["F"], t.opts(jitter=1, spp=2))
    assert (ns == 2).all() and st.samples == 2 * 64 * 64
    plain, _ = t.render(cfg["F"], t.opts(jitter=1, spp=2))
    assert np.array_equal(img, plain) and not img.any()
    t.close()


@pytest.mark.parametrize("prec", [bh.PREC_FP64, bh.PREC_MIXED])
def test_adaptive_sky_refines_edges(prec):
    """With the checkerboard sky the samples of a pixel differ only where it straddles a square's edge
    or the shadow's: those pixels (and their 3 x 3 neighbours) are refined, the rest keep 2 samples,
    and every pixel equals the plain render at its sample count."""
    cfg = setup(shot("c1_vacuum_256.ini"), W=64, H=64)
    t = tracer(cfg, 2, 32, 0.03)
    img, ns, st = t.render_adaptive(cfg["F"], t.opts(precision=prec, jitter=1, spp=2))
    assert st.samples == int(ns.sum())
    refined = (ns > 2).mean()
    print(f"prec {prec}: {refined:.3f} of the pixels refined, counts {np.unique(ns).tolist()}")
    assert 0.0 < refined < 0.6
    for m in np.unique(ns):
        plain, _ = t.render(cfg["F"], t.opts(precision=prec, jitter=1, spp=int(m)))
        sel = ns == m
        assert np.array_equal(img[sel], plain[sel]), m
    t.close()
