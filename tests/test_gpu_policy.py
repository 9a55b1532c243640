"""The mixed tracer's region policy, swept with the shipped options on the 1024^2 230 GHz frame and
measured against the FP64 frame (itself bit-identical to the oracle):

- the emission region's radius (`r_emit`; default: the radius outside which the flow model emits
  less than 1e-6 of its emission, 13.8 M for this scene), with the outer flow traced as vacuum
  (default) or with emission at tol_mid (`full_flow`): a smaller region loses emission, or moves
  emitting steps to the looser tolerance, and the frame error grows;
- the far-field tolerance (`tol_far`; default 1e-4 at this pixel size): 1e-6 takes more attempts
  for the same frame error.

Each case prints its line of the table (flux, pixel median and p99 error, attempts per ray)."""
import numpy as np
import pytest

import bhrt_amd as bh
from helpers import setup, shot, tracer_for
from test_gpu_mixed import frame_errors

pytestmark = pytest.mark.gpu

CASES = [   # (label, options)
    ("default (r_emit 13.8, outer flow vacuum, tol_far 1e-4)", {}),
    ("r_emit 10", {"r_emit": 10.0}),
    ("r_emit 8", {"r_emit": 8.0}),
    ("full_flow (emission to the flow's edge at tol_mid)", {"full_flow": 1}),
    ("full_flow, r_emit 8", {"full_flow": 1, "r_emit": 8.0}),
    ("full_flow, r_emit 5", {"full_flow": 1, "r_emit": 5.0}),
    ("tol_far 1e-6", {"tol_far": 1e-6}),
]


@pytest.fixture(scope="module")
def frame():
    cfg = setup(shot("c2_radio_1024.ini"))
    cfg["S"].n0 = 291464.1043528089   # the 0.6 Jy normalisation of this frame
    t = tracer_for(cfg)
    ref, st64 = t.render(cfg["F"], bh.Tracer.opts(precision=bh.PREC_FP64))
    yield t, cfg, ref, st64
    t.close()


def test_region_policy_sweep(frame):
    t, cfg, ref, st64 = frame
    res = {}
    print(f"\nFP64: {st64.step_attempts / st64.samples:.1f} attempts per ray")
    for label, kw in CASES:
        img, st = t.render(cfg["F"], bh.Tracer.opts(precision=bh.PREC_MIXED, **kw))
        flux, med, p99 = frame_errors(img, ref)
        att = st.step_attempts / st.samples
        res[label] = (flux, med, p99, att)
        print(f"| {label} | {att:.1f} | {flux:.1e} | {med:.1e} | {p99:.1e} |", flush=True)
    d = res[CASES[0][0]]
    # the default is within the bars with margin
    assert d[0] < 1e-4 and d[1] < 1e-4
    # a smaller emission region traced as vacuum loses the emission outside it
    assert res["r_emit 10"][0] > 3 * d[0] and res["r_emit 8"][0] > res["r_emit 10"][0]
    # the outer flow traced with emission: closer to FP64, more attempts
    ff = res["full_flow (emission to the flow's edge at tol_mid)"]
    assert ff[0] < d[0] * 2 and ff[3] > d[3] * 1.2
    # emitting steps at the looser tolerance: the error grows as the inner region shrinks
    assert res["full_flow, r_emit 5"][0] > 3 * ff[0]
    # a tighter far field costs attempts and does not buy accuracy at this pixel size
    tf = res["tol_far 1e-6"]
    assert tf[3] > d[3] * 1.1 and tf[1] < 2 * d[1]
