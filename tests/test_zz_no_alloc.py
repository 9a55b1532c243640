"""bh_render_noalloc (the frame kernel writing into caller memory, partials in the library's fixed
device array) against bh_render on the frame kernel: same pixels bit for bit and the same counters,
on two calls in a row, the second on a smaller frame, in both precisions."""
import ctypes as C

import numpy as np
import pytest

import bhrt_amd as bh
from helpers import setup, shot, tracer_for

pytestmark = pytest.mark.gpu

COUNTERS = ("samples", "step_attempts", "steps_accepted", "flow_steps", "step_limited", "fp64_steps", "flux_sum")


class DeviceBuffer:
    """A device allocation made by the test (the entry point under test allocates nothing)."""

    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptr = C.c_void_p()
        self.nbytes = nbytes
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes)) == 0
        assert self.hip.hipMemset(self.ptr, 0xFF, C.c_size_t(nbytes)) == 0   # NaN pattern: unwritten pixels show

    def read(self, shape):
        out = np.empty(shape, dtype=np.float32)
        assert out.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        if self.ptr.value:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


@pytest.fixture(scope="module")
def radio():
    cfg = setup(shot("c2_radio_1024.ini"), W=1024, H=1024)
    t = tracer_for(cfg)
    yield cfg, t
    t.close()


def frame_of(cfg, W, H):
    F = bh.Frame.from_buffer_copy(cfg["F"])
    # same camera and field of view, fewer pixels: the pixel scale grows with the shrink
    F.pix = cfg["F"].pix * cfg["F"].W / W
    F.W, F.H = W, H
    return F


@pytest.mark.parametrize("prec", [bh.PREC_FP64, bh.PREC_MIXED])
def test_noalloc_matches_render(radio, prec):
    cfg, t = radio
    buf = DeviceBuffer(384 * 256 * 16)
    try:
        for W, H in ((384, 256), (150, 97)):   # second call smaller (ragged tiles)
            F = frame_of(cfg, W, H)
            print(f"noalloc: precision {prec}, frame {W}x{H}", flush=True)
            o = bh.Tracer.opts(precision=prec, engine=bh.ENGINE_FRAME)
            ref, st_ref = t.render(F, o)
            st = t.render_noalloc(F, buf.ptr.value, o)
            img = buf.read((H, W, 4))
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), f"{W}x{H}: pixels differ"
            for k in COUNTERS:
                assert getattr(st, k) == getattr(st_ref, k), (W, H, k)
            assert st.step_attempts > 0 and st.flux_sum > 0
    finally:
        buf.free()


def test_noalloc_rejects_oversized_frames(radio):
    cfg, t = radio
    F = frame_of(cfg, 4112, 4096)   # 257 x 256 tiles -> 65792 workgroups, over the 65536 limit
    with pytest.raises(bh.BhError, match="scratch array holds 65536"):
        t.render_noalloc(F, 0x1000, bh.Tracer.opts())   # rejected on the host before any launch
    with pytest.raises(bh.BhError, match="d_px is NULL"):
        t.render_noalloc(frame_of(cfg, 64, 64), 0, bh.Tracer.opts())
