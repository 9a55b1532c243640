"""Shared setup for the parity tests: the same scene, tables and frame for the product (libbhrt)
and the oracle, built once by the product's host layer and copied field by field."""
import ctypes as C
import math
import os
import struct
import zlib

import numpy as np

import bhrt_amd as bh
from oracle_ctypes import OrFrame, make_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHOTS = os.path.join(ROOT, "spacetime-optics_amd", "shots")
D2R = math.pi / 180


def shot(name):
    return os.path.join(SHOTS, name)


def scene_dict(S):
    return {name: getattr(S, name) for name, _ in S._fields_}


def oracle_frame(F):
    O = OrFrame()
    C.memmove(C.addressof(O), C.addressof(F), C.sizeof(OrFrame))
    return O


_GRID_CACHE = {}


def setup(ini, sets="", with_grid=True, W=None, H=None):
    """(Scene, k2, (l0, dl), grid or None, info, OrScene, Frame, OrFrame) for a shot file; the
    frame is frame 0 of the shot at camera.width x camera.height (or W x H), centre rays."""
    S, k2, (l0, dl), info, opts = bh.build_scene(ini, sets)
    grid = None
    if with_grid and S.mode == bh.MODE_RADIO and S.flow_turb_amp > 0:
        key = (ini, sets)
        if key not in _GRID_CACHE:
            n = int(bh.config_str(ini, "turbulence.grid", sets) or 128)
            half = float(bh.config_str(ini, "turbulence.half_m", sets) or S.r_flow_max)
            corr = float(bh.config_str(ini, "turbulence.corr_m", sets) or 5.0)
            seed = int(bh.config_str(ini, "turbulence.seed", sets) or 1)
            _GRID_CACHE[key] = (bh.turbulence(n, 1, half, S.flow_turb_amp, corr, 30.0, 10.0, seed)[0], half)
        grid, half = _GRID_CACHE[key]
    else:
        half = 60.0
    Wc = W or int(bh.config_str(ini, "camera.width", sets))
    Hc = H or int(bh.config_str(ini, "camera.height", sets))
    r = float(bh.config_str(ini, "camera.r", sets))
    incl = float(bh.config_str(ini, "camera.incl", sets)) * D2R
    fov = float(bh.config_str(ini, "camera.fov_m", sets))
    F = bh.make_frame(S.a, r, incl, fov=fov, W=Wc, H=Hc, spp=1, seed=0x9E3779B9)
    OS = make_scene(scene_dict(S), k2=k2, grid=grid, turb_half=half)
    OS.k2_l0 = l0
    OS.k2_dl = dl
    return dict(S=S, k2=k2, l0=l0, dl=dl, grid=grid, half=half, info=info, OS=OS, F=F, OF=oracle_frame(F))


def tracer_for(cfg, device=0):
    return bh.Tracer(cfg["S"], k2=cfg["k2"], k2_l0=cfg["l0"], k2_dl=cfg["dl"], grid=cfg["grid"],
                     turb_half=cfg["half"], device=device)


def pixel_grid(w, h, x0=0, y0=0):
    py, px = np.mgrid[y0:y0 + h, x0:x0 + w]
    return px.ravel().astype(np.int32), py.ravel().astype(np.int32)


def read_png(path):
    """Decodes an 8-bit PNG written with Sub-filtered rows, checking every chunk's CRC."""
    data = open(path, "rb").read()
    assert data[:8] == bytes([137, 80, 78, 71, 13, 10, 26, 10])
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, comp, filt, inter) == (8, 0, 0, 0)
    ch = {0: 1, 2: 3, 6: 4}[ctype]
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    rows = np.frombuffer(raw, dtype=np.uint8).reshape(h, 1 + w * ch)
    out = np.zeros((h, w * ch), dtype=np.uint8)
    for y in range(h):
        assert rows[y, 0] == 1   # Sub filter: each byte adds the reconstructed byte one pixel left
        out[y] = (np.cumsum(rows[y, 1:].astype(np.int64).reshape(w, ch), axis=0) & 255).ravel()
    return out.reshape(h, w, ch) if ch > 1 else out.reshape(h, w)
