"""Benchmark: Mrays/s on the 1024 x 1024, 230 GHz Kerr frame (a = 0.94, i = 17 deg, 128^3 seeded
log-normal turbulence, flux normalised to 0.6 Jy), plain FP64 tracer vs mixed FP32/FP64 + MLP
emissivity, on one MI355X.

A step is one full-frame trace (1,048,576 pixel-centre rays) with the scene, tables and turbulence
grid already resident in HBM; the frame stays in device memory (bh_render_resident). Prints one
JSON line. The CPU baseline is the oracle's glibc build (the reference integrator's arithmetic,
restated in C++) on a bounded random pixel sample of the same frame, over all host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--modes fp64,mixed]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("spacetime-optics_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import bhrt_amd as bh  # noqa: E402

INI = os.path.join(ROOT, "spacetime-optics_amd", "shots", "c2_radio_1024.ini")
WORKLOAD = ("c2_radio_1024: 1024x1024 Kerr frame at 230 GHz, a=0.94, i=17 deg, r_obs=1000 M, FoV +-20 M, "
            "DP54 tol 1e-10, H/R=0.3, Theta_e~r^-0.84, beta=10, 128^3 log-normal turbulence (sigma 0.3, 5 M), "
            "flux normalised to 0.6 Jy, pixel-centre rays")
# Algorithmic work per unit (one DP54 step attempt; extra work of an accepted step inside the flow)
FLOP_PER_ATTEMPT = 600.0
FLOP_PER_FLOW_STEP = 200.0
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (= FP64 matrix) peak


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def bench_mode(tracer, F, opts, steps, warmup):
    for _ in range(warmup):
        tracer.render_resident(F, opts)
    kernel_ms, stats = [], None
    t0 = time.perf_counter()
    for _ in range(steps):
        _, st = tracer.render_resident(F, opts)   # synchronous: returns after the frame is done
        kernel_ms.append(st.ms_kernel)
        stats = st
    wall = time.perf_counter() - t0
    rays = F.W * F.H
    work = FLOP_PER_ATTEMPT * stats.step_attempts + FLOP_PER_FLOW_STEP * stats.flow_steps
    avg_kernel_s = float(np.mean(kernel_ms)) * 1e-3
    return {
        "mrays_per_s": rays * steps / wall / 1e6,
        "ms_per_step": wall / steps * 1e3,
        "kernel_ms": float(np.mean(kernel_ms)),
        "kernel_mrays_per_s": rays / avg_kernel_s / 1e6,
        "step_attempts": int(stats.step_attempts),
        "steps_per_ray": stats.step_attempts / rays,
        "flow_steps": int(stats.flow_steps),
        "fp64_attempts": int(stats.fp64_steps),
        "step_limited": int(stats.step_limited),
        "flux_sum": stats.flux_sum,
        "work_flop": work,
        "achieved_tflops": work / avg_kernel_s / 1e12,
    }


def cpu_baseline(S, k2, l0, dl, F, n_pix, seed=11):
    """Oracle (glibc build) on a seeded random pixel sample of the same frame, all host cores."""
    from oracle_ctypes import OrFrame, Oracle, make_scene
    import ctypes as C

    o = Oracle("libm")
    grid = bh.turbulence(int(bh.config_str(INI, "turbulence.grid")), 1, float(bh.config_str(INI, "turbulence.half_m")),
                         S.flow_turb_amp, float(bh.config_str(INI, "turbulence.corr_m")), 30.0, 10.0,
                         int(bh.config_str(INI, "turbulence.seed")))[0]
    OS = make_scene({name: getattr(S, name) for name, _ in S._fields_}, k2=k2, grid=grid,
                    turb_half=float(bh.config_str(INI, "turbulence.half_m")))
    OS.k2_l0, OS.k2_dl = l0, dl
    OF = OrFrame()
    C.memmove(C.addressof(OF), C.addressof(F), C.sizeof(OrFrame))
    rng = np.random.default_rng(seed)
    idx = rng.choice(F.W * F.H, size=n_pix, replace=False)
    px, py = (idx % F.W).astype(np.int32), (idx // F.W).astype(np.int32)
    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    _, rays, used = o.trace_pixels(OS, OF, px, py, spp=1, jitter=False, nthreads=threads)
    wall = time.perf_counter() - t0
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": n_pix / wall / 1e6, "unit": "Mrays/s", "cores": used, "kind": "port",
            "sample": f"{n_pix} random pixel-centre rays of the same 1024^2 frame (seed {seed}), "
                      f"{wall:.1f} s wall on {used} threads ({cpu_model}); glibc build of the oracle, "
                      f"mean {rays['steps'].mean():.1f} DP54 steps/ray"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-pixels", type=int, default=1048576)
    ap.add_argument("--modes", default="fp64,mixed")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (rays shard with no "
              f"exchange; see DESIGN.md for how the path would shard)", file=sys.stderr)
        sys.exit(2)
    if bh.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)

    log("bench: building scene, 128^3 turbulence, flux normalisation to 0.6 Jy")
    tracer, S, F, opts, jy = bh.Tracer.from_ini(INI, normalize=True)
    S = tracer.scene()
    _, k2, (l0, dl), _, _ = bh.build_scene(INI)
    log(f"bench: n0 = {S.n0:.6g} cm^-3, flux {jy:.4f} Jy")

    results = {}
    modes = [m for m in args.modes.split(",") if m]
    for m in modes:
        o = bh.Tracer.opts(precision=bh.PREC_FP64 if m == "fp64" else bh.PREC_MIXED,
                           emissivity=bh.EMIS_MLP if m == "mixed_mlp" else bh.EMIS_EXACT, jitter=0, spp=1)
        try:
            results[m] = bench_mode(tracer, F, o, args.steps, args.warmup)
            log(f"bench: {m}: {results[m]['mrays_per_s']:.3f} Mrays/s, kernel {results[m]['kernel_ms']:.2f} ms, "
                f"{results[m]['achieved_tflops']:.2f} TFLOP/s")
        except bh.BhError as e:
            log(f"bench: {m}: not available ({e})")

    # accuracy of the mixed tracer against the FP64 tracer (bit-identical to the oracle) on this frame
    pix_err = None
    if "mixed" in results:
        ref, _ = tracer.render(F, bh.Tracer.opts(precision=bh.PREC_FP64, jitter=0, spp=1))
        mix, _ = tracer.render(F, bh.Tracer.opts(precision=bh.PREC_MIXED, emissivity=bh.EMIS_EXACT, jitter=0, spp=1))
        Ia, Ib = mix[..., 0].astype(np.float64), ref[..., 0].astype(np.float64)
        mask = Ib > 1e-3 * Ib.max()
        rel = np.abs(Ia[mask] - Ib[mask]) / Ib[mask]
        pix_err = {"flux_rel": abs(Ia.sum() - Ib.sum()) / Ib.sum(), "pixel_median_rel": float(np.median(rel)),
                   "pixel_p99_rel": float(np.percentile(rel, 99)), "pixels": int(mask.sum())}
        log(f"bench: mixed vs fp64 frame: {pix_err}")

    head = results.get("mixed") or results["fp64"]
    head_mode = "mixed" if "mixed" in results else "fp64"
    fp64 = results["fp64"]
    flux_err = None
    if "mixed" in results:
        flux_err = abs(results["mixed"]["flux_sum"] - fp64["flux_sum"]) / abs(fp64["flux_sum"])

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_c2.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(head_mode)
        except (OSError, ValueError):
            traffic = None

    cpu = None
    if not args.no_cpu:
        log(f"bench: CPU baseline on {args.cpu_pixels} pixels")
        cpu = cpu_baseline(S, k2, l0, dl, F, args.cpu_pixels)

    achieved = head["achieved_tflops"]
    line = {
        "metric": "Mrays/s, 1024^2 230GHz Kerr frame: plain FP64 vs mixed+MLP at <0.1% flux err",
        "value": round(head["mrays_per_s"], 4),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head_mode == "fp64" else "f32/f64 + bf16 MLP",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "precision": head_mode, "rays_per_step": F.W * F.H,
                   "n0_cm3": S.n0, "flux_jy": jy},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic,
                     "note": "FP64 vector ALU roof; algorithmic work = 600 flop-units per DP54 step attempt "
                             "+ 200 per in-flow step, counted by the kernel"},
        "cpu_baseline": cpu,
        "fp64": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in fp64.items()},
    }
    if "mixed" in results:
        line["mixed"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in results["mixed"].items()}
        line["speedup_mixed_vs_fp64"] = round(results["mixed"]["mrays_per_s"] / fp64["mrays_per_s"], 3)
        line["flux_rel_err_mixed_vs_fp64"] = flux_err
        line["mixed_vs_fp64_frame_error"] = pix_err
    if cpu:
        line["gpu_vs_cpu"] = round(head["mrays_per_s"] / cpu["value"], 1)
    tracer.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
