"""Benchmark: Mrays/s on the 1024 x 1024, 230 GHz Kerr frame (a = 0.94, i = 17 deg, 128^3 seeded
log-normal turbulence, flux normalised to 0.6 Jy), plain FP64 tracer vs mixed FP32/FP64 + MLP
emissivity, on one MI355X.

A step is one full-frame trace (1,048,576 pixel-centre rays) with the scene, tables and turbulence
grid already resident in HBM; the frame stays in device memory (bh_render_resident). Prints one
JSON line. The CPU baseline is the oracle's glibc build (the reference integrator's arithmetic,
restated in C++) on a bounded random pixel sample of the same frame, over all host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--modes fp64,mixed,...] [--no-autotune]
                    [--batch-frames N]

Modes: fp64 / mixed / mixed_mlp run one launch per frame (bh_render_resident); p_* run the same
tracers through the persistent batched kernel (bh_trace_batch, graph replay); p_mixed_f32far is
the mixed tracer with FP32 stages outside the flow, p_mixed_fullflow the mixed tracer with emission,
absorption and the transfer step cap over the whole flow (default: inside the emission region). `value` is the
metric's own mode, mixed precision + MLP emissivity: the fastest surrogate mode whose frame is within 0.1% flux
and 1% median per-pixel error of the FP64 frame (the fastest passing mode when the surrogate modes are skipped);
`exact_emission` reports the fastest passing mode without the surrogate beside it. The speed-up over the
fastest FP64 mode is split into its two sources: fewer step attempts per ray (`attempts_ratio`) and less
time per attempt (`per_attempt_speedup`).
`animation_batch` reports the same mode over N frames of an animation (camera azimuth stepped per
frame) in one persistent launch, beside the one-frame-per-step headline.
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
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector peak
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector peak


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def summarise(kernel_ms, wall, steps, rays, stats, prec):
    """Throughput, counters and the roofline position of one mode. Work is counted in flop-units
    per DP54 step attempt and per in-flow step; attempts the kernel ran in FP32 (mixed far field)
    and the mixed tracer's FP32 log-space emission are priced at the FP32 vector peak, the rest at
    the FP64 peak, so the roof is the time the same work would take at the two peaks."""
    avg_kernel_s = float(np.mean(kernel_ms)) * 1e-3
    att = stats.step_attempts
    a64 = stats.fp64_steps if prec == bh.PREC_MIXED else att
    w64 = FLOP_PER_ATTEMPT * a64 + (FLOP_PER_FLOW_STEP * stats.flow_steps if prec == bh.PREC_FP64 else 0.0)
    w32 = FLOP_PER_ATTEMPT * (att - a64) + (FLOP_PER_FLOW_STEP * stats.flow_steps if prec == bh.PREC_MIXED else 0.0)
    t_roof = w64 / (FP64_PEAK_TFLOPS * 1e12) + w32 / (FP32_PEAK_TFLOPS * 1e12)
    return {
        "mrays_per_s": rays * steps / wall / 1e6,
        "ms_per_step": wall / steps * 1e3,
        "kernel_ms": float(np.mean(kernel_ms)),
        "kernel_mrays_per_s": rays / avg_kernel_s / 1e6,
        "step_attempts": int(att),
        "steps_per_ray": att / rays,
        "flow_steps": int(stats.flow_steps),
        "fp64_attempts": int(a64),
        "step_limited": int(stats.step_limited),
        "flux_sum": stats.flux_sum,
        "work_flop": w64 + w32,
        "achieved_tflops": (w64 + w32) / avg_kernel_s / 1e12,
        "peak_tflops": (w64 + w32) / t_roof / 1e12,
        "roof_frac": t_roof / avg_kernel_s,
    }


def bench_mode(tracer, F, S, opts, steps, warmup, persistent, cfg=None):
    """Times `steps` full-frame traces; persistent: one-frame bh_trace_batch (graph replay, launch
    configuration cfg), else bh_render_resident. Both return after the frame is done."""
    run = ((lambda: tracer.trace_batch([(S, F)], opts, cfg=cfg, resident=True)) if persistent
           else (lambda: tracer.render_resident(F, opts)))
    for _ in range(warmup):
        run()
    kernel_ms, stats = [], None
    t0 = time.perf_counter()
    for _ in range(steps):
        _, st = run()
        kernel_ms.append(st.ms_kernel)
        stats = st
    wall = time.perf_counter() - t0
    return summarise(kernel_ms, wall, steps, F.W * F.H, stats, opts.precision)


def calib_copy(mib):
    """One device-to-device copy of `mib` MiB (the runtime's copy kernel): a transfer of known size
    that scales the HBM counters of a profiling run (FETCH_SIZE / WRITE_SIZE)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = int(mib) << 20
    a, b = C.c_void_p(), C.c_void_p()
    ok = hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    ok = ok and hip.hipMemset(a, 1, C.c_size_t(n)) == 0 and hip.hipDeviceSynchronize() == 0
    ok = ok and hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0 and hip.hipDeviceSynchronize() == 0
    for p in (a, b):
        if p.value:
            hip.hipFree(p)
    if not ok:
        raise RuntimeError("calibration copy failed")
    log(f"bench: calibration copy of {mib} MiB device to device")


def frame_error(img, ref):
    """Flux and per-pixel errors of an image against the FP64 image (pixels above 1e-3 of the peak)."""
    Ia, Ib = img[..., 0].astype(np.float64), ref[..., 0].astype(np.float64)
    mask = Ib > 1e-3 * Ib.max()
    rel = np.abs(Ia[mask] - Ib[mask]) / Ib[mask]
    return {"flux_rel": float(abs(Ia.sum() - Ib.sum()) / Ib.sum()), "pixel_median_rel": float(np.median(rel)),
            "pixel_p99_rel": float(np.percentile(rel, 99)), "pixels": int(mask.sum())}


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


MODES = {   # name: (precision, emissivity, persistent, FP32 far-field stages, emission over the whole flow)
    "fp64": (bh.PREC_FP64, bh.EMIS_EXACT, False, 0, 0),
    "mixed": (bh.PREC_MIXED, bh.EMIS_EXACT, False, 0, 0),
    "mixed_mlp": (bh.PREC_MIXED, bh.EMIS_MLP, False, 0, 0),
    "p_fp64": (bh.PREC_FP64, bh.EMIS_EXACT, True, 0, 0),
    "p_mixed": (bh.PREC_MIXED, bh.EMIS_EXACT, True, 0, 0),
    "p_mixed_mlp": (bh.PREC_MIXED, bh.EMIS_MLP, True, 0, 0),
    "p_mixed_f32far": (bh.PREC_MIXED, bh.EMIS_EXACT, True, 1, 0),
    "p_mixed_fullflow": (bh.PREC_MIXED, bh.EMIS_EXACT, True, 0, 1),
}
FLUX_GATE, PIXEL_GATE = 1e-3, 1e-2   # a mode counts only below 0.1% flux and 1% median per-pixel error


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-pixels", type=int, default=1048576)
    ap.add_argument("--modes", default=",".join(MODES))
    ap.add_argument("--no-autotune", action="store_true", help="time the persistent modes with the default launch")
    ap.add_argument("--no-mlp-train", action="store_true", help="skip MLP modes")
    ap.add_argument("--mlp-save", default="", help="write the trained surrogate (parameters, normalisation) to this .npz")
    ap.add_argument("--mlp-load", default="", help="load the surrogate from this .npz instead of training (profiling)")
    ap.add_argument("--mlp-epochs", type=int, default=0,
                    help="training epochs of the surrogate (0: the library's default; fewer for a short kernel trace)")
    ap.add_argument("--copy-calib", type=int, default=0, help="MiB of a device copy before the modes (profiling)")
    ap.add_argument("--no-adaptive", action="store_true", help="skip the adaptive-sampling render")
    ap.add_argument("--batch-frames", type=int, default=4,
                    help="frames of the animation batch traced in one persistent launch (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the tracer runs on one GPU; --gpus {args.gpus} is not supported (rays shard with no "
              f"exchange: one process per GPU would take frames or tile ranges from a shared counter)", file=sys.stderr)
        sys.exit(2)
    if bh.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    modes = [m for m in args.modes.split(",") if m]
    for m in modes:
        if m not in MODES:
            print(f"bench.py: unknown mode {m}; known: {', '.join(MODES)}", file=sys.stderr)
            sys.exit(2)

    log("bench: building scene, 128^3 turbulence, flux normalisation to 0.6 Jy")
    tracer, S, F, opts, jy = bh.Tracer.from_ini(INI, normalize=True)
    S = tracer.scene()
    _, k2, (l0, dl), _, _ = bh.build_scene(INI)
    log(f"bench: n0 = {S.n0:.6g} cm^-3, flux {jy:.4f} Jy")

    setup = {}
    if args.no_mlp_train:
        modes = [m for m in modes if MODES[m][1] != bh.EMIS_MLP]
    if any(MODES[m][1] == bh.EMIS_MLP for m in modes):
        t0 = time.perf_counter()
        if args.mlp_load:
            z = np.load(args.mlp_load)
            tracer.mlp_set(z["params"], z["norm"])
            rep = {"loss_first": None, "loss_last": None}
            setup["mlp_loaded"] = args.mlp_load
        else:
            rep = tracer.mlp_train(epochs=args.mlp_epochs)
        if args.mlp_save:
            p, nrm = tracer.mlp_get()
            np.savez(args.mlp_save, params=p, norm=nrm)
        t1 = time.perf_counter()
        bake = tracer.mlp_bake()
        t2 = time.perf_counter()
        setup["mlp_train_s"] = round(t1 - t0, 3)
        setup["mlp_bake_s"] = round(t2 - t1, 3)
        setup["mlp_loss_first_last"] = [rep["loss_first"], rep["loss_last"]]
        setup["mlp_table_cells"] = bake["cells"]
        setup["mlp_hot_cells"] = bake["hot_cells"]
        setup["mlp_failed_hot_cells"] = bake["failed_hot_cells"]
        log(f"bench: surrogate trained in {t1 - t0:.2f} s, baked in {t2 - t1:.2f} s: {setup}")
    if args.copy_calib > 0:
        calib_copy(args.copy_calib)

    results = {}
    for m in modes:
        prec, emis, persistent, f32, full = MODES[m]
        o = bh.Tracer.opts(precision=prec, emissivity=emis, jitter=0, spp=1, far_fp32=f32, full_flow=full,
                           engine=bh.ENGINE_AUTO if persistent else bh.ENGINE_FRAME)
        try:
            cfg = None
            if persistent and not args.no_autotune:
                t0 = time.perf_counter()
                cfg, ms, tried = tracer.autotune([(S, F)], o)
                setup[f"autotune_{m}"] = {"best": cfg.as_dict(), "ms": round(ms, 3), "configs": len(tried),
                                          "s": round(time.perf_counter() - t0, 2)}
                log(f"bench: {m}: autotuned {cfg.as_dict()} ({ms:.2f} ms, {len(tried)} configs)")
            results[m] = bench_mode(tracer, F, S, o, args.steps, args.warmup, persistent, cfg)
            log(f"bench: {m}: {results[m]['mrays_per_s']:.3f} Mrays/s, kernel {results[m]['kernel_ms']:.2f} ms, "
                f"{results[m]['achieved_tflops']:.2f} TFLOP/s ({100 * results[m]['roof_frac']:.1f}% of roof)")
        except bh.BhError as e:
            log(f"bench: {m}: not available ({e})")

    # accuracy of every mode against the FP64 tracer (bit-identical to the oracle) on this frame
    ref, _ = tracer.render(F, bh.Tracer.opts(precision=bh.PREC_FP64, jitter=0, spp=1, engine=bh.ENGINE_FRAME))
    for m, r in results.items():
        prec, emis, persistent, f32, full = MODES[m]
        o = bh.Tracer.opts(precision=prec, emissivity=emis, jitter=0, spp=1, far_fp32=f32, full_flow=full,
                           engine=bh.ENGINE_AUTO if persistent else bh.ENGINE_FRAME)
        cfg = None
        if f"autotune_{m}" in setup:
            cfg = bh.BatchCfg(**setup[f"autotune_{m}"]["best"])
        img = tracer.trace_batch([(S, F)], o, cfg=cfg)[0][0] if persistent else tracer.render(F, o)[0]
        r["error_vs_fp64"] = frame_error(img, ref)
        r["passes_gate"] = bool(r["error_vs_fp64"]["flux_rel"] < FLUX_GATE and
                                r["error_vs_fp64"]["pixel_median_rel"] < PIXEL_GATE)
        log(f"bench: {m} vs fp64: {r['error_vs_fp64']}")

    passing = [m for m in results if results[m]["passes_gate"]]
    passing_mlp = [m for m in passing if MODES[m][1] == bh.EMIS_MLP]
    passing_exact = [m for m in passing if MODES[m][1] == bh.EMIS_EXACT]
    head_mode = max(passing_mlp or passing, key=lambda m: results[m]["mrays_per_s"])
    head = results[head_mode]
    fp64 = results.get("fp64")

    # bytes the dominant kernel moves between L2 and memory per launch, from the committed counter
    # profile of the same mode (FETCH_SIZE + WRITE_SIZE, calibrated on a 1 GiB copy)
    traffic, traffic_note = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_c2.json")
    if os.path.exists(tpath):
        try:
            t = json.load(open(tpath)).get(head_mode)
            if t:
                traffic = int(t["read_bytes"] + t["write_bytes"])
                traffic_note = (f"{t['kernel']}: {t['read_bytes'] / 1e6:.1f} MB read + {t['write_bytes'] / 1e6:.1f} MB "
                                f"written per launch ({t['counters']}; profiles/c2_radio_1024_counters.md)")
        except (OSError, ValueError, KeyError):
            traffic = None

    # bhrt's production render of this frame (adaptive passes, spp 8 -> up to 128, jittered), for reference
    adaptive = {}
    if not args.no_adaptive:
        for name, prec in (("fp64", bh.PREC_FP64), ("mixed", bh.PREC_MIXED)):
            o = bh.Tracer.opts(precision=prec, jitter=1, spp=8)
            tracer.render_adaptive(F, o)   # warm-up (first launch of these kernels)
            _, ns, st = tracer.render_adaptive(F, o)
            adaptive[name] = {"ms": round(st.ms_kernel, 2), "samples": int(st.samples),
                              "msamples_per_s": round(st.samples / st.ms_kernel * 1e-3, 2),
                              "pixels_refined": int((ns > 8).sum())}
            log(f"bench: adaptive render ({name}): {st.ms_kernel:.1f} ms, {st.samples / 1e6:.2f} M samples "
                f"({st.samples / st.ms_kernel * 1e-3:.1f} Msamples/s), {int((ns > 8).sum())} pixels refined")

    # the batching path: an animation of several frames of the same workload (camera azimuth stepped
    # by 0.05 rad per frame) in one persistent launch, with the head mode's launch configuration; the
    # longest rays of one frame overlap the bulk of the next, so the per-frame time drops below the
    # single-frame one (reported beside the headline, which stays one frame per step)
    batch = None
    if args.batch_frames > 0 and MODES[head_mode][2]:
        prec_h, emis_h, _, f32_h, full_h = MODES[head_mode]
        o = bh.Tracer.opts(precision=prec_h, emissivity=emis_h, jitter=0, spp=1, far_fp32=f32_h, full_flow=full_h)
        cfg = bh.BatchCfg(**setup[f"autotune_{head_mode}"]["best"]) if f"autotune_{head_mode}" in setup else None
        frames = []
        for k in range(args.batch_frames):
            Fk = type(F).from_buffer_copy(F)
            Fk.ph = F.ph + 0.05 * k
            frames.append((S, Fk))
        tracer.trace_batch(frames, o, cfg=cfg, resident=True)
        kms = []
        for _ in range(max(args.steps, 1)):
            _, st = tracer.trace_batch(frames, o, cfg=cfg, resident=True)
            kms.append(st.ms_kernel)
        ms = float(np.mean(kms))
        rays = args.batch_frames * F.W * F.H
        batch = {"mode": head_mode, "frames": args.batch_frames, "kernel_ms": round(ms, 3),
                 "ms_per_frame": round(ms / args.batch_frames, 3), "mrays_per_s": round(rays / ms * 1e-3, 2)}
        log(f"bench: batch of {args.batch_frames} frames ({head_mode}): {ms:.2f} ms, "
            f"{ms / args.batch_frames:.3f} ms per frame, {rays / ms * 1e-3:.1f} Mrays/s")

    cpu = None
    if not args.no_cpu:
        log(f"bench: CPU baseline on {args.cpu_pixels} pixels")
        cpu = cpu_baseline(S, k2, l0, dl, F, args.cpu_pixels)

    prec = MODES[head_mode][0]
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
        "dtype": "f64" if prec == bh.PREC_FP64 else ("f32/f64 + bf16-MLP table" if MODES[head_mode][1] else "f32/f64"),
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "mode": head_mode, "rays_per_step": F.W * F.H, "n0_cm3": S.n0,
                   "flux_jy": jy},
        "roofline": {"bound": "valu", "achieved": round(head["achieved_tflops"], 3),
                     "peak": round(head["peak_tflops"], 2), "unit": "TFLOP/s", "frac": round(head["roof_frac"], 4),
                     "traffic": traffic,
                     "traffic_note": traffic_note,
                     "note": "vector-ALU roof (FP64 78.6, FP32 157.3 TFLOP/s, weighted by the work each precision "
                             "did); work = 600 flop-units per DP54 step attempt + 200 per in-flow step, counted "
                             "by the kernel; HIP-event kernel time"},
        "cpu_baseline": cpu,
        "modes": {m: {k: (round(v, 5) if isinstance(v, float) else v) for k, v in r.items()} for m, r in results.items()},
        "setup": setup,
        "adaptive_render_spp8_128": adaptive or None,
        "animation_batch": batch,
    }
    if fp64:
        line["speedup_vs_fp64"] = round(head["mrays_per_s"] / fp64["mrays_per_s"], 3)
    fp64_modes = [m for m in results if MODES[m][0] == bh.PREC_FP64]
    if fp64_modes:
        best64 = max(fp64_modes, key=lambda m: results[m]["mrays_per_s"])
        line["speedup_vs_best_fp64"] = {"mode": best64,
                                        "x": round(head["mrays_per_s"] / results[best64]["mrays_per_s"], 3)}
        base = results[best64]
        # where the speed-up comes from: attempts per ray, and kernel time per attempt
        att_ratio = base["steps_per_ray"] / head["steps_per_ray"]
        per_att = (base["kernel_ms"] / base["step_attempts"]) / (head["kernel_ms"] / head["step_attempts"])
        line["attempts_ratio"] = round(att_ratio, 3)
        line["per_attempt_speedup"] = round(per_att, 3)
    if passing_exact:
        ex = max(passing_exact, key=lambda m: results[m]["mrays_per_s"])
        line["exact_emission"] = {"mode": ex, "mrays_per_s": round(results[ex]["mrays_per_s"], 3),
                                  "kernel_ms": round(results[ex]["kernel_ms"], 4),
                                  "surrogate_gain": round(head["mrays_per_s"] / results[ex]["mrays_per_s"], 4)
                                  if head_mode in passing_mlp else None}
    if cpu:
        line["gpu_vs_cpu"] = round(head["mrays_per_s"] / cpu["value"], 1)
    tracer.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
