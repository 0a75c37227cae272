"""Benchmark: Mrays/s of the Kerr geodesic renderer on one MI355X.

Workload: 1920x1080, spin 0.998, inclination 85 deg, observer at 100 M,
FOV 20 deg, DP5(4) tolerance 1e-9, up to 20000 steps per ray, disk from the
ISCO to 30 M, FBM noise 6 octaves seed 1, bloom sigma 12 px, ACES + sRGB.
One step = one full frame: trace (FP64) + bloom + display transform, with
every buffer resident in HBM.  Rays are synthetic (pinhole camera), spp 1.

Prints one JSON line.  `value` = rays per second over a run of whole frames
(Mrays/s) issued as a renderer streams them: consecutive frames take two or
three HIP streams in turn, each with its own HDR, bloom and image buffers, so
several frames are in flight (as in the reference's double-buffered frame
loop) and one frame's long-ray tail overlaps the next frame's bulk; the
stream count and launch options are tuned for that (`pipelined`).  `serial`
gives the same frames one at a time (`frame_latency_ms`), `trace_mrays_s`
the tracer alone.  `roofline` prices the tracer's
algorithmic FP64 FLOPs (exact per-unit counts from
tests/golden/flop_constants.json times the work counters the kernel writes)
against the FP64 vector peak, over the tracer launch alone (`achieved`) and
over the streamed frame time (`streamed_achieved`); `post_roofline` prices bloom + display
against HBM bandwidth.  `cpu_baseline` times the FP64 CPU oracle on a
bounded row sample of the same frame on the host cores.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_VECTOR = 78.6e12   # MI355X dense FP64 vector FLOP/s
PEAK_FP32_VECTOR = 157.3e12  # MI355X FP32 vector FLOP/s
PEAK_HBM = 8.0e12            # bytes/s
# tolerance of the converged FP64 frame the 3840x2160 frames are also
# compared with (four decades below the workload's 1e-7)
CONVERGED_TOL = 1e-11


def pmc_traffic(name="trace_fp64_hd1080_a0998_traffic.json"):
    """HBM bytes per launch (tracer) or per frame (post passes) measured with
    rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on this workload (profiles/,
    with the commands that made them), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def flop_constants():
    with open(os.path.join(ROOT, "tests", "golden", "flop_constants.json")) as fh:
        return json.load(fh)


def trace_flops(st, fc):
    return st["rhs_evals"] * fc["rhs"] + st["attempts"] * fc["attempt_overhead"] + st["accepted"] * fc["commit"]


def cpu_baseline(cfg, cam, budget_s=20.0):
    """FP64 oracle on a row sample of the frame, all host cores."""
    from oracle import oracle as O
    threads = O.host_threads()
    # probe: a sparse row sample to estimate the per-row cost
    step = max(1, cfg.height // 16)
    _, _, _, st = O.render_rows(cfg, cam, (0, cfg.height, step), threads)
    per_row = st.seconds / max(1, len(range(0, cfg.height, step)))
    rows_budget = max(1, int(budget_s / max(per_row, 1e-6)))
    step2 = max(1, -(-cfg.height // rows_budget))
    t0 = time.time()
    _, _, _, st2 = O.render_rows(cfg, cam, (0, cfg.height, step2), threads)
    nrows = len(range(0, cfg.height, step2))
    return {
        "value": st2.rays / st2.seconds / 1e6,
        "unit": "Mrays/s",
        "cores": int(st2.threads),
        "kind": "port",
        "sample": f"every {step2}th row of the frame ({nrows} rows x {cfg.width} px = {st2.rays} rays), "
                  f"FP64 oracle, {st2.threads} threads, {st2.seconds:.1f} s",
        "seconds": round(time.time() - t0, 2),
    }


def timed_traces(r, cam, hdr, steps):
    """Mean tracer-kernel time (s) over `steps` frames, from the renderer's
    HIP events around the kernel launch, and the last frame's counters."""
    ms = []
    st = None
    for _ in range(steps):
        r.trace(cam, hdr=hdr, with_class=False)
        st = r.stats()
        ms.append(st["kernel_event_ms"])
    return sum(ms) / len(ms) / 1e3, st


def pipelined_frames(r, cam, bufs, streams, steps):
    """Frames k = 0..steps-1 on streams[k % n] into bufs[k % n] (hdr, bloom,
    ldr), n = len(streams): trace then bloom + display transform, up to n
    frames in flight.  Returns the elapsed ms from the first issue to the end
    of the last frame (HIP events on the current stream, which waits for all
    of them)."""
    import torch
    cur = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(cur)
    for s in streams:
        s.wait_event(e0)
    n = len(streams)
    for k in range(steps):
        s = streams[k % n]
        hdr, bloom, ldr = bufs[k % n]
        r.trace(cam, hdr=hdr, with_class=False, stream=s)
        r.post_process(hdr, ldr=ldr, bloom=bloom, stream=s)
    for s in streams:
        e = torch.cuda.Event()
        e.record(s)
        cur.wait_event(e)
    e1.record(cur)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


STREAM_CANDIDATES = [(w, t, 8, 1, 256, 1) for w in (6, 7, 8) for t in (16, 24, 32)] + \
    [(6, 24, 8, 1, 256, 0), (7, 24, 8, 1, 256, 0), (7, 24, 8, 1, 256, 2)]


def tune_streamed(r, cam, bufs, streams, first, counts=(2, 3), frames=6):
    """Launch options and stream count for streamed frames: the frame time
    of several frames in flight favours other options than a frame alone
    (more resident waves, rays kept longer per wave).  Tries `first` and
    STREAM_CANDIDATES on each stream count in `counts`, each on `frames`
    streamed frames, and returns the fastest (options dict, stream count)."""
    keys = ("waves_per_cu", "refill_threshold", "queue_chunk", "cost_order", "long_ray_attempts", "isolate_level")
    cands = [tuple(first[k] for k in keys)] if first else []
    cands += [c for c in STREAM_CANDIDATES if c not in cands]
    best, best_n, best_ms = None, None, float("inf")
    for c in cands:
        r.set_launch_opts(*c)
        for n in counts:
            pipelined_frames(r, cam, bufs[:n], streams[:n], n)
            ms = pipelined_frames(r, cam, bufs[:n], streams[:n], frames) / frames
            if ms < best_ms:
                best, best_n, best_ms = c, n, ms
    return dict(zip(keys, best)), best_n


def mixed_leg(cfg, cam, r64, hdr64, args, fc, t64, launch_opts):
    """The same frame through the mixed-precision tracer at tolerances 1e-9
    (the workload's), 1e-8 and 1e-7: its time against the FP64 tracer at the
    same tolerance (and against the FP64 frame at 1e-9), and its error
    against the FP64 frame of the same tolerance (per-pixel hit class,
    sub-class, relative HDR error on matching non-boundary pixels)."""
    import numpy as np
    import torch

    from kerrtrace_amd import config as K
    from kerrtrace_amd.renderer import HipRenderer
    from tests.parity import compare
    out = {}
    for tol in (cfg.tolerance, 1e-8, 1e-7):
        frames, times, stats = {}, {}, {}
        for prec in (K.KT_PRECISION_FP64, K.KT_PRECISION_MIXED):
            c = cfg.copy()
            c.precision = prec
            c.tolerance = tol
            if prec == K.KT_PRECISION_FP64 and tol == cfg.tolerance:
                rr, hh, t, own = r64, hdr64, t64, False
            else:
                rr = HipRenderer(c)
                if launch_opts:
                    rr.set_launch_opts(**launch_opts)
                hh, own = rr.alloc_hdr(), True
                rr.trace(cam, hdr=hh, with_class=False)
                torch.cuda.synchronize()
                t, _ = timed_traces(rr, cam, hh, args.steps)
            cl = torch.empty((c.height, c.width), dtype=torch.uint8, device="cuda")
            rr.trace(cam, hdr=hh, cls=cl)
            torch.cuda.synchronize()
            stats[prec] = rr.stats()
            frames[prec] = (hh.cpu().numpy().astype(np.float64), cl.cpu().numpy())
            times[prec] = t
            if own:
                rr.close()
        st = stats[K.KT_PRECISION_MIXED]
        tm, tf = times[K.KT_PRECISION_MIXED], times[K.KT_PRECISION_FP64]
        m = compare(*frames[K.KT_PRECISION_MIXED], *frames[K.KT_PRECISION_FP64])
        rays = cfg.width * cfg.height * cfg.spp
        key = "same_tol" if tol == cfg.tolerance else f"tol_{tol:g}"
        out[key] = {
            "tolerance": tol,
            "fp32_rhs": tol >= cfg.mixed_min_tolerance,
            "trace_ms": round(tm * 1e3, 4),
            "trace_mrays_s": round(rays / tm / 1e6, 4),
            "fp64_same_tol_trace_ms": round(tf * 1e3, 4),
            "speedup_vs_fp64_same_tol": round(tf / tm, 4),
            "speedup_vs_fp64_tol_1e-9": round(t64 / tm, 4),
            "attempts_per_ray": round(st["attempts"] / rays, 3),
            "fp64_attempts_per_ray": round(stats[K.KT_PRECISION_FP64]["attempts"] / rays, 3),
            "max_attempts_ray": st["max_attempts_ray"],
            "algorithmic_tflops": round(trace_flops(st, fc) / tm / 1e12, 4),
            "vs_fp64_frame_same_tol": {k: m[k] for k in ("class_mismatch", "class_mismatch_frac",
                                                          "class_mismatch_non_boundary", "subclass_mismatch",
                                                          "max_rel_non_boundary", "rms_rel", "p999_rel")},
        }
    return out


def uhd_leg(args, launch_opts):
    """3840x2160, i = 60 deg, tol 1e-7 at spins 0, 0.5, 0.9 and 0.998: the
    mixed tracer against the FP64 tracer on the same frame (the FP64 frame
    is the FP64 oracle's bit for bit), each timed with HIP events around
    its launches, with launch options from the autotuner on each frame
    (`launch_opts` when given, the defaults with --no-autotune)."""
    import numpy as np
    import torch

    from kerrtrace_amd import config as K
    from kerrtrace_amd.renderer import HipRenderer
    from tests.parity import compare
    out = {"workload": "3840x2160, i=60deg, r_obs=100M, fov=20deg, tol=1e-7, FBM 6 octaves, mixed vs FP64"}
    for spin in (0.0, 0.5, 0.9, 0.998):
        cfg, cam = K.preset("uhd2160_i60", spin=spin)
        rays = cfg.width * cfg.height * cfg.spp
        res = {}
        for prec in (K.KT_PRECISION_FP64, K.KT_PRECISION_MIXED):
            c = cfg.copy()
            c.precision = prec
            with HipRenderer(c) as r:
                h = r.alloc_hdr()
                if launch_opts:
                    r.set_launch_opts(**launch_opts)
                    opts = launch_opts
                elif args.no_autotune:
                    opts = "default"
                else:
                    opts = r.autotune(cam, hdr=h)
                cls = torch.empty((c.height, c.width), dtype=torch.uint8, device="cuda")
                r.trace(cam, hdr=h, cls=cls)
                torch.cuda.synchronize()
                t, st = timed_traces(r, cam, h, args.steps)
                r.trace(cam, hdr=h, cls=cls)
                torch.cuda.synchronize()
                res[prec] = (t, st, h.cpu().numpy().astype(np.float64), cls.cpu().numpy(), opts)
        t64, st64, h64, c64, o64 = res[K.KT_PRECISION_FP64]
        tm, stm, hm, cm, om = res[K.KT_PRECISION_MIXED]
        m = compare(hm, cm, h64, c64)
        # both frames against the converged frame (FP64 at CONVERGED_TOL):
        # the FP64 tracer's own error at tolerance 1e-7 is the scale the
        # mixed frame's difference from it is measured against
        c = cfg.copy()
        c.precision = K.KT_PRECISION_FP64
        c.tolerance = CONVERGED_TOL
        with HipRenderer(c) as r:
            hc, cc = r.trace(cam)
            torch.cuda.synchronize()
            hc, cc = hc.cpu().numpy().astype(np.float64), cc.cpu().numpy()
        keys = ("class_mismatch", "class_mismatch_non_boundary", "subclass_mismatch_non_boundary",
                "max_rel_non_boundary", "p999_rel", "rms_rel")
        m64c, mmc = compare(h64, c64, hc, cc), compare(hm, cm, hc, cc)
        out[f"spin_{spin:g}"] = {
            "fp64_trace_ms": round(t64 * 1e3, 4), "fp64_mrays_s": round(rays / t64 / 1e6, 4),
            "mixed_trace_ms": round(tm * 1e3, 4), "mixed_mrays_s": round(rays / tm / 1e6, 4),
            "speedup": round(t64 / tm, 4),
            "launch_opts": {"fp64": o64, "mixed": om},
            "attempts_per_ray": {"fp64": round(st64["attempts"] / rays, 3), "mixed": round(stm["attempts"] / rays, 3)},
            "vs_fp64_frame": {k: m[k] for k in ("class_mismatch", "class_mismatch_frac", "class_mismatch_non_boundary",
                                                 "subclass_mismatch", "max_rel", "max_rel_non_boundary", "rms_rel")},
            "vs_converged": {"tolerance": CONVERGED_TOL, "fp64": {k: m64c[k] for k in keys},
                             "mixed": {k: mmc[k] for k in keys}},
        }
        print(f"[bench] 4K spin {spin}: fp64 {t64 * 1e3:.2f} ms, mixed {tm * 1e3:.2f} ms, "
              f"class mismatches {m['class_mismatch']}; max rel. error vs the converged frame: "
              f"fp64 {m64c['max_rel_non_boundary']:.2e}, mixed {mmc['max_rel_non_boundary']:.2e}",
              file=sys.stderr, flush=True)
    return out


def batch_leg(args, precision, tol):
    """The 64-render seed-7 batch (kerrtrace_amd/presets/batch64_seed7.json)
    at one precision and tolerance: launch options from the autotuner, then
    one persistent launch over every render's tiles vs 64 launches of one
    render each (same options), both timed with HIP events around the
    tracer launches."""
    import torch

    from kerrtrace_amd import config as K
    from kerrtrace_amd.renderer import HipBatch, HipRenderer
    renders = K.batch_preset()
    for c, _ in renders:
        c.precision = precision
        c.tolerance = tol
    rays = sum(c.width * c.height * c.spp for c, _ in renders)
    b = HipBatch(renders)
    hdr, _ = b.alloc(with_class=False)
    opts = b.autotune(hdr=hdr)
    torch.cuda.synchronize()
    t_batch = []
    for _ in range(args.steps):
        b.trace(hdr=hdr)
        st = b.stats()
        t_batch.append(st["kernel_event_ms"])
    singles = [HipRenderer(c) for c, _ in renders]
    for r, (_, cam), h in zip(singles, renders, hdr):
        r.set_launch_opts(**opts)
        r.trace(cam, hdr=h, with_class=False)
    torch.cuda.synchronize()
    t_seq = []
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for r, (_, cam), h in zip(singles, renders, hdr):
            r.trace(cam, hdr=h, with_class=False)
        e1.record()
        torch.cuda.synchronize()
        t_seq.append(e0.elapsed_time(e1))
    for r in singles:
        r.close()
    b.close()
    tb = sum(t_batch) / len(t_batch)
    ts = sum(t_seq) / len(t_seq)
    return {
        "precision": "mixed" if precision == K.KT_PRECISION_MIXED else "fp64",
        "tolerance": tol,
        "renders": len(renders),
        "rays": rays,
        "launch_opts": opts,
        "batched_ms": round(tb, 4),
        "batched_mrays_s": round(rays / tb / 1e3, 4),
        "sequential_ms": round(ts, 4),
        "sequential_mrays_s": round(rays / ts / 1e3, 4),
        "speedup": round(ts / tb, 4),
        "attempts_per_ray": round(st["attempts"] / rays, 3),
        "max_attempts_ray": st["max_attempts_ray"],
        "tail_ms": round(st["tail_us"] / 1e3, 4),
    }


def grad_leg(args):
    """The differentiable 512x512 scene: recording forward pass and adjoint
    pass at the fit's start point against the target frame (L2 loss on
    linear HDR), their HIP-event times and ratio, the plain FP64 trace of
    the same frame, then an Adam fit from the start point."""
    import torch

    from kerrtrace_amd import config as K
    from kerrtrace_amd.grad import GradRenderer, fit_adam
    from kerrtrace_amd.renderer import HipRenderer
    cfg, _ = K.preset("grad512")
    out = {"workload": "512x512 observer scene, FP64, tol 1e-7, target (a 0.7, i 70deg, r 60M, mdot 1.0), "
                       "start (0.5, 55deg, 70M, 0.8), checkpoint every 32 steps"}
    with GradRenderer(cfg, checkpoint_interval=32) as g:
        target, _ = g.forward(K.GRAD_TARGET)
        target = target.clone()
        fwd, bwd = [], []
        for it in range(args.warmup + args.steps):
            hdr, _ = g.forward(K.GRAD_START)
            dl = 2.0 * (hdr - target)
            dl[..., 3] = 0.0
            g.backward(dl)
            st = g.stats()
            if it >= args.warmup:
                fwd.append(st["forward_ms"])
                bwd.append(st["backward_ms"])
        tf, tb = sum(fwd) / len(fwd), sum(bwd) / len(bwd)
        out.update({"forward_ms": round(tf, 4), "backward_ms": round(tb, 4),
                    "backward_over_forward": round(tb / tf, 4), "log_chunks": st["chunks"],
                    "forward_attempts": st["attempts"]})
        # the same frame through the plain (non-recording) FP64 tracer
        c2 = cfg.copy()
        c2.spin, c2.mdot_scale, c2.escape_radius = K.GRAD_START[0], K.GRAD_START[3], K.GRAD_START[2]
        cam = K.observer_camera(K.GRAD_START[2], math.degrees(K.GRAD_START[1]), K.GRAD_FOV_DEG)
        with HipRenderer(c2) as r:
            h2 = r.alloc_hdr()
            r.trace(cam, hdr=h2, with_class=False)
            plain = []
            for _ in range(args.steps):
                r.trace(cam, hdr=h2, with_class=False)
                plain.append(r.stats()["kernel_event_ms"])
        out["plain_forward_ms"] = round(sum(plain) / len(plain), 4)
        from kerrtrace_amd.grad import autotune_checkpoint
        best_k, times = autotune_checkpoint(cfg, K.GRAD_START)
        out["checkpoint_autotune"] = {"best_interval": best_k,
                                      "backward_ms": {str(k): round(v, 4) for k, v in times.items()},
                                      "backward_over_forward": round(times[best_k] / tf, 4)}
        out["backward_over_plain_forward"] = round(tb / out["plain_forward_ms"], 4)
        # the other precision pairs: FP64 forward + FP32 reversal, and the
        # workload's stated mixed forward + FP32 adjoint (FP64 crossing and
        # ray-start derivatives in both)
        from kerrtrace_amd._abi import KT_ADJOINT_FP32
        out["precision_modes"] = {}
        for name, prec in (("fp64_forward_fp32_adjoint", K.KT_PRECISION_FP64),
                           ("mixed_forward_fp32_adjoint", K.KT_PRECISION_MIXED)):
            cm = cfg.copy()
            cm.precision = prec
            with GradRenderer(cm, checkpoint_interval=32, adjoint_precision=KT_ADJOINT_FP32) as gm:
                fw, bw = [], []
                for it in range(args.warmup + args.steps):
                    hdr, _ = gm.forward(K.GRAD_START)
                    dl = 2.0 * (hdr - target)
                    dl[..., 3] = 0.0
                    gm.backward(dl)
                    stm = gm.stats()
                    if it >= args.warmup:
                        fw.append(stm["forward_ms"])
                        bw.append(stm["backward_ms"])
                tfm, tbm = sum(fw) / len(fw), sum(bw) / len(bw)
                out["precision_modes"][name] = {"forward_ms": round(tfm, 4), "backward_ms": round(tbm, 4),
                                                "backward_over_forward": round(tbm / tfm, 4)}
        if args.adam_iters > 0:
            def adam_fit(renderer):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                theta, hist = fit_adam(renderer, target, K.GRAD_START, iters=args.adam_iters)
                torch.cuda.synchronize()
                wall = time.perf_counter() - t0
                losses = [h[0] for h in hist]
                t = K.GRAD_TARGET
                return {"iters": args.adam_iters, "seconds": round(wall, 3),
                        "schedule": "per-coordinate steps (0.01, 0.01 rad, 0.5 M, 0.01); lr 1 for 60% of the "
                                    "iterations, then linear to 0.1; coarse to fine: L2 of the frames blurred by a "
                                    "Gaussian of sigma 8 px falling linearly to 0 (plain L2 at the end); gradient "
                                    "components clipped to 3x their median over the last 10 iterations; Adam "
                                    "betas (0.9, 0.99); best iterate by the plain L2 loss",
                        "loss_first": losses[0], "loss_last": losses[-1], "loss_min": min(losses),
                        "theta_best": [round(v, 6) for v in theta],
                        "error_best": {"spin": round(abs(theta[0] - t[0]), 6),
                                       "inclination_deg": round(math.degrees(abs(theta[1] - t[1])), 6),
                                       "r_obs_M": round(abs(theta[2] - t[2]), 6),
                                       "mdot": round(abs(theta[3] - t[3]), 6)},
                        "theta_last": [round(v, 6) for v in hist[-1][1]],
                        "theta_start": [round(v, 6) for v in K.GRAD_START],
                        "theta_target": [round(v, 6) for v in t],
                        "loss_every_20": [float(f"{v:.6e}") for v in losses[::20]]}
            out["adam"] = adam_fit(g)
            # the workload's stated precision: mixed forward, FP32 adjoint
            cm = cfg.copy()
            cm.precision = K.KT_PRECISION_MIXED
            with GradRenderer(cm, checkpoint_interval=32, adjoint_precision=KT_ADJOINT_FP32) as gm:
                out["adam_mixed_forward_fp32_adjoint"] = adam_fit(gm)
    return out


def copy_calibration(mib, reps=3):
    """A device-to-device copy of known size: under rocprofv3 --pmc
    FETCH_SIZE (then WRITE_SIZE) each copy dispatch should report mib * 1024
    KiB; the ratio is the scale applied to the tracer's and the post
    passes' counts."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[bench] copied {mib} MiB x {reps} ({mib * 1024} KiB read and written per copy)", flush=True)


TRACE_PRESETS = ("cpu256_a09", "hd1080_a0998", "uhd2160_i60", "grad512")


def parse_trace_spec(spec):
    """--trace-only's workload, preset[:tolerance[:fp64|mixed[:spin]]], e.g.
    hd1080_a0998:1e-7:mixed or uhd2160_i60:1e-7:fp64:0.998, as a dict
    (name, tolerance or None, precision or None, spin or None); exits with
    a message on a malformed spec."""
    parts = spec.split(":")
    if not 1 <= len(parts) <= 4 or parts[0] not in TRACE_PRESETS:
        sys.exit(f"--trace-only takes preset[:tolerance[:fp64|mixed[:spin]]] with a preset of {TRACE_PRESETS}")
    out = {"name": parts[0], "tolerance": None, "precision": None, "spin": None}
    try:
        if len(parts) > 1 and parts[1]:
            out["tolerance"] = float(parts[1])
        if len(parts) > 3:
            out["spin"] = float(parts[3])
    except ValueError:
        sys.exit("--trace-only: the tolerance and the spin are numbers")
    if out["tolerance"] is not None and not out["tolerance"] > 0.0:
        sys.exit("--trace-only: the tolerance must be positive")
    if out["spin"] is not None and not -1.0 < out["spin"] < 1.0:
        sys.exit("--trace-only: the spin must lie in (-1, 1)")
    if len(parts) > 2:
        if parts[2] not in ("fp64", "mixed"):
            sys.exit("--trace-only: the precision is fp64 or mixed")
        out["precision"] = parts[2]
    return out


def trace_only(spec, steps, launch_opts, lone=False):
    """Trace one workload (parse_trace_spec) `steps` times and print its
    work counters: the command profiles run under rocprofv3.  lone: trace
    only the frame's longest ray (lone_ray_scene)."""
    import torch

    from kerrtrace_amd import config as K
    from kerrtrace_amd.renderer import HipRenderer
    kw = {"spin": spec["spin"]} if spec["spin"] is not None else {}
    cfg, cam = K.preset(spec["name"], **kw)
    if spec["tolerance"] is not None:
        cfg.tolerance = spec["tolerance"]
    if spec["precision"] is not None:
        cfg.precision = {"fp64": K.KT_PRECISION_FP64, "mixed": K.KT_PRECISION_MIXED}[spec["precision"]]
    if lone:
        cfg, cam = lone_ray_scene(cfg, cam, launch_opts)
        launch_opts = (1, 1, 1, 2, 256, 0)
    with HipRenderer(cfg) as r:
        if launch_opts:
            r.set_launch_opts(*launch_opts)
        hdr = r.alloc_hdr()
        for k in range(steps):
            r.trace(cam, hdr=hdr, with_class=False)
            st = r.stats()
            print(f"[bench] trace {k}: {st['kernel_event_ms']:.3f} ms, {st['attempts']} attempts, "
                  f"{st['rhs_evals']} rhs, {st['accepted']} accepted, max {st['max_attempts_ray']} per ray",
                  flush=True)
        torch.cuda.synchronize()


def lone_ray_scene(cfg, cam, launch_opts):
    """A 1x1 render of the frame's longest ray: the frame is traced once
    with per-pixel attempt counts, then the camera is aimed along the exact
    jittered direction of its longest pixel with fov_scale 0, so the single
    ray repeats that pixel's arithmetic (and attempts) bit for bit, alone
    on the GPU: the frame's critical path."""
    import torch

    from kerrtrace_amd._abi import KtCamera
    from kerrtrace_amd.renderer import HipRenderer
    with HipRenderer(cfg) as r:
        if launch_opts:
            r.set_launch_opts(*launch_opts)
        att = torch.zeros((cfg.height, cfg.width), dtype=torch.int32, device="cuda")
        r.trace(cam, with_class=False, attempts=att)
        a = att.cpu()
    idx = int(a.argmax())
    y, x = divmod(idx, cfg.width)

    def pcg32(state):
        m64 = (1 << 64) - 1
        v = (state * 6364136223846793005 + 1442695040888963407) & m64
        xs = (((v >> 18) ^ v) >> 27) & 0xFFFFFFFF
        rot = v >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF
    base = (y * cfg.width + x) << 32
    rx, ry = pcg32(base) * 2.0 ** -32, pcg32(base ^ 0xda3e39cb94b95bdb) * 2.0 ** -32
    u = ((x + rx) * (2.0 / cfg.width) - 1.0) * (cfg.width / cfg.height) * cam.fov_scale
    v = (1.0 - (y + ry) * (2.0 / cfg.height)) * cam.fov_scale
    one = KtCamera()
    for k in range(3):
        one.pos[k], one.rgt[k], one.up[k] = cam.pos[k], cam.rgt[k], cam.up[k]
        one.fwd[k] = u * cam.rgt[k] + v * cam.up[k] + cam.fwd[k]
    one.fov_scale = 0.0
    c1 = cfg.copy()
    c1.width = c1.height = 1
    print(f"[bench] lone ray: pixel ({x}, {y}), {int(a[y, x])} attempts in the frame", flush=True)
    return c1, one


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-render batch leg")
    ap.add_argument("--no-uhd", action="store_true", help="skip the 3840x2160 four-spin mixed-vs-FP64 leg")
    ap.add_argument("--no-grad", action="store_true", help="skip the differentiable-render leg")
    ap.add_argument("--no-autotune", action="store_true", help="default launch options for the 1080p frame")
    ap.add_argument("--launch-opts", type=str, default=None,
                    help="fixed launch options for the 1080p frame instead of the autotuner: waves per CU, "
                         "refill threshold, tile edge, cost order, long-ray attempts[, isolate level] "
                         "(e.g. 6,32,8,1,256,1)")
    ap.add_argument("--adam-iters", type=int, default=200, help="Adam iterations of the fit (0: none)")
    ap.add_argument("--streams", type=int, default=0,
                    help="HIP streams of the streamed frames, 1 to 3 (default: tuned between 2 and 3; "
                         "2 with --launch-opts or --no-autotune)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB three times and exit: the "
                         "calibration of the FETCH_SIZE / WRITE_SIZE counters (a copy of N MiB reads "
                         "and writes N MiB)")
    ap.add_argument("--trace-only", type=str, default=None,
                    help="only trace one workload --steps times with --launch-opts and print its counters "
                         "(for profiler runs): preset[:tolerance[:fp64|mixed[:spin]]]")
    ap.add_argument("--lone-ray", action="store_true",
                    help="with --trace-only: trace only the frame's longest ray, alone (its critical path)")
    args = ap.parse_args()
    # every argument is checked before anything touches a GPU
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU: rays are independent, so the frame shards with no exchange; "
                 "the renderer traces on one GPU (use --gpus 1)")
    fixed_opts = None
    if args.launch_opts:
        try:
            fixed_opts = [int(v) for v in args.launch_opts.split(",")]
        except ValueError:
            fixed_opts = []
        if len(fixed_opts) not in (5, 6):
            sys.exit("--launch-opts takes five or six integers")
    if args.steps < 1 or args.warmup < 0:
        sys.exit("--steps must be at least 1 and --warmup at least 0")
    if not 0 <= args.streams <= 3:
        sys.exit("--streams takes 1, 2 or 3 (four streams would share a hardware queue)")
    if args.lone_ray and not args.trace_only:
        sys.exit("--lone-ray traces one workload's longest ray: it needs --trace-only")
    if args.trace_only:
        trace_only(parse_trace_spec(args.trace_only), args.steps, fixed_opts, lone=args.lone_ray)
        return
    if args.copy_mib > 0:
        copy_calibration(args.copy_mib)
        return

    import torch

    from kerrtrace_amd import config as K
    from kerrtrace_amd.renderer import HipRenderer

    torch.cuda.set_device(0)
    cfg, cam = K.preset("hd1080_a0998")
    W, H = cfg.width, cfg.height
    rays = W * H * cfg.spp
    fc = flop_constants()
    r = HipRenderer(cfg)
    hdr = r.alloc_hdr()
    bloom = torch.empty_like(hdr)
    ldr = torch.empty((H, W), dtype=torch.int32, device="cuda")
    if fixed_opts:
        vals = fixed_opts
        r.set_launch_opts(*vals)
        launch_opts = dict(zip(("waves_per_cu", "refill_threshold", "queue_chunk", "cost_order", "long_ray_attempts",
                                "isolate_level"), vals + [0] * (6 - len(vals))))
    else:
        launch_opts = None if args.no_autotune else r.autotune(cam, hdr=hdr)
    for _ in range(args.warmup):
        r.trace(cam, hdr=hdr, with_class=False)
        r.post_process(hdr, ldr=ldr, bloom=bloom)
    torch.cuda.synchronize()
    print(f"[bench] warmup done ({args.warmup} frames)", file=sys.stderr, flush=True)

    # timed frames: events around each frame's trace and post; the
    # renderer's own events bracket the tracer kernel launch alone
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    kernel_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record()
        r.trace(cam, hdr=hdr, with_class=False)
        ev[k][1].record()
        r.post_process(hdr, ldr=ldr, bloom=bloom)
        ev[k][2].record()
        st = r.stats()  # synchronises the frame
        kernel_ms.append(st["kernel_event_ms"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    frame_ms = [e[0].elapsed_time(e[2]) for e in ev]
    trace_ms = [e[0].elapsed_time(e[1]) for e in ev]
    post_ms = [e[1].elapsed_time(e[2]) for e in ev]
    latency_ms = sum(frame_ms) / args.steps
    serial_hdr, serial_ldr = hdr.clone(), ldr.clone()

    # the same frames streamed: two or three streams (a process has four
    # hardware queues, one taken by the current stream), as many frames in
    # flight, with the stream count and launch options tuned for streaming
    # (unless they were given)
    n_max = max(3, args.streams)
    streams = [torch.cuda.Stream() for _ in range(n_max)]
    bufs = [(hdr, bloom, ldr)] + [(r.alloc_hdr(), torch.empty_like(bloom), torch.empty_like(ldr))
                                  for _ in range(n_max - 1)]
    stream_opts = launch_opts
    n_streams = args.streams or 2
    if not args.launch_opts and not args.no_autotune:
        counts = (args.streams,) if args.streams else (2, 3)
        stream_opts, n_streams = tune_streamed(r, cam, bufs, streams, launch_opts, counts)
        r.set_launch_opts(**stream_opts)
    streams, bufs = streams[:n_streams], bufs[:n_streams]
    pipelined_frames(r, cam, bufs, streams, n_streams * max(1, args.warmup))
    print("[bench] pipelined warmup done", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    total_ms = pipelined_frames(r, cam, bufs, streams, args.steps)
    wall += time.perf_counter() - t1
    ms_per_step = total_ms / args.steps
    same = all(torch.equal(b[0], serial_hdr) and torch.equal(b[2], serial_ldr) for b in bufs)
    if launch_opts:
        r.set_launch_opts(**launch_opts)  # the later legs reuse the single-frame options
    t_trace = sum(trace_ms) / len(trace_ms) / 1e3
    t_kernel = sum(kernel_ms) / len(kernel_ms) / 1e3
    t_post = sum(post_ms) / len(post_ms) / 1e3
    flops = trace_flops(st, fc)
    traffic = pmc_traffic()
    post_traffic = pmc_traffic("post_hd1080_a0998_traffic.json")
    achieved = flops / t_kernel
    post_bytes = 20.0 * W * H  # compulsory: float4 HDR read + u32 write per pixel
    out = {
        "metric": "Mrays/s at 1080p a=0.998: FP64 vs FP32-mixed; backward/forward time ratio",
        "value": round(rays / (ms_per_step / 1e3) / 1e6, 4),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "1920x1080 a=0.998 i=85deg r_obs=100M fov=20deg tol=1e-9 FP64 trace + bloom(sigma 12) + ACES/sRGB",
                   "rays": rays, "parallelism": f"single GPU, frames streamed on {n_streams} HIP streams",
                   "launch_opts": launch_opts},
        "frame_latency_ms": round(latency_ms, 4),
        "serial": {"ms_per_frame": round(latency_ms, 4), "mrays_s": round(rays / (latency_ms / 1e3) / 1e6, 4)},
        "pipelined": {"streams": n_streams, "ms_per_frame": round(ms_per_step, 4), "frames_bitwise_equal_serial": same,
                      "launch_opts": stream_opts,
                      "speedup_vs_serial": round(latency_ms / ms_per_step, 4)},
        "trace_mrays_s": round(rays / t_trace / 1e6, 4),
        "trace_ms": round(t_trace * 1e3, 4),
        "trace_kernel_ms": round(t_kernel * 1e3, 4),
        "post_ms": round(t_post * 1e3, 4),
        "wall_s": round(wall, 3),
        "work": {"rhs_evals": st["rhs_evals"], "attempts": st["attempts"], "accepted": st["accepted"],
                 "crossings": st["crossings"], "max_attempts_ray": st["max_attempts_ray"],
                 "attempts_per_ray": round(st["attempts"] / rays, 3),
                 "kernel_ms": round(st["kernel_us"] / 1e3, 4), "tail_ms": round(st["tail_us"] / 1e3, 4)},
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 4), "peak": PEAK_FP64_VECTOR / 1e12,
                     "compute_unit": "FP64 VALU: the tracer has no matrix-shaped work and issues no MFMA; priced "
                                     "against MI355X's FP64 vector peak",
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_VECTOR, 4),
                     "traffic": (traffic or {}).get("bytes_per_launch"),
                     "kernel": "trace_fp64_persistent",
                     "flop_per_frame": flops,
                     "streamed_achieved": round(flops / (ms_per_step / 1e3) / 1e12, 4),
                     "streamed_frac": round(flops / (ms_per_step / 1e3) / PEAK_FP64_VECTOR, 4)},
        "post_roofline": {"bound": "hbm", "achieved": round(post_bytes / t_post / 1e9, 2), "peak": PEAK_HBM / 1e9,
                          "unit": "GB/s", "frac": round(post_bytes / t_post / PEAK_HBM, 4),
                          "traffic": (post_traffic or {}).get("bytes_per_frame"),
                          "traffic_over_compulsory": (post_traffic or {}).get("traffic_over_compulsory"),
                          "kernels": "bloom_h + post_v"},
    }
    print(f"[bench] timed {args.steps} frames: {latency_ms:.2f} ms/frame one at a time, {ms_per_step:.2f} ms/frame "
          f"streamed (bitwise equal: {same})", file=sys.stderr, flush=True)
    if not args.no_mixed:
        out["mixed"] = mixed_leg(cfg, cam, r, hdr, args, fc, t_kernel, launch_opts)
    if not args.no_uhd:
        print("[bench] 4K four-spin leg", file=sys.stderr, flush=True)
        out["uhd2160_i60"] = uhd_leg(args, launch_opts if args.launch_opts else None)
    if not args.no_batch:
        print("[bench] batch leg", file=sys.stderr, flush=True)
        out["batch64"] = {"fp64_tol_1e-9": batch_leg(args, K.KT_PRECISION_FP64, 1e-9),
                          "mixed_tol_1e-7": batch_leg(args, K.KT_PRECISION_MIXED, 1e-7)}
    if not args.no_grad:
        print("[bench] gradient leg", file=sys.stderr, flush=True)
        out["grad512"] = grad_leg(args)
        out["backward_forward_ratio"] = out["grad512"]["backward_over_forward"]

    if not args.no_cpu:
        print("[bench] cpu baseline leg", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cfg, cam, args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    r.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
