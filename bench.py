"""GI benchmarks on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 1080p]
                    [--form fp32|mixed|both] [--slots 3] [--no-cpu]

Default (workload 1080p): a step is one GI frame of the many-lights hot path
on inputs already resident in HBM -- VPL extraction from the RSM, ISM atlas
clear + point splat, pull-push, interleaved final gathering -- at 1920x1080,
4,096 VPLs (128^2 RSM), 1M splat points into 4,096 64^2 ISMs of a 4096^2
atlas, 4 pull-push levels, 8x8 interleave, 40k-triangle procedural room.
Four forms run on the same inputs:
  fp32          the plain FP32 form through the single-frame entry point
                (mlgi_frame, GIStage::process's path)
  mixed_single  the mixed-precision form through the same entry point
  fp32_batched  the FP32 form through the batched kernels
                (mlgi_batch_run of the one frame)
  mixed         the mixed-precision form through the batched kernels
The timed region is exactly K frames of a form bracketed by a device
synchronize on both sides, with no timing events inside; ms_per_step is
wall time / K. A second timed region of the same K frames then records HIP
events on the context stream around every stage of every frame (the
per-stage spans and the kernels' rooflines come from it; its wall time per
frame is reported as profiled_ms_per_frame). Batched forms are also
replayed from a captured HIP graph (graph_ms_per_frame). Then, with
--slots S > 1 (default 3), the batched forms run K frames once more with S
frames in flight (fp32_pipelined, mixed_pipelined): S frame slots, each
with its own buffers and its own context and stream, frame i on slot
i mod S, so that one frame's last workgroups of a stage share the GPU with
the next frame's work; every frame does the whole work. `value` is
mixed_pipelined (config.form, config.frames_in_flight); K frames in flight
take about (K - 1) steady-state frames plus one frame's latency, so the
default K = 200 keeps that fill and drain to a few percent of the figure
(K = 50 reports ~5% more per frame). serial_ms_per_frame
is the mixed form one frame after another on one stream, through the faster
of its two paths (both drop-in entry points for one frame). Precision is
compared within a path: mixed_speedup_vs_fp32 = fp32 / mixed_single,
mixed_speedup_same_path = fp32_batched / mixed (and its graph replays),
mixed_speedup_pipelined = fp32_pipelined / mixed_pipelined.
The default 1080p run then runs the other configurations briefly and adds
their headline figures as other_configs (the 4K stress frames, the 64-frame
batch, a 10 s served trace; --no-others leaves them out).

roofline: for the kernel with the largest mean duration in the timed frames:
its VALU issue rate ("valu": SQ_INSTS_VALU per launch from the committed
rocprofv3 counters, profiles/sq_counters_<workload>.json, over the live kernel
time, against the issue rate of an FMA probe kernel measured in the same run)
and
pull-push against HBM (6 bytes per atlas texel: the u32 atlas read once, the
unorm16 ISM written once), the final gather against the FP32 vector peak (84
flops per lit pixel-VPL pair, pairs counted on the device), the splat (mask,
tile and spill kernels together) against HBM (16 bytes per point record read
plus 4 bytes per atlas texel written). Every kernel's figure is under
"kernels"; "traffic" is the HBM bytes per launch measured with rocprofv3
counters (profiles/traffic_<workload>.json), null until measured.
cpu_baseline: the oracle's FP32 instantiation of the same four stages on the
host for one full frame, splat, pull-push and final gathering on
--cpu-threads host threads (default: OMP_NUM_THREADS, else the CPUs this
process may run on). The reference shaders under Mesa llvmpipe are not
measured (no OpenGL context can be created in this image); the line says so.

Other workloads (one JSON line each, not the default):
  small     512x512, 1,024 VPLs, 256k points, 2048^2 atlas, 4x4 interleave
  4k        3840x2160, 16,384 VPLs, 8M points, 16,384 32^2 ISMs, 5 levels,
            160k triangles, mixed batched form, light pose changed every
            frame (device RSM) and the frame replayed from a HIP graph
  batch64   64 independent frames of the 40k scene (resolutions 128^2..1024^2,
            256..4,096 VPLs, random light poses), autotuned, one batch
            replayed from a HIP graph; ms per batch and per frame
  serve     the request server over 3 resident scenes (4k/40k/160k triangles),
            Poisson arrivals at --rate req/s for --seconds, HBM-budget
            admission; req/s and p50/p99 latency
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "many-lights-gi_amd"))

FLOPS_PER_LIT_PAIR = 84
FP32_VECTOR_PEAK_TFLOPS = 157.3
# 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 VALU instruction
VALU_PEAK_GINST_S = 614.4
HBM_PEAK_GBS = 8000.0
# the served workload's stated p99 latency budget (arrival to result), at the
# configured 200 requests/s
SERVE_P99_BUDGET_MS = 5.0
DATA = "synthetic: seeded procedural room, ray-cast G-buffer/RSM, area-proportional points"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def upload_frame(ctx, inp):
    import mlgi

    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    return rsm, pts, gb


def run_form(ctx, wl, inp, form: str, steps: int, warmup: int):
    """fp32 / mixed_single: single-frame path in FP32 / mixed precision;
    fp32_batched / mixed: batched path (batch of one)."""
    import mlgi

    prec = mlgi.PRECISION_MIXED if form in ("mixed", "mixed_single") else mlgi.PRECISION_FP32
    st = mlgi.GIStage(ctx, wl.params(precision=prec), wl.n_vpl, wl.width, wl.height)
    rsm, pts, gb = upload_frame(ctx, inp)
    batch = mlgi.Batch(ctx, [(st, rsm, pts, gb)]) if form in ("fp32_batched", "mixed") else None
    step = batch.run if batch else (lambda: st.process(rsm, pts, gb))
    for _ in range(warmup):
        step()
    ctx.sync()
    # the headline region: K frames, no events on the stream
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    # the same K frames with HIP events around every stage (stage spans and
    # kernel rooflines)
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    prof_ms = (time.perf_counter() - t0) * 1e3 / steps
    stages = ctx.profile_read()
    ctx.profile(False)
    out = {"st": st, "wall_ms": wall_ms, "profiled_ms": prof_ms, "stages": stages, "lit": st.count_lit_pairs(gb),
           "path": "batched (mlgi_batch_run)" if batch else "single frame (mlgi_frame)"}
    if batch:
        g = batch.capture()
        for _ in range(warmup):
            g.launch()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.launch()
        ctx.sync()
        out["graph_ms"] = (time.perf_counter() - t0) * 1e3 / steps
        g.close()
        batch.close()
    return out


def run_pipelined(ctx, wl, inp, prec, steps: int, warmup: int, slots: int):
    """K frames through the batched kernels with `slots` frames in flight:
    each slot has its own frame buffers (GIStage, uploaded inputs) and its
    own context, hence its own stream, and frame i runs on slot i mod slots.
    Every frame does the whole work; a frame's slot is free again once the
    frame `slots` earlier on it has finished (stream order), and frames on
    different slots overlap on the device, so one frame's last workgroups
    of a stage share the GPU with the next frame's work. Returns wall ms per
    frame."""
    import mlgi

    ctxs = [ctx] + [mlgi.Context(ctx.device) for _ in range(slots - 1)]
    batches = []
    try:
        for c in ctxs:
            st = mlgi.GIStage(c, wl.params(precision=prec), wl.n_vpl, wl.width, wl.height)
            rsm, pts, gb = upload_frame(c, inp)
            batches.append(mlgi.Batch(c, [(st, rsm, pts, gb)]))
        # untimed: every slot's first frames (at least 8 per slot)
        for i in range(max(warmup, 8 * slots)):
            batches[i % slots].run()
        for c in ctxs:
            c.sync()
        t0 = time.perf_counter()
        for i in range(steps):
            batches[i % slots].run()
        for c in ctxs:
            c.sync()
        return (time.perf_counter() - t0) * 1e3 / steps
    finally:
        for b in batches:
            b.close()
        for c in ctxs[1:]:
            c.close()


def valu_roof(kernel_ms, sq, peak_fma_s):
    """VALU issue roofline of a kernel: VALU wave-instructions per launch
    (rocprofv3 SQ_INSTS_VALU, committed profile) over the live kernel time,
    against the measured FP32 FMA issue rate (one wave64 instruction = 64
    lane FMAs), and against the nominal FMA issue rate (spec_peak: one wave64
    FMA per SIMD every 4 cycles at 2.4 GHz), which the probe reaches to
    within a few percent; instruction mixes with compares and selects can
    issue faster than either."""
    if not sq or kernel_ms <= 0:
        return None
    peak = peak_fma_s / 64.0 / 1e9 if peak_fma_s else VALU_PEAK_GINST_S
    ach = sq["SQ_INSTS_VALU"] / (kernel_ms * 1e-3) / 1e9
    return {"insts_per_launch": sq["SQ_INSTS_VALU"], "achieved": round(ach, 1), "peak": round(peak, 1),
            "unit": "G wave64-inst/s", "frac": round(ach / peak, 4),
            "peak_source": "measured FMA probe" if peak_fma_s else "spec",
            "spec_peak": VALU_PEAK_GINST_S, "spec_frac": round(ach / VALU_PEAK_GINST_S, 4)}


def kernel_rooflines(wl, stages, lit, sq=None, peak_fma_s=None, single=False, splats=None, spills=None):
    """Per-kernel rooflines of one form's timed frames; `single` selects the
    single-frame kernels' counters (profile keys with a _single suffix);
    splats / spills: the frame's LDS and global atomic mins of the splat."""
    sq = sq or {}
    sfx = "_single" if single else ""
    a2 = wl.atlas_side * wl.atlas_side
    out = {}
    t = stages["pullpush_ms"] * 1e-3
    if t > 0:
        gbs = 6.0 * a2 / t / 1e9
        out["pullpush"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": 6 * a2,
                           "kernel_ms": round(stages["pullpush_ms"], 4),
                           "valu": valu_roof(stages["pullpush_ms"], sq.get("pullpush" + sfx), peak_fma_s)}
    t = stages["ism_splat_ms"] * 1e-3
    if t > 0:
        # ISM-major splat: point records read once, the atlas written once
        # (that write is the clear); the 2-byte masks stay in L2/MALL
        b = 16 * wl.n_points + 4 * a2
        gbs = b / t / 1e9
        out["ism_splat"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": b,
                            "kernel_ms": round(stages["ism_splat_ms"], 4)}
        if splats is not None:
            # the tile pass's LDS atomic mins (one per accepted (point, ISM)
            # splat) over the whole splat span (mask + tiles + spill)
            out["ism_splat"]["lds_atomic_min_per_launch"] = splats
            out["ism_splat"]["lds_atomics_per_s"] = round(splats / t, 1)
            out["ism_splat"]["global_atomic_min_per_launch"] = spills
    t = stages["final_gather_ms"] * 1e-3
    if t > 0:
        tf = FLOPS_PER_LIT_PAIR * lit / t / 1e12
        out["final_gather"] = {"bound": "valu", "achieved": round(tf, 3), "peak": FP32_VECTOR_PEAK_TFLOPS,
                               "unit": "TFLOP/s", "frac": round(tf / FP32_VECTOR_PEAK_TFLOPS, 4),
                               "lit_pairs": lit, "flops_per_lit_pair": FLOPS_PER_LIT_PAIR,
                               "kernel_ms": round(stages["final_gather_ms"], 4),
                               "valu": valu_roof(stages["final_gather_ms"], sq.get("final_gather" + sfx), peak_fma_s)}
    return out


def host_threads() -> int:
    """Host threads for the CPU baseline: OMP_NUM_THREADS when set (the
    process's share of the host), else the CPUs this process may run on."""
    n = os.environ.get("OMP_NUM_THREADS", "")
    if n.isdigit() and int(n) > 0:
        return int(n)
    return len(os.sched_getaffinity(0))


LLVMPIPE = ("not measured: the reference shaders need an OpenGL 4.3 context, and this image has Mesa's "
            "llvmpipe libraries but no X server, EGL or OSMesa to create one")


def cpu_baseline(wl, inp, threads: int):
    import numpy as np
    import mlgi
    import oracle

    sh = mlgi.shuffle_indices(wl.n_vpl)
    p = wl.params()
    t = {}
    t0 = time.perf_counter()
    pos, nrm, pk, col = oracle.vpl_extract(inp.rsm_diffuse, inp.rsm_normal, inp.rsm_depth, inp.inv_bvp,
                                           p.light_intensity, sh, wl.n_vpl, np.float32)
    t["vpl"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    atlas = oracle.ism_splat(inp.points, inp.bin_offsets, pos, nrm, wl.atlas_side, p.z_far, threads=threads)
    t["splat"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    r16 = oracle.pullpush(atlas, wl.ism_side, wl.levels, p.z_far, threads=threads)
    t["pullpush"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.final_gather(inp.gb_depth, inp.gb_normal, inp.vp_inv, pos, nrm, col, r16, wl.atlas_side, p.z_far,
                        p.gi_intensity_factor, p.vpl_clamping_value, wl.interleave, threads=threads,
                        dtype=np.float32)
    t["gather"] = time.perf_counter() - t0
    total_ms = sum(t.values()) * 1e3
    return {"value": round(total_ms, 3), "unit": "ms/frame", "cores": threads, "kind": "port",
            "sample": (f"one full {wl.name} frame: oracle FP32 instantiation; splat, pull-push and final "
                       f"gathering on {threads} host threads, VPL extraction on 1"),
            "stages_s": {k: round(v, 4) for k, v in t.items()}, "llvmpipe": LLVMPIPE}


def config_of(wl, form):
    return {"workload": wl.name, "width": wl.width, "height": wl.height, "vpls": wl.n_vpl,
            "interleave": wl.interleave, "points": wl.n_points, "atlas": wl.atlas_side, "ism": wl.ism_side,
            "pull_push_levels": wl.levels, "tile": f"{wl.tile[0]}x{wl.tile[1]}", "form": form,
            "triangles": wl.n_random + 12}


def frame_bench(a, ctx, wl):
    from mlgi import workloads as W

    log(f"[bench] building inputs for {wl.name} ...")
    t0 = time.perf_counter()
    inp = W.build_inputs(wl)
    log(f"[bench] inputs ready in {time.perf_counter() - t0:.1f}s")
    forms = {"fp32": ["fp32"], "mixed": ["mixed"], "both": ["fp32", "mixed_single", "fp32_batched", "mixed"]}[a.form]
    res = {}
    for form in forms:
        log(f"[bench] form {form}: warmup {a.warmup}, steps {a.steps}")
        res[form] = run_form(ctx, wl, inp, form, a.steps, a.warmup)
        log(f"[bench] {form}: {res[form]['wall_ms']:.4f} ms/frame, stages {res[form]['stages']}")
    # frames in flight: the batched forms again with a.slots frame slots
    piped = {}
    if a.slots > 1:
        import mlgi

        for form, prec in (("fp32_pipelined", mlgi.PRECISION_FP32), ("mixed_pipelined", mlgi.PRECISION_MIXED)):
            if form.split("_")[0] in [f.split("_")[0] for f in forms]:
                log(f"[bench] form {form}: {a.slots} frame slots, warmup {a.warmup}, steps {a.steps}")
                piped[form] = run_pipelined(ctx, wl, inp, prec, a.steps, a.warmup, a.slots)
                log(f"[bench] {form}: {piped[form]:.4f} ms/frame")
    mixed_forms = [f for f in ("mixed_single", "mixed") if f in res]
    head = min(mixed_forms, key=lambda f: res[f]["wall_ms"]) if mixed_forms else "fp32"
    h = res[head]
    single = head in ("mixed_single", "fp32")
    # value: the mixed form's frames in flight when measured, else its
    # faster one-frame-at-a-time path
    vform = "mixed_pipelined" if "mixed_pipelined" in piped else head
    if "mixed" not in forms and "mixed_single" not in forms:
        vform = "fp32_pipelined" if "fp32_pipelined" in piped else head
    value_ms = piped[vform] if vform in piped else h["wall_ms"]
    peak_fma_s = ctx.probe_valu()
    log(f"[bench] FP32 FMA probe: {peak_fma_s * 2 / 1e12:.1f} TFLOP/s")
    sq = None
    sq_path = ROOT / "profiles" / f"sq_counters_{wl.name}.json"
    if sq_path.exists():
        sq = {k: v for k, v in json.loads(sq_path.read_text()).items() if not k.startswith("_")}
    # the splat's atomics: the last eager single-frame splat of these inputs
    # (the single-frame forms ran it through mlgi_ism_splat on this context)
    splats = spills = None
    eager = [f for f in ("mixed_single", "fp32") if f in res]
    if eager:
        ism = res[eager[0]]["st"].ism
        splats = ism.splat_count()
        spills = ism.spill_stats()[0]
        log(f"[bench] splat: {splats} LDS atomic mins, {spills} spilled to global atomics")
    kernels = kernel_rooflines(wl, h["stages"], h["lit"], sq, peak_fma_s, single, splats, spills)
    dominant = max(kernels, key=lambda k: kernels[k]["kernel_ms"])
    roof = dict(kernels[dominant])
    roof["kernel"] = dominant
    roof["traffic"] = None
    tpath = ROOT / "profiles" / f"traffic_{wl.name}.json"
    if tpath.exists():
        tr = json.loads(tpath.read_text()).get(dominant + ("_single" if single else ""))
        if tr:
            roof["traffic"] = tr["bytes_per_launch"]
    forms_out = {}
    for k, v in res.items():
        forms_out[k] = {"ms_per_frame": round(v["wall_ms"], 4), "path": v["path"],
                        "profiled_ms_per_frame": round(v["profiled_ms"], 4),
                        "stages_ms": {s: round(x, 4) for s, x in v["stages"].items() if s != "frames"}}
        if "graph_ms" in v:
            forms_out[k]["graph_ms_per_frame"] = round(v["graph_ms"], 4)
    for k, v in piped.items():
        forms_out[k] = {"ms_per_frame": round(v, 4),
                        "path": f"batched (mlgi_batch_run), {a.slots} frames in flight (one frame slot and "
                                f"stream each)"}
    out = {
        "metric": "GI ms/frame 1080p·4096 VPL (mixed vs FP32 HIP)",
        "value": round(value_ms, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(value_ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16x2+fp32" if vform.startswith("mixed") else "fp32",
        "data": DATA,
        "config": dict(config_of(wl, vform), frames_in_flight=a.slots if vform in piped else 1),
        # one frame after another on one stream: the mixed form's faster path
        "serial_ms_per_frame": round(h["wall_ms"], 4),
        "forms": forms_out,
        "roofline": roof,
        "kernels": kernels,
    }
    out["fp32_fma_probe_tflops"] = round(peak_fma_s * 2 / 1e12, 2)
    if "fp32" in res and "mixed_single" in res:
        # precision on the single-frame path
        out["mixed_speedup_vs_fp32"] = round(res["fp32"]["wall_ms"] / res["mixed_single"]["wall_ms"], 4)
    if "fp32_batched" in res and "mixed" in res:
        # precision alone: both through the batched kernels
        out["mixed_speedup_same_path"] = round(res["fp32_batched"]["wall_ms"] / res["mixed"]["wall_ms"], 4)
        g32, gmx = res["fp32_batched"].get("graph_ms"), res["mixed"].get("graph_ms")
        if g32 and gmx:
            out["mixed_speedup_same_path_graph"] = round(g32 / gmx, 4)
    if "fp32_pipelined" in piped and "mixed_pipelined" in piped:
        out["mixed_speedup_pipelined"] = round(piped["fp32_pipelined"] / piped["mixed_pipelined"], 4)
    if not a.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(wl, inp, a.cpu_threads or host_threads())
    return out


def stress_bench(a, ctx, wl):
    """4K stress: the light pose changes every frame; the frame's RSM is
    rendered on the device and the mixed batched frame replays from a graph.
    `value`: the RSM of frame i + 1 renders on a second stream beside frame
    i's GI (mlgi.RsmPipeline, double-buffered RSM); serial_ms_per_frame: the
    render and the frame one after the other on one stream."""
    import numpy as np
    import mlgi
    from mlgi import workloads as W

    log(f"[bench] building inputs for {wl.name} ...")
    inp = W.build_inputs(wl)
    scene = W._scene(wl.seed, wl.n_random)
    dev = scene.upload(ctx)
    st = mlgi.GIStage(ctx, wl.params(precision=mlgi.PRECISION_MIXED), wl.n_vpl, wl.width, wl.height)
    _, pts, gb = upload_frame(ctx, inp)
    r = wl.rsm
    dif, nrm, dep = ctx.empty((r, r, 3), np.float32), ctx.empty((r, r, 3), np.float32), ctx.empty((r, r), np.float32)
    rsm = mlgi.rsm_desc(dif, nrm, dep, inp.inv_bvp)
    batch = mlgi.Batch(ctx, [(st, rsm, pts, gb)])
    g = batch.capture()

    def frame(i):
        inv = dev.render_rsm(mlgi.Camera.light(1000 + i), dif, nrm, dep)
        batch.update_light(0, inv)
        g.launch()

    for i in range(a.warmup):
        frame(i)
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(a.steps):
        frame(a.warmup + i)
    ctx.sync()
    serial_ms = (time.perf_counter() - t0) * 1e3 / a.steps
    g.close()
    batch.close()
    dev.close()
    log(f"[bench] 4k serial: {serial_ms:.4f} ms/frame")

    pipe = mlgi.RsmPipeline(ctx, scene, st, pts, gb, r, lambda i: mlgi.Camera.light(1000 + i))
    for _ in range(a.warmup):
        pipe.frame()
    pipe.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        pipe.frame()
    pipe.sync()
    ms = (time.perf_counter() - t0) * 1e3 / a.steps
    pipe.close()
    log(f"[bench] 4k pipelined: {ms:.4f} ms/frame")
    return {"metric": "GI ms/frame 4K·16384 VPL, light pose per frame (mixed batched, graph replay)",
            "value": round(ms, 4), "unit": "ms/frame", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": round(ms, 4), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp16x2+fp32", "data": DATA, "serial_ms_per_frame": round(serial_ms, 4),
            "config": dict(config_of(wl, "mixed"), light="new pose per frame, device RSM",
                           rsm="next frame's RSM rendered on a second stream beside this frame's GI")}


def batch_bench(a, ctx):
    """64 frames of the 40k scene with random shapes and light poses, autotuned,
    one batch replayed from a HIP graph (mlgi.workloads.Batch64)."""
    import mlgi
    from mlgi import workloads as W

    log("[bench] building the 64-frame batch (device G-buffers and RSMs, autotuning) ...")
    b64 = W.Batch64(ctx)
    batch = mlgi.Batch(ctx, b64.frames)
    g = batch.capture()
    for _ in range(a.warmup):
        g.launch()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        g.launch()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / a.steps
    tuned = {}
    for sp, tune in zip(b64.specs, b64.tune):
        tuned[f"{sp.size}x{sp.size}/{sp.n_vpl}"] = {"interleave": tune["interleave"], "tile": "%dx%d" % tuple(tune["tile"])}
    g.close()
    batch.close()
    b64.close()
    return {"metric": "GI ms per batch of 64 mixed-shape frames (mixed, batched kernels, graph replay)",
            "value": round(ms, 4), "unit": "ms/batch", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": round(ms, 4), "ms_per_frame": round(ms / 64, 4), "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp16x2+fp32", "data": DATA,
            "config": {"workload": "batch64", "frames": 64, "resolutions": list(W.BATCH64_SIZES),
                       "vpls": list(W.BATCH64_VPLS), "triangles": 40_012, "points": W.Batch64.POINTS,
                       "autotuned": tuned}}


def serve_bench(a, ctx, sweep_seconds: float = 2.0):
    import numpy as np
    import mlgi
    from mlgi import workloads as W

    scenes = [(W._scene(1, 4_000), 256_000, 21), (W._scene(1, 40_000), 1_000_000, 22),
              (W._scene(1, 160_000), 4_000_000, 23)]
    sizes = [128, 256, 512, 1024]
    vpls = [256, 1024, 2048, 4096]
    log("[bench] server create (resident scenes, point binnings, G-buffers, tuning) ...")
    t0 = time.perf_counter()
    srv = mlgi.Server(ctx, scenes, vpls, [(s, s) for s in sizes], hbm_budget=a.budget_gib << 30,
                      max_inflight=64, max_batch=32, max_queued_batches=3, streams=3)
    log(f"[bench] server ready in {time.perf_counter() - t0:.1f}s")

    def trace(rate, seconds, seed):
        rng = np.random.default_rng(seed)
        t, reqs = 0.0, []
        while True:
            t += rng.exponential(1.0 / rate)
            if t >= seconds:
                return reqs
            s = int(rng.choice(sizes))
            reqs.append((int(rng.integers(3)), s, s, int(rng.choice(vpls)), int(rng.integers(1 << 40)), t))

    reqs = trace(a.rate, a.seconds, 7)
    log(f"[bench] serving {len(reqs)} requests over {a.seconds:.0f} s ...")
    st, lat = srv.run(reqs)
    # higher arrival rates (short traces): where batching and overlap engage
    sweep = []
    for rate in (500.0, 1000.0, 2000.0, 4000.0, 8000.0):
        r = trace(rate, sweep_seconds, 100 + int(rate))
        s2, _ = srv.run(r)
        sweep.append({"arrival_rate_per_s": rate, "requests": len(r), "req_per_s": round(s2["req_per_s"], 2),
                      "latency_p50_ms": round(s2["latency_p50_ms"], 3), "latency_p99_ms": round(s2["latency_p99_ms"], 3),
                      "mean_batch": round(s2["mean_batch"], 3), "batches": s2["batches"],
                      "overlapped_batches": s2["overlapped_batches"]})
        log(f"[bench] serve sweep {sweep[-1]}")
    srv.close()
    engages = [x["arrival_rate_per_s"] for x in sweep if x["mean_batch"] > 1.05]
    return {"metric": "served GI req/s + p99 latency (3 resident scenes, Poisson arrivals, mixed form)",
            "value": round(st["req_per_s"], 3), "unit": "req/s", "n_gpus": 1, "steps": len(reqs), "warmup": 0,
            "ms_per_step": round(st["latency_mean_ms"], 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp16x2+fp32", "data": DATA,
            "latency_ms": {"p50": round(st["latency_p50_ms"], 3), "p99": round(st["latency_p99_ms"], 3),
                           "mean": round(st["latency_mean_ms"], 3), "max": round(st["latency_max_ms"], 3)},
            "latency_budget_ms": {"p99": SERVE_P99_BUDGET_MS, "met": st["latency_p99_ms"] <= SERVE_P99_BUDGET_MS},
            "server": {k: st[k] for k in ("completed", "rejected", "batches", "mean_batch", "max_inflight_seen",
                                          "peak_bytes", "wall_s", "overlapped_batches", "streams")},
            "rate_sweep": sweep, "batching_engages_at_per_s": engages[0] if engages else None,
            "config": {"workload": "serve", "arrival_rate_per_s": a.rate, "seconds": a.seconds,
                       "sweep_seconds": sweep_seconds,
                       "scenes_triangles": [4_012, 40_012, 160_012],
                       "scene_points": [256_000, 1_000_000, 4_000_000], "resolutions": sizes, "vpls": vpls,
                       "hbm_budget_gib": a.budget_gib, "rsm": 128}}


def other_configs(a, ctx) -> dict:
    """The other configurations, briefly, in the default run, so that its one
    line carries them too: the 4K stress frames (K frames each way), the
    64-frame batch (50 replays) and a 10 s served trace at the arrival rate
    (1 s per sweep rate); each entry holds the headline figures of that
    workload's own line (`--workload 4k | batch64 | serve` runs them in full)."""
    import argparse
    from mlgi import workloads as W

    def keep(d, keys):
        return {k: d[k] for k in keys if k in d}

    b = argparse.Namespace(**vars(a))
    b.seconds = 10.0
    k4 = stress_bench(b, ctx, W.WORKLOADS["4k"])
    bb = argparse.Namespace(**vars(a))
    bb.steps = 50
    b64 = batch_bench(bb, ctx)
    sv = serve_bench(b, ctx, sweep_seconds=1.0)
    return {"4k": keep(k4, ("metric", "value", "unit", "steps", "serial_ms_per_frame")),
            "batch64": keep(b64, ("metric", "value", "unit", "steps", "ms_per_frame")),
            "serve": dict(keep(sv, ("metric", "value", "unit", "latency_ms", "latency_budget_ms",
                                    "batching_engages_at_per_s")),
                          seconds=b.seconds, sweep_seconds=1.0,
                          mean_batch=sv["server"]["mean_batch"],
                          overlapped_batches=sv["server"]["overlapped_batches"],
                          completed=sv["server"]["completed"], rejected=sv["server"]["rejected"])}


# ---------------------------------------------------------------- summaries
# rocprofv3 output -> the summaries under profiles/ (no GPU needed).
# Profile keys and the kernel each one names (batched kernels; _single: the
# single-frame entry point's kernels of the mixed form).
PROFILE_KERNELS = {
    "pullpush": "batch_pullpush_kernel<", "final_gather": "batch_gather_kernel<true>",
    "final_gather_fp32": "batch_gather_kernel<false>", "ism_mask": "batch_mask_kernel",
    "ism_tiles": "batch_tiles_kernel", "ism_spill": "batch_spill_kernel", "vpl": "batch_vpl_kernel",
    "pullpush_single": "pullpush_kernel<", "final_gather_single": "final_gather_kernel<16, 16, true, 0, true>",
    "final_gather_fp32_single": "final_gather_kernel<16, 16, true, 0, false>", "ism_mask_single": "ism_mask_kernel",
    "ism_tiles_single": "ism_tiles_kernel", "ism_spill_single": "ism_spill_kernel", "vpl_single": "vpl_extract_kernel",
}


def short_kernel_name(name: str) -> str:
    """'void mlgi::(anonymous namespace)::batch_gather_kernel<true>(args)' -> 'batch_gather_kernel<true>'."""
    n = name.replace("(anonymous namespace)", "").split("(")[0]
    return n.replace("void ", "").split("::")[-1].strip()


def counter_dispatches(root: str) -> dict:
    """{kernel: {dispatch: {counter: value}}} of a --pmc run."""
    import csv
    import glob
    from collections import defaultdict
    out = defaultdict(lambda: defaultdict(lambda: defaultdict(float)))
    for f in glob.glob(f"{root}/**/*counter_collection.csv", recursive=True):
        for r in csv.DictReader(open(f)):
            out[short_kernel_name(r["Kernel_Name"])][(f, r["Dispatch_Id"])][r["Counter_Name"]] += float(
                r["Counter_Value"])
    return out


def counter_means(root: str) -> dict:
    """{kernel: {"dispatches": n, counter: mean per dispatch}} of a --pmc run."""
    out = {}
    for k, disp in counter_dispatches(root).items():
        names = sorted({c for d in disp.values() for c in d})
        out[k] = {"dispatches": len(disp), **{c: sum(d.get(c, 0.0) for d in disp.values()) / len(disp) for c in names}}
    return out


def profile_rows(means: dict) -> dict:
    out = {}
    for key, prefix in PROFILE_KERNELS.items():
        hits = [k for k in means if k.startswith(prefix)]
        if hits:
            out[key] = {"kernel": hits[0], **means[hits[0]]}
    return out


def summarize(kind: str, dirs: list, how: str) -> dict:
    """kind sq: SQ counters per profile key; traffic: FETCH_SIZE (dirs[0])
    and WRITE_SIZE (dirs[1]) in bytes per launch, each scaled by what the
    counter reads for the --calibrate-copy-mib copy (dirs[2]: its MiB)."""
    if kind == "sq":
        return {"_how": how, **profile_rows(counter_means(dirs[0]))}
    fetch, write, mib = counter_means(dirs[0]), counter_means(dirs[1]), int(dirs[2])
    copy = next(k for k in fetch if "copyBuffer" in k)
    kib = mib * 1024.0
    # the calibration copy: the largest single dispatch of the copy kernel
    fk = max(d["FETCH_SIZE"] for d in counter_dispatches(dirs[0])[copy].values())
    wk = max(d["WRITE_SIZE"] for d in counter_dispatches(dirs[1])[copy].values())
    fs, ws = kib / fk, kib / wk
    out = {"_how": how, "_calibration": {"copy_kernel": copy, "fetch_kib_read": fk, "fetch_scale": round(fs, 4),
                                         "write_kib_read": wk, "write_scale": round(ws, 4)}}
    for key, row in profile_rows(fetch).items():
        w = write.get(row["kernel"], {}).get("WRITE_SIZE")
        if w is None:
            continue
        f = row["FETCH_SIZE"]
        out[key] = {"kernel": row["kernel"], "fetch_kib": round(f, 1), "write_kib": round(w, 1),
                    "bytes_per_launch": int(round((f * fs + w * ws) * 1024))}
    return out


def trace_table(root: str) -> str:
    """Markdown table of a --kernel-trace --stats run (its kernel_stats.csv)."""
    import csv
    import glob
    rows = []
    for f in glob.glob(f"{root}/**/*kernel_stats.csv", recursive=True):
        rows += list(csv.DictReader(open(f)))
    rows.sort(key=lambda r: -float(r["Percentage"]))
    out = ["| kernel | calls | avg us | min us | max us | % of GPU time |", "|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| `{short_kernel_name(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                   f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200,
                    help="timed steps K per form (200: the pipelined forms' fill and drain, about three "
                         "frames' time, then weigh ~2%% in the per-frame figure)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="1080p", choices=["1080p", "small", "4k", "batch64", "serve"])
    ap.add_argument("--form", default="both", choices=["fp32", "mixed", "both"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-others", action="store_true",
                    help="1080p: leave out the brief runs of the 4k, batch64 and serve workloads (other_configs)")
    ap.add_argument("--slots", type=int, default=3,
                    help="frame workloads: frames in flight for the *_pipelined forms (1: not run)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0: host_threads())")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: Poisson arrival rate (req/s)")
    ap.add_argument("--seconds", type=float, default=60.0, help="serve: length of the arrival trace")
    ap.add_argument("--budget-gib", type=int, default=16, help="serve: HBM budget of admitted requests")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="before the frames, copy this many MiB device-to-device (a plain copy kernel of "
                         "known size for scaling the HBM traffic counters)")
    ap.add_argument("--pmc-summary", nargs="+", metavar="ARG",
                    help="no GPU: summarise rocprofv3 output for profiles/ and exit: 'sq DIR' (SQ counters), "
                         "'traffic FETCH_DIR WRITE_DIR COPY_MIB' or 'trace DIR' (kernel-trace stats table)")
    ap.add_argument("--how", default="", help="--pmc-summary: the commands that made the input, for the summary")
    a = ap.parse_args()
    if a.pmc_summary:
        if a.pmc_summary[0] == "trace":
            print(trace_table(a.pmc_summary[1]))
        else:
            print(json.dumps(summarize(a.pmc_summary[0], a.pmc_summary[1:], a.how), indent=1))
        return 0

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(the sharding design is documented, not built).", file=sys.stderr)
        return 2

    import numpy as np
    import mlgi
    from mlgi import workloads as W

    with mlgi.Context(0) as ctx:
        if a.calibrate_copy_mib > 0:
            n = a.calibrate_copy_mib << 20
            src, dst = mlgi.DeviceArray(ctx, n, np.uint8), mlgi.DeviceArray(ctx, n, np.uint8)
            dst.copy_from_device(src)
            ctx.sync()
            log(f"[bench] calibration copy of {a.calibrate_copy_mib} MiB done")
            src.free()
            dst.free()
        if a.workload == "batch64":
            out = batch_bench(a, ctx)
        elif a.workload == "serve":
            out = serve_bench(a, ctx)
        elif a.workload == "4k":
            out = stress_bench(a, ctx, W.WORKLOADS["4k"])
        else:
            out = frame_bench(a, ctx, W.WORKLOADS[a.workload])
            if a.workload == "1080p" and not a.no_others:
                out["other_configs"] = other_configs(a, ctx)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
