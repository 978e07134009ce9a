"""GI frame benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 1080p]
                    [--form fp32|mixed|both] [--no-cpu]

A step is one GI frame of the many-lights hot path on inputs already resident
in HBM: VPL extraction from the RSM, ISM atlas clear + point splat, pull-push,
and interleaved final gathering (mlgi_frame). The timed region is exactly K
frames bracketed by a device synchronize on both sides; ms_per_step is wall
time / K. Rank 0 prints one JSON line.

The default workload is the 1920x1080, 4,096-VPL configuration (40k-triangle
procedural room, 128^2 RSM, 1M splat points into 4,096 64^2 ISMs of a 4096^2
atlas, 4 pull-push levels, 8x8 interleave, 16x16 tiles).

roofline: for the kernel with the largest mean duration (HIP events on the
context stream over the timed frames): pull-push against HBM (6 bytes per
atlas texel: the u32 atlas read once, the unorm16 ISM written once); the
final gather against the FP32 vector peak (84 flops per lit pixel-VPL pair,
pairs counted on the device). Every kernel's figure is under "kernels";
"traffic" is the HBM bytes per launch measured with rocprofv3 counters
(profiles/traffic_<workload>.json), null until measured.
cpu_baseline: the oracle's FP32 instantiation of the same four stages run on
the host for one full frame (final gathering threaded, the rest on one
thread), scaled to ms/frame.
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
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(ctx, wl, inp, precision: int, steps: int, warmup: int):
    import mlgi

    p = wl.params(precision=precision)
    st = mlgi.GIStage(ctx, p, wl.n_vpl, wl.width, wl.height)
    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    for _ in range(warmup):
        st.process(rsm, pts, gb)
    ctx.sync()
    # timed region: K frames between device synchronisations; HIP events on
    # the context stream bracket every stage of every frame
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        st.process(rsm, pts, gb)
    ctx.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    stages = ctx.profile_read()
    ctx.profile(False)
    lit = st.count_lit_pairs(gb)
    return {"st": st, "rsm": rsm, "pts": pts, "gb": gb, "wall_ms": wall_ms, "stages": stages, "lit": lit}


def kernel_rooflines(wl, stages, lit):
    """Per-kernel roofline figures from the live stage times.

    pull-push: reads the u32 atlas once and writes the unorm16 ISM once,
      6 bytes per atlas texel, against HBM.
    splat: reads every 16-byte point record once (lower bound: the accepted
      atomicMins add at most 4 x 4 bytes per point), against HBM.
    final gather: 84 flops per lit pixel-VPL pair, against the FP32 vector
      peak.
    """
    a2 = wl.atlas_side * wl.atlas_side
    out = {}
    t = stages["pullpush_ms"] * 1e-3
    if t > 0:
        gbs = 6.0 * a2 / t / 1e9
        out["pullpush"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": 6 * a2,
                           "kernel_ms": round(stages["pullpush_ms"], 4)}
    t = stages["ism_splat_ms"] * 1e-3
    if t > 0:
        gbs = 16.0 * wl.n_points / t / 1e9
        out["ism_splat"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": 16 * wl.n_points,
                            "kernel_ms": round(stages["ism_splat_ms"], 4)}
    t = stages["final_gather_ms"] * 1e-3
    if t > 0:
        tf = FLOPS_PER_LIT_PAIR * lit / t / 1e12
        out["final_gather"] = {"bound": "valu", "achieved": round(tf, 3), "peak": FP32_VECTOR_PEAK_TFLOPS,
                               "unit": "TFLOP/s", "frac": round(tf / FP32_VECTOR_PEAK_TFLOPS, 4),
                               "lit_pairs": lit, "flops_per_lit_pair": FLOPS_PER_LIT_PAIR,
                               "kernel_ms": round(stages["final_gather_ms"], 4)}
    return out


def cpu_baseline(wl, inp, rec, r16, threads: int):
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
    atlas = oracle.ism_splat(inp.points, inp.bin_offsets, pos, nrm, wl.atlas_side, p.z_far)
    t["splat"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    r16c = oracle.pullpush(atlas, wl.ism_side, wl.levels, p.z_far)
    t["pullpush"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.final_gather(inp.gb_depth, inp.gb_normal, inp.vp_inv, pos, nrm, col, r16c, wl.atlas_side, p.z_far,
                        p.gi_intensity_factor, p.vpl_clamping_value, wl.interleave, threads=threads,
                        dtype=np.float32)
    t["gather"] = time.perf_counter() - t0
    total_ms = sum(t.values()) * 1e3
    return {"value": round(total_ms, 3), "unit": "ms/frame", "cores": threads, "kind": "port",
            "sample": (f"one full {wl.name} frame: oracle FP32 instantiation; final gathering on {threads} "
                       "threads, VPL extraction/splat/pull-push on 1 thread"),
            "stages_s": {k: round(v, 4) for k, v in t.items()}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="1080p")
    ap.add_argument("--form", default="both", choices=["fp32", "mixed", "both"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="before the frames, copy this many MiB device-to-device (a plain copy kernel of "
                         "known size for scaling the HBM traffic counters)")
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(the sharding design is documented, not built).", file=sys.stderr)
        return 2

    import numpy as np
    import mlgi
    from mlgi import workloads as W

    wl = W.WORKLOADS[a.workload]
    log(f"[bench] building inputs for {wl.name} ...")
    t0 = time.perf_counter()
    inp = W.build_inputs(wl)
    log(f"[bench] inputs ready in {time.perf_counter() - t0:.1f}s")
    forms = {"fp32": [mlgi.PRECISION_FP32], "mixed": [mlgi.PRECISION_MIXED],
             "both": [mlgi.PRECISION_FP32, mlgi.PRECISION_MIXED]}[a.form]
    res = {}
    with mlgi.Context(0) as ctx:
        if a.calibrate_copy_mib > 0:
            n = a.calibrate_copy_mib << 20
            src, dst = mlgi.DeviceArray(ctx, n, np.uint8), mlgi.DeviceArray(ctx, n, np.uint8)
            dst.copy_from_device(src)
            ctx.sync()
            log(f"[bench] calibration copy of {a.calibrate_copy_mib} MiB done")
            src.free()
            dst.free()
        for prec in forms:
            name = "fp32" if prec == mlgi.PRECISION_FP32 else "mixed"
            log(f"[bench] form {name}: warmup {a.warmup}, steps {a.steps}")
            r = run_form(ctx, wl, inp, prec, a.steps, a.warmup)
            r["rec"] = r["st"].vpl_processor.host_records()
            r["r16"] = r["st"].ism.pushpull_result.numpy()
            r["img"] = r["st"].gi_buffer.numpy()
            log(f"[bench] {name}: {r['wall_ms']:.4f} ms/frame, stages {r['stages']}")
            res[name] = r
    head = "mixed" if "mixed" in res else "fp32"
    h = res[head]
    kernels = kernel_rooflines(wl, h["stages"], h["lit"])
    dominant = max(kernels, key=lambda k: kernels[k]["kernel_ms"])
    roof = dict(kernels[dominant])
    roof["kernel"] = dominant
    roof["traffic"] = None
    tpath = ROOT / "profiles" / f"traffic_{wl.name}.json"
    if tpath.exists():
        tr = json.loads(tpath.read_text()).get(dominant)
        if tr:
            roof["traffic"] = tr["bytes_per_launch"]
    out = {
        "metric": "GI ms/frame 1080p·4096 VPL (mixed vs FP32 HIP)",
        "value": round(h["wall_ms"], 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(h["wall_ms"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16x2+fp32" if head == "mixed" else "fp32",
        "data": "synthetic: seeded procedural room, ray-cast G-buffer/RSM, area-proportional points",
        "config": {"workload": wl.name, "width": wl.width, "height": wl.height, "vpls": wl.n_vpl,
                   "interleave": wl.interleave, "points": wl.n_points, "atlas": wl.atlas_side,
                   "ism": wl.ism_side, "pull_push_levels": wl.levels, "tile": f"{wl.tile[0]}x{wl.tile[1]}",
                   "form": head, "triangles": wl.n_random + 12},
        "forms": {k: {"ms_per_frame": round(v["wall_ms"], 4),
                      "stages_ms": {s: round(x, 4) for s, x in v["stages"].items() if s != "frames"}}
                  for k, v in res.items()},
        "roofline": roof,
        "kernels": kernels,
    }
    if not a.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(wl, inp, h["rec"], h["r16"], a.cpu_threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
