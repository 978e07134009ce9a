"""GPU: per-stage timing (mlgi_profile_*), the PerfCounter analogue.

Adjacent stage boundaries share one HIP event, so the stage spans of a
frame must add up to the frame's span, and every stage must be counted
once per frame. Work a caller enqueues on the same stream between two calls
(here a 256 MiB device-to-device copy, about a frame's time or more) must
not be attributed to any stage or frame: the frame and stage times with the
copies in between must match those without.
"""
import numpy as np
import pytest

import mlgi
from mlgi import workloads as W

pytestmark = pytest.mark.gpu

STAGES = ("vpl_ms", "ism_splat_ms", "pullpush_ms", "final_gather_ms")


def upload_frame(ctx, inp):
    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    return rsm, pts, gb


def timed(ctx, step, frames, between=None):
    for _ in range(2):
        step()
    ctx.sync()
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(frames):
        step()
        if between:
            between()
    ctx.sync()
    t = ctx.profile_read()
    ctx.profile(False)
    return t


def test_stage_times_add_up_and_exclude_caller_work(ctx):
    wl = W.WORKLOADS["small"]
    inp = W.build_inputs(wl)
    rsm, pts, gb = upload_frame(ctx, inp)
    st = mlgi.GIStage(ctx, wl.params(precision=mlgi.PRECISION_MIXED), wl.n_vpl, wl.width, wl.height)
    batch = mlgi.Batch(ctx, [(st, rsm, pts, gb)])
    big = 256 << 20
    src, dst = ctx.empty(big, np.uint8), ctx.empty(big, np.uint8)
    try:
        for path, step in (("single", lambda: st.process(rsm, pts, gb)), ("batched", batch.run)):
            plain = timed(ctx, step, 10)
            busy = timed(ctx, step, 10, between=lambda: dst.copy_from_device(src))
            print(f"[profile] {path}: plain {plain}", flush=True)
            print(f"[profile] {path}: with copies {busy}", flush=True)
            for t in (plain, busy):
                assert t["frames"] == 10, path
                assert all(t[s] > 0.0 for s in STAGES), (path, t)
                total = sum(t[s] for s in STAGES)
                # spans share their boundaries: stages never exceed the frame, and leave
                # at most a few microseconds per frame outside them (argument checks)
                assert total <= t["frame_ms"] * 1.001 + 1e-4, (path, t)
                assert total >= t["frame_ms"] - 0.02, (path, t)
            # a 256 MiB copy takes ~0.1 ms: had a shared event spanned it, the
            # frame (or its first stage) would have grown by that much
            assert busy["frame_ms"] <= plain["frame_ms"] * 1.25 + 0.02, path
            assert busy["vpl_ms"] <= plain["vpl_ms"] * 1.5 + 0.02, path
    finally:
        batch.close()
