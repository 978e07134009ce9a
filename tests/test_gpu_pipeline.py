"""GPU: frames with the light moving every frame, the next frame's RSM
rendered on a second context beside the current frame's GI (RsmPipeline,
mlgi_context_wait). Every frame must give the bits of rendering the RSM and
then running the frame on one stream: VPL records, atlas, ISM and irradiance
bit-identical. Checked frame by frame, and at the end of a burst of frames
enqueued without any host synchronisation, where a render that overtook the
previous frame's GI would have overwritten the RSM buffer that frame reads.
"""
import numpy as np
import pytest

import mlgi
from mlgi import workloads as W

pytestmark = pytest.mark.gpu


def light_of(i):
    return mlgi.Camera.light(2000 + i)


def frame_results(st):
    return (st.vpl_processor.host_records(), st.ism.softrender.numpy(), st.ism.pushpull_result.numpy(),
            st.gi_buffer.numpy())


def assert_same(got, want, what):
    for name, g, w in zip(("VPL records", "atlas", "ISM", "irradiance"), got, want):
        assert np.array_equal(g, w), f"{what}: {name} differs"


def serial_reference(ctx, wl, inp, scene, frames):
    """Frame i's results, rendering each RSM and then running the frame."""
    dev = scene.upload(ctx)
    st = mlgi.GIStage(ctx, wl.params(precision=mlgi.PRECISION_MIXED), wl.n_vpl, wl.width, wl.height)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    r = wl.rsm
    dif, nrm, dep = ctx.empty((r, r, 3), np.float32), ctx.empty((r, r, 3), np.float32), ctx.empty((r, r), np.float32)
    batch = mlgi.Batch(ctx, [(st, mlgi.rsm_desc(dif, nrm, dep, inp.inv_bvp), pts, gb)])
    out = {}
    try:
        for i in frames:
            inv = dev.render_rsm(light_of(i), dif, nrm, dep)
            batch.update_light(0, inv)
            batch.run()
            ctx.sync()
            out[i] = frame_results(st)
    finally:
        batch.close()
        dev.close()
    return out


@pytest.mark.parametrize("name", ["tiny", "small"])
def test_pipelined_rsm_frames_match_serial(ctx, name):
    wl = W.WORKLOADS[name]
    inp = W.build_inputs(wl)
    scene = W._scene(wl.seed, wl.n_random)
    print(f"[pipeline] {name}: {wl}", flush=True)
    ref = serial_reference(ctx, wl, inp, scene, list(range(4)) + [7])

    st = mlgi.GIStage(ctx, wl.params(precision=mlgi.PRECISION_MIXED), wl.n_vpl, wl.width, wl.height)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    pipe = mlgi.RsmPipeline(ctx, scene, st, pts, gb, wl.rsm, light_of)
    try:  # release the second context before any failure leaves the test
        for want in range(4):  # one frame at a time
            i = pipe.frame()
            assert i == want
            ctx.sync()
            assert_same(frame_results(st), ref[i], f"{name} frame {i}")
        for _ in range(4):  # frames 4..7 back to back
            i = pipe.frame()
        pipe.sync()
        assert i == 7
        assert_same(frame_results(st), ref[7], f"{name} frame 7 after a burst")
    finally:
        pipe.close()


def test_context_wait_rejects_null_and_accepts_self(ctx):
    assert mlgi.lib.mlgi_context_wait(None, ctx.h) != 0
    assert mlgi.lib.mlgi_context_wait(ctx.h, None) != 0
    ctx.wait_for(ctx)  # one stream: nothing to order
