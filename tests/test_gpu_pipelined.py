"""GPU: frames in flight. Two frame slots, each with its own buffers and its
own context (stream), take frames in turn, as the bench's *_pipelined forms
do; the slots' work overlaps on the device. Every slot's results must be
bit-identical to the same frame run alone on one context, in both precision
forms."""
import numpy as np
import pytest

import mlgi
from mlgi import workloads as W

pytestmark = pytest.mark.gpu


def slot(ctx, wl, inp, prec):
    st = mlgi.GIStage(ctx, wl.params(precision=prec), wl.n_vpl, wl.width, wl.height)
    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    return st, mlgi.Batch(ctx, [(st, rsm, pts, gb)])


@pytest.mark.parametrize("prec", [mlgi.PRECISION_FP32, mlgi.PRECISION_MIXED])
def test_two_slots_match_one_frame_alone(ctx, prec):
    wl = W.WORKLOADS["small"]
    inp = W.build_inputs(wl)
    print(f"[pipelined] {wl.name} precision {prec}", flush=True)
    st0, b0 = slot(ctx, wl, inp, prec)
    b0.run()
    ctx.sync()
    want = (st0.gi_buffer.numpy(), st0.ism.pushpull_result.numpy(), st0.ism.softrender.numpy())
    assert np.abs(want[0]).sum() > 0
    b0.close()
    with mlgi.Context(ctx.device) as c1, mlgi.Context(ctx.device) as c2:
        slots = [slot(c1, wl, inp, prec), slot(c2, wl, inp, prec)]
        try:
            for i in range(8):
                slots[i % 2][1].run()
            c1.sync()
            c2.sync()
            for st, _ in slots:
                got = (st.gi_buffer.numpy(), st.ism.pushpull_result.numpy(), st.ism.softrender.numpy())
                for g, w in zip(got, want):
                    assert np.array_equal(g, w)
        finally:  # the batches go before their contexts
            for _, b in slots:
                b.close()
