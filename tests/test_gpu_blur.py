"""GPU: geometry-aware blur of the irradiance (GIStage::blur, gi_blur.frag).

The fused x/y kernel is bit-exact against the oracle's FP32 instantiation
(same operation order, no contraction) and within 1e-6 relative of FP64, on
the irradiance the gather produced, at tile-aligned and ragged sizes.
"""
import numpy as np
import pytest

import mlgi
import oracle
from mlgi import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,size", [("tiny", None), ("small", None), ("tiny", (101, 67)), ("tiny", (20, 5))])
def test_blur_against_oracle(ctx, name, size):
    wl = W.WORKLOADS[name]
    if size:
        wl = wl.with_(width=size[0], height=size[1])
    inp = W.build_inputs(wl)
    st = mlgi.GIStage(ctx, wl.params(), wl.n_vpl, wl.width, wl.height)
    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    st.process(rsm, pts, gb)
    proj = inp.camera.matrices()[1]
    st.blur(gb, proj)
    ctx.sync()
    gi = st.gi_buffer.numpy()
    got = st.gi_blur_final.numpy()
    ref32 = oracle.gi_blur(gi, inp.gb_normal, inp.gb_depth, proj, np.float32)
    ref64 = oracle.gi_blur(gi, inp.gb_normal, inp.gb_depth, proj, np.float64)
    print(f"[blur] {wl.name} {wl.width}x{wl.height}: mismatches vs FP32 twin {int((got != ref32).sum())}")
    assert np.array_equal(got, ref32)
    assert float(np.linalg.norm(got - ref64) / np.linalg.norm(ref64)) <= 1e-6
    assert not np.array_equal(got, gi)  # it did blur


def test_blur_rejects_in_place(ctx):
    wl = W.WORKLOADS["tiny"]
    inp = W.build_inputs(wl)
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    img = ctx.empty((wl.height, wl.width, 3), np.float32)
    im = mlgi.N.Image()
    im.rgb, im.width, im.height = img.ptr, wl.width, wl.height
    proj = np.ascontiguousarray(inp.camera.matrices()[1], np.float32)
    import ctypes as C
    rc = mlgi.lib.mlgi_gi_blur(ctx.h, C.byref(gb), proj.ctypes.data, C.byref(im), C.byref(im))
    assert rc == mlgi.N.MLGI_ERR_INVALID_ARGUMENT
