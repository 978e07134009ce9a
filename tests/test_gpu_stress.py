"""GPU: the 4K stress configuration.

3840x2160 G-buffer, 160k-triangle scene, 16,384 VPLs, 8M splat points into
16,384 ISMs of 32^2 (4096^2 atlas), 5 pull-push levels, 8x8 interleave,
mixed form through the batched kernels replayed from a HIP graph, with the
light pose changed every frame (RSM rendered on the device). Frames 0 and 30
of 31 are checked stage-wise against the oracle: VPL records, atlas and ISM
bit-exact, irradiance within the mixed form's rel-L2 2e-3 of the FP64
transcription. On frame 30 the single-frame gather then runs on the same VPLs
and ISM in the plain FP32 form (rel-L2 <= 1e-4 of the same transcription)
and, in both forms, with the VPL culling switched off: culled and brute-force
sums are equal bit for bit.
"""
import numpy as np
import pytest

import mlgi
import oracle
from mlgi import workloads as W

pytestmark = pytest.mark.gpu


def test_4k_light_per_frame_graph_replays(ctx):
    wl = W.WORKLOADS["4k"]
    inp = W.build_inputs(wl)
    scene = W._scene(wl.seed, wl.n_random)
    dev = scene.upload(ctx)
    p = wl.params(precision=mlgi.PRECISION_MIXED)
    st = mlgi.GIStage(ctx, p, wl.n_vpl, wl.width, wl.height)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    r = wl.rsm
    dif, nrm, dep = ctx.empty((r, r, 3), np.float32), ctx.empty((r, r, 3), np.float32), ctx.empty((r, r), np.float32)
    rsm = mlgi.rsm_desc(dif, nrm, dep, inp.inv_bvp)
    batch = mlgi.Batch(ctx, [(st, rsm, pts, gb)])
    g = batch.capture()
    print(f"[4k] {wl}", flush=True)
    sh = mlgi.shuffle_indices(wl.n_vpl)
    for i in range(31):
        light = mlgi.Camera.light(1000 + i)
        inv = dev.render_rsm(light, dif, nrm, dep)
        batch.update_light(0, inv)
        g.launch()
        if i % 30:
            continue
        ctx.sync()
        rec = st.vpl_processor.host_records()
        atlas = st.ism.softrender.numpy()
        r16 = st.ism.pushpull_result.numpy()
        img = st.gi_buffer.numpy()
        hd, hn, hdep, hinv = scene.render_rsm(light, r, r)
        assert np.array_equal(hinv, inv)
        pos, vn, pk, col = oracle.vpl_extract(hd, hn, hdep, hinv, p.light_intensity, sh, wl.n_vpl, np.float32)
        assert np.array_equal(rec[:, 0:3], pos) and np.array_equal(rec[:, 16:19], col), i
        assert np.array_equal(atlas, oracle.ism_splat(inp.points, inp.bin_offsets, pos, vn, wl.atlas_side,
                                                      p.z_far)), i
        assert np.array_equal(r16, oracle.pullpush(atlas, wl.ism_side, wl.levels, p.z_far)), i
        ref, lit = oracle.final_gather(inp.gb_depth, inp.gb_normal, inp.vp_inv, pos, vn, col, r16, wl.atlas_side,
                                       p.z_far, p.gi_intensity_factor, p.vpl_clamping_value, wl.interleave,
                                       threads=16)
        err = float(np.linalg.norm(img - ref) / np.linalg.norm(ref))
        print(f"[4k] frame {i}: rel-L2 {err:.3e}, lit pairs {lit}", flush=True)
        assert err <= 2e-3, i
    ctx.sync()
    for name, prec, bar in (("fp32", mlgi.PRECISION_FP32, 1e-4), ("mixed", mlgi.PRECISION_MIXED, 2e-3)):
        imgs = []
        for cull_off in (0, 1):
            st.params.precision = prec
            st.params.disable_culling = cull_off
            st.gi_buffer.copy_from(np.full((wl.height, wl.width, 3), -1.0, np.float32))
            st.compute_final_gathering(gb)
            ctx.sync()
            imgs.append(st.gi_buffer.numpy())
        st.params.precision, st.params.disable_culling = mlgi.PRECISION_MIXED, 0
        err = float(np.linalg.norm(imgs[0] - ref) / np.linalg.norm(ref))
        diff = int((imgs[0] != imgs[1]).sum())
        print(f"[4k] frame 30 single-frame {name}: rel-L2 {err:.3e}, culled vs brute force {diff} differing", flush=True)
        assert err <= bar, name
        assert diff == 0, name
    g.close()
    batch.close()
    dev.close()
