"""GPU: the VPL-range forms (scaleISMs, pointsOnlyIntoScaledISMs) against the oracle.

On the 96x64 workload (256 VPLs, 512^2 atlas, 3 pull-push levels), for VPL
ranges with either flag or both: the points are sampled with the range's bin
and ISM counts (as ism.geom bins them), the splat (the point-major form,
which these flags take) and the pull-push are bit-exact against the oracle,
the irradiance is within the form's rel-L2 bar of the FP64 transcription
(every VPL gathered, ISM index v - start with scaleISMs, normalised by the
range's count), and mlgi_frame gives the stage-by-stage bits. The batched
path rejects the flags.
"""
import numpy as np
import pytest

import mlgi
import oracle
from mlgi import _native as N
from mlgi import workloads as W

pytestmark = pytest.mark.gpu

# (the last: 192 VPLs before the range sample negative ISM indices, 32 after
# it write tiles past the range's 32 ISMs)
CASES = [(32, 96, True, False), (32, 96, True, True), (64, 128, False, True), (0, 256, True, True),
         (192, 224, True, False)]


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("start,end,scale,only", CASES)
def test_vpl_range_against_oracle(ctx, start, end, scale, only):
    wl = W.WORKLOADS["tiny"]
    inp = W.build_inputs(wl)
    n, count = wl.n_vpl, end - start
    n_bins = count if only else n
    n_isms = count if scale else n
    pts_h, off_h = W._scene(wl.seed, wl.n_random).sample_points(wl.n_points, n_bins, n_isms, 77)
    p = wl.params(vpl_start_index=start, vpl_end_index=end, scale_isms=int(scale),
                  points_only_into_scaled_isms=int(only))
    side = wl.atlas_side // mlgi.ism_indices_1d(n_isms)
    print(f"[range] [{start}, {end}) scale={scale} points_only={only}: {n_bins} bins, {n_isms} ISMs of {side}^2",
          flush=True)
    st = mlgi.GIStage(ctx, p, n, wl.width, wl.height)
    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    pts = mlgi.points_desc(ctx.upload(pts_h), ctx.upload(off_h))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    st.vpl_processor.process(rsm, p.light_intensity, bool(p.shuffle_lights), p)
    rec = st.vpl_processor.host_records()
    st.ism.render(pts, st.vpl_processor, p)
    assert st.ism.ism_side == side
    atlas = st.ism.softrender.numpy()
    st.ism.pullpush(p)
    r16 = st.ism.pushpull_result.numpy()
    pos, nrm, col = rec[:, 0:3], rec[:, 4:7], rec[:, 16:19]
    ref_atlas = oracle.ism_splat(pts_h, off_h, pos, nrm, wl.atlas_side, p.z_far, vpl_range=(start, end),
                                 scale_isms=scale, points_only=only)
    assert (ref_atlas != 0xFFFFFFFF).any()
    assert np.array_equal(atlas, ref_atlas)
    assert np.array_equal(r16, oracle.pullpush(atlas, side, wl.levels, p.z_far))
    ref, lit = oracle.final_gather(inp.gb_depth, inp.gb_normal, inp.vp_inv, pos, nrm, col, r16, wl.atlas_side,
                                   p.z_far, p.gi_intensity_factor, p.vpl_clamping_value, wl.interleave,
                                   vpl_start=start, vpl_end=end, scale_isms=scale)
    assert lit > 0
    for prec, bar in ((mlgi.PRECISION_FP32, 1e-4), (mlgi.PRECISION_MIXED, 2e-3)):
        st.params.precision = prec
        st.compute_final_gathering(gb)
        ctx.sync()
        err = rel_l2(st.gi_buffer.numpy(), ref)
        print(f"[range] precision {prec}: rel-L2 {err:.3e}", flush=True)
        assert err <= bar, (prec, err)
    # the whole frame through mlgi_frame: the same bits as the stages
    img = st.gi_buffer.numpy()
    st.gi_buffer.copy_from(np.full(img.shape, -1.0, np.float32))
    st.process(rsm, pts, gb)
    ctx.sync()
    assert np.array_equal(st.ism.softrender.numpy(), atlas)
    assert np.array_equal(st.gi_buffer.numpy(), img)


def test_batches_reject_the_range_flags(ctx):
    wl = W.WORKLOADS["tiny"]
    inp = W.build_inputs(wl)
    p = wl.params(vpl_start_index=32, vpl_end_index=96, scale_isms=1)
    st = mlgi.GIStage(ctx, p, wl.n_vpl, wl.width, wl.height)
    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    with pytest.raises(mlgi.MlgiError) as e:
        mlgi.Batch(ctx, [(st, rsm, pts, gb)])
    assert e.value.code == N.MLGI_ERR_UNSUPPORTED
