"""GPU: batched frames.

A batch of frames with different G-buffer sizes, VPL counts, ISM sides,
interleaves, gather strips, light poses and both precisions runs as one set
of launches, one per stage (one workgroup per work item of all frames).
Every frame's VPL records, atlas and unorm16 ISM must be bit-identical to
the same frame run alone through the single-frame path, and its irradiance
within rel-L2 1e-5 of it (the batched gather runs one wave per strip, the
single-frame kernel one workgroup per tile, so the FP32 forms may contract
differently); every output must be bit-identical to the frame run as a
batch of one, to a second run, and to replays of the captured batch graph;
one frame is also checked against the oracle.
"""
import numpy as np
import pytest

import mlgi
import oracle
from mlgi import workloads as W

pytestmark = pytest.mark.gpu

SHAPES = [  # (width, height, n_vpl, interleave, atlas_side, tile_w, precision, light_seed)
    (96, 64, 256, 4, 512, 8, mlgi.PRECISION_MIXED, 7),
    (128, 128, 1024, 8, 1024, 16, mlgi.PRECISION_MIXED, 11),
    (64, 200, 256, 8, 512, 32, mlgi.PRECISION_FP32, 13),
    (101, 67, 1024, 4, 2048, 16, mlgi.PRECISION_FP32, 17),
    (160, 90, 2048, 8, 4096, 8, mlgi.PRECISION_MIXED, 19),
]


def make_frame(ctx, shape):
    w, h, n, il, atlas, tw, prec, seed = shape
    wl = W.WORKLOADS["tiny"].with_(width=w, height=h, n_vpl=n, interleave=il, atlas_side=atlas,
                                   light_seed=seed, n_points=20_000 if n <= 1024 else 60_000, rsm=64)
    inp = W.build_inputs(wl)
    # single-frame workgroup tile whose waves are the batched 2^w x 64/2^w strips
    p = wl.params(precision=prec, tile_w=tw, tile_h={8: 8, 16: 4, 32: 8}[tw], pull_push_levels=3)
    st = mlgi.GIStage(ctx, p, n, w, h)
    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    return wl, inp, st, rsm, pts, gb


def outputs(st):
    return (st.vpl_processor.host_records(), st.ism.softrender.numpy(), st.ism.pushpull_result.numpy(),
            st.gi_buffer.numpy())


def clobber(st, wl):
    st.gi_buffer.copy_from(np.full((wl.height, wl.width, 3), -1.0, np.float32))
    st.ism.pushpull_result.copy_from(np.zeros((wl.atlas_side, wl.atlas_side), np.uint16))


@pytest.fixture(scope="module")
def frames(ctx):
    fr = [make_frame(ctx, s) for s in SHAPES]
    print(f"[batch] {len(fr)} frames: {SHAPES}", flush=True)
    return fr


def single_path_outputs(ctx, frames):
    ref = []
    for wl, inp, st, rsm, pts, gb in frames:
        # the single-frame gather takes workgroup tiles; the batched strip
        # 2^w x 64/2^w is one wave of the same tile shape
        st.process(rsm, pts, gb)
        ctx.sync()
        ref.append(outputs(st))
    return ref


def test_batch_matches_single_frames(ctx, frames):
    ref = single_path_outputs(ctx, frames)
    for wl, inp, st, rsm, pts, gb in frames:
        clobber(st, wl)
    b = mlgi.Batch(ctx, [(st, rsm, pts, gb) for wl, inp, st, rsm, pts, gb in frames])
    b.run()
    ctx.sync()
    first = [outputs(f[2]) for f in frames]
    for i, (wl, inp, st, rsm, pts, gb) in enumerate(frames):
        got = first[i]
        for name, x, y in zip(("vpl", "atlas", "ism"), got[:3], ref[i][:3]):
            assert np.array_equal(x, y), (i, name, SHAPES[i])
        err = float(np.linalg.norm(got[3] - ref[i][3]) / np.linalg.norm(ref[i][3]))
        print(f"[batch] frame {i} vs single-frame path: rel-L2 {err:.2e}")
        assert err <= 1e-5, (i, err)
    b.run()  # a second run gives the same bits
    ctx.sync()
    for i, f in enumerate(frames):
        for x, y in zip(outputs(f[2]), first[i]):
            assert np.array_equal(x, y), i
    b.close()


# ISM side 16, three levels, 4 and 256 ISMs: one pull-push launch whose
# workgroups take 4 ISMs in the first frame and 16 in the second.
GROUP_SHAPES = [
    (48, 32, 4, 2, 32, 8, mlgi.PRECISION_MIXED, 23),
    (64, 48, 256, 4, 256, 8, mlgi.PRECISION_MIXED, 29),
]


def test_batch_pullpush_groups_of_different_sizes(ctx):
    fr = [make_frame(ctx, s) for s in GROUP_SHAPES]
    assert [f[0].ism_side for f in fr] == [16, 16]
    ref = single_path_outputs(ctx, fr)
    for wl, inp, st, rsm, pts, gb in fr:
        clobber(st, wl)
    b = mlgi.Batch(ctx, [(st, rsm, pts, gb) for wl, inp, st, rsm, pts, gb in fr])
    b.run()
    ctx.sync()
    for i, (wl, inp, st, rsm, pts, gb) in enumerate(fr):
        got = outputs(st)
        for name, x, y in zip(("vpl", "atlas", "ism"), got[:3], ref[i][:3]):
            assert np.array_equal(x, y), (i, name, GROUP_SHAPES[i])
        assert np.array_equal(got[2], oracle.pullpush(got[1], 16, 3, 50.0)), i
    b.close()


def test_batch_of_one_and_graph_replays(ctx, frames):
    big = mlgi.Batch(ctx, [(st, rsm, pts, gb) for wl, inp, st, rsm, pts, gb in frames])
    big.run()
    ctx.sync()
    mixed = [outputs(f[2]) for f in frames]
    for i, (wl, inp, st, rsm, pts, gb) in enumerate(frames):
        clobber(st, wl)
        one = mlgi.Batch(ctx, [(st, rsm, pts, gb)])
        one.run()
        ctx.sync()
        for x, y in zip(outputs(st), mixed[i]):
            assert np.array_equal(x, y), i
        one.close()
    g = big.capture()
    for rep in range(3):
        for wl, inp, st, rsm, pts, gb in frames:
            clobber(st, wl)
        g.launch()
        ctx.sync()
        for i, f in enumerate(frames):
            for x, y in zip(outputs(f[2]), mixed[i]):
                assert np.array_equal(x, y), (rep, i)
        big.run()  # eager work between replays
        ctx.sync()
    g.close()
    big.close()


def test_batch_frame_against_oracle(ctx, frames):
    wl, inp, st, rsm, pts, gb = frames[1]
    b = mlgi.Batch(ctx, [(f[2], f[3], f[4], f[5]) for f in frames])
    b.run()
    ctx.sync()
    rec, atlas, r16, img = outputs(st)
    pos, nrm, col = rec[:, 0:3], rec[:, 4:7], rec[:, 16:19]
    assert np.array_equal(atlas, oracle.ism_splat(inp.points, inp.bin_offsets, pos, nrm, wl.atlas_side, 50.0))
    assert np.array_equal(r16, oracle.pullpush(atlas, wl.ism_side, 3, 50.0))
    p = st.params
    ref, _ = oracle.final_gather(inp.gb_depth, inp.gb_normal, inp.vp_inv, pos, nrm, col, r16, wl.atlas_side,
                                 p.z_far, p.gi_intensity_factor, p.vpl_clamping_value, wl.interleave)
    err = float(np.linalg.norm(img - ref) / np.linalg.norm(ref))
    print(f"[batch] frame 1 (mixed) rel-L2 {err:.3e}")
    assert err <= 2e-3
    b.close()


def test_batch_rejects_bad_frames(ctx, frames):
    wl, inp, st, rsm, pts, gb = frames[0]
    st.params.tile_w = 64
    try:
        with pytest.raises(mlgi.MlgiError) as e:
            mlgi.Batch(ctx, [(st, rsm, pts, gb)])
        assert e.value.code == mlgi.N.MLGI_ERR_UNSUPPORTED
    finally:
        st.params.tile_w = 8


def test_autotune_picks_and_caches(ctx):
    shape = (128, 96, 1024, 4, 1024, 8, mlgi.PRECISION_MIXED, 23)
    wl, inp, st, rsm, pts, gb = make_frame(ctx, shape)
    r1 = mlgi.autotune(ctx, st, rsm, pts, gb)
    print(f"[tune] {r1}")
    assert r1["interleave"] == 4  # 1024 / 8^2 = 16 < 32 VPLs per class
    assert r1["tile"] in [(8, 8), (16, 4), (32, 8)] and not r1["cached"]
    assert all(t > 0 for t in r1["gather_ms"])
    r2 = mlgi.autotune(ctx, st, rsm, pts, gb)
    assert r2["cached"] and r2["tile"] == r1["tile"]
    b = mlgi.Batch(ctx, [(st, rsm, pts, gb)])
    b.run()
    ctx.sync()
    rec, atlas, r16, img = outputs(st)
    p = st.params
    ref, _ = oracle.final_gather(inp.gb_depth, inp.gb_normal, inp.vp_inv, rec[:, 0:3], rec[:, 4:7], rec[:, 16:19],
                                 r16, wl.atlas_side, p.z_far, p.gi_intensity_factor, p.vpl_clamping_value, 4)
    assert float(np.linalg.norm(img - ref) / np.linalg.norm(ref)) <= 2e-3
    b.close()
