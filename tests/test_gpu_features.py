"""GPU: graphs, kernel variants, edge cases and error handling.

* A captured frame graph replayed several times, with eager work and host
  synchronisation in between, reproduces the eager frame bit for bit.
* Every final-gather tile shape and VPL-fetch variant, and every splat
  workgroup size, gives bit-identical results (same arithmetic, different
  scheduling).
* Edge cases the reference meets: an empty point set (atlas stays cleared,
  ISM all far), image sizes that are not multiples of the interleave or the
  tile (ragged borders), interleave 1/2/4/8, shadowing off, VPL counts that
  are not perfect squares (RSM rows outside the texture read as 0).
* Invalid descriptors are rejected with MLGI_ERR_INVALID_ARGUMENT /
  MLGI_ERR_UNSUPPORTED before any launch.
"""
import ctypes as C

import numpy as np
import pytest

import mlgi
import oracle
from mlgi import _native as N
from mlgi import workloads as W

pytestmark = pytest.mark.gpu


def setup_frame(ctx, wl, **pkw):
    inp = W.build_inputs(wl)
    p = wl.params(**pkw)
    st = mlgi.GIStage(ctx, p, wl.n_vpl, wl.width, wl.height)
    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    print(f"[setup] {wl.name} {wl.width}x{wl.height} N={wl.n_vpl} I={wl.interleave} {pkw}", flush=True)
    return inp, p, st, rsm, pts, gb


def outputs(st):
    return (st.vpl_processor.host_records(), st.ism.softrender.numpy(), st.ism.pushpull_result.numpy(),
            st.gi_buffer.numpy())


def check_against_oracle(inp, wl, p, rec, atlas, r16, img, tol=1e-4, points=None, bin_offsets=None):
    pos, nrm, col = rec[:, 0:3], rec[:, 4:7], rec[:, 16:19]
    points = inp.points if points is None else points
    bin_offsets = inp.bin_offsets if bin_offsets is None else bin_offsets
    ref_atlas = oracle.ism_splat(points, bin_offsets, pos, nrm, wl.atlas_side, p.z_far)
    assert np.array_equal(atlas, ref_atlas)
    ref_r16 = oracle.pullpush(atlas, wl.ism_side, wl.levels, p.z_far)
    assert np.array_equal(r16, ref_r16)
    il = wl.interleave if p.use_interleaving else 1
    ref, _ = oracle.final_gather(inp.gb_depth, inp.gb_normal, inp.vp_inv, pos, nrm, col, r16, wl.atlas_side,
                                 p.z_far, p.gi_intensity_factor, p.vpl_clamping_value, il,
                                 shadowing=bool(p.enable_shadowing))
    err = np.linalg.norm(img - ref) / max(np.linalg.norm(ref), 1e-30)
    assert err <= tol, err


def test_graph_replays_match_eager(ctx):
    wl = W.WORKLOADS["small"]
    inp, p, st, rsm, pts, gb = setup_frame(ctx, wl)
    st.process(rsm, pts, gb)
    ctx.sync()
    eager = outputs(st)
    g = st.capture(rsm, pts, gb)
    for rep in range(4):
        # clobber the outputs, replay, compare
        st.gi_buffer.copy_from(np.zeros((wl.height, wl.width, 3), np.float32))
        st.ism.pushpull_result.copy_from(np.zeros((wl.atlas_side, wl.atlas_side), np.uint16))
        g.launch()
        ctx.sync()
        got = outputs(st)
        for a, b in zip(got, eager):
            assert np.array_equal(a, b), rep
        st.process(rsm, pts, gb)  # eager work between replays
        ctx.sync()
    g.close()


def test_variants_bit_identical(ctx):
    wl = W.WORKLOADS["small"]
    inp, p, st, rsm, pts, gb = setup_frame(ctx, wl)
    st.process(rsm, pts, gb)
    ctx.sync()
    ref = outputs(st)
    for tile in [(8, 8), (16, 4), (16, 16), (32, 8)]:
        for fetch in (0, 1):
            st.params.tile_w, st.params.tile_h, st.params.gather_fetch = tile[0], tile[1], fetch
            st.compute_final_gathering(gb)
            ctx.sync()
            assert np.array_equal(st.gi_buffer.numpy(), ref[3]), (tile, fetch)
    for blk in (64, 128, 256):
        st.params.splat_block = blk
        st.ism.render(pts, st.vpl_processor, st.params)
        ctx.sync()
        assert np.array_equal(st.ism.softrender.numpy(), ref[1]), blk


def test_empty_point_set(ctx):
    wl = W.WORKLOADS["tiny"]
    inp, p, st, rsm, pts, gb = setup_frame(ctx, wl)
    zeros = np.zeros(wl.n_vpl + 1, np.uint32)
    empty = mlgi.points_desc(ctx.upload(np.zeros((1, 4), np.float32)), ctx.upload(zeros))
    empty.count = 0
    st.process(rsm, empty, gb)
    ctx.sync()
    rec, atlas, r16, img = outputs(st)
    assert np.all(atlas == 0xFFFFFFFF) and np.all(r16 == 65535)
    check_against_oracle(inp, wl, p, rec, atlas, r16, img, points=np.zeros((0, 4), np.float32),
                         bin_offsets=zeros)


@pytest.mark.parametrize("il", [1, 2, 8])
def test_interleave_and_ragged_sizes(ctx, il):
    wl = W.WORKLOADS["tiny"].with_(width=101, height=67, interleave=il, n_vpl=256)
    inp, p, st, rsm, pts, gb = setup_frame(ctx, wl, use_interleaving=1 if il > 1 else 0)
    st.process(rsm, pts, gb)
    ctx.sync()
    check_against_oracle(inp, wl, p, *outputs(st))


def test_shadowing_off(ctx):
    wl = W.WORKLOADS["tiny"]
    inp, p, st, rsm, pts, gb = setup_frame(ctx, wl, enable_shadowing=0)
    st.process(rsm, pts, gb)
    ctx.sync()
    check_against_oracle(inp, wl, p, *outputs(st))


def test_non_square_vpl_count(ctx):
    # N = 2048: sqrt grid 46 x 44, the last rows fall outside the RSM (read 0)
    wl = W.WORKLOADS["tiny"].with_(n_vpl=2048, atlas_side=2048, rsm=64, interleave=4)
    inp, p, st, rsm, pts, gb = setup_frame(ctx, wl)
    st.process(rsm, pts, gb)
    ctx.sync()
    rec, atlas, r16, img = outputs(st)
    sh = mlgi.shuffle_indices(wl.n_vpl)
    pos, nrm, pk, col = oracle.vpl_extract(inp.rsm_diffuse, inp.rsm_normal, inp.rsm_depth, inp.inv_bvp,
                                           p.light_intensity, sh, wl.n_vpl, np.float32)
    assert np.array_equal(rec[:, 0:3], pos) and np.array_equal(rec[:, 16:19], col)
    check_against_oracle(inp, wl, p, rec, atlas, r16, img)


def test_invalid_arguments_rejected(ctx):
    wl = W.WORKLOADS["tiny"]
    inp, p, st, rsm, pts, gb = setup_frame(ctx, wl)
    bad = mlgi.default_params(interleave=3)
    st.params = bad
    with pytest.raises(mlgi.MlgiError) as e:
        st.compute_final_gathering(gb)
    assert e.value.code == N.MLGI_ERR_INVALID_ARGUMENT
    st.params = wl.params(use_push_pull=0)
    with pytest.raises(mlgi.MlgiError) as e:
        st.ism.render(pts, st.vpl_processor, st.params)
    assert e.value.code == N.MLGI_ERR_UNSUPPORTED
    st.params = wl.params(scale_isms=1, vpl_start_index=200, vpl_end_index=300)  # past the 256 VPLs
    with pytest.raises(mlgi.MlgiError) as e:
        st.ism.render(pts, st.vpl_processor, st.params)
    assert e.value.code == N.MLGI_ERR_INVALID_ARGUMENT
    # a far plane outside [2^-60, 2^60] (infinite, denormal, zero): the
    # splat's divisions by zFar are exact only within that range
    for z in (float("inf"), 1e-40, 0.0, 2.0 ** 61):
        st.params = wl.params(z_far=z)
        with pytest.raises(mlgi.MlgiError) as e:
            st.ism.render(pts, st.vpl_processor, st.params)
        assert e.value.code == N.MLGI_ERR_INVALID_ARGUMENT, z
    st.params = wl.params()
    wrong_bins = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(np.zeros(11, np.uint32)))
    with pytest.raises(mlgi.MlgiError) as e:
        st.ism.render(wrong_bins, st.vpl_processor, st.params)
    assert e.value.code == N.MLGI_ERR_INVALID_ARGUMENT
    d = st.ism.depth_desc()
    d.ism_side = 7
    im = st.image_desc()
    rc = mlgi.lib.mlgi_final_gather(ctx.h, C.byref(gb), C.byref(st.vpl_processor.desc()), C.byref(d),
                                    C.byref(st.params), C.byref(im))
    assert rc == N.MLGI_ERR_INVALID_ARGUMENT
    # a pull-push output that overlaps the atlas (written while other
    # workgroups still read it), alone and in a batch
    aliased = st.ism.depth_desc()
    aliased.texels = st.ism.softrender.ptr + 4 * wl.atlas_side  # inside the atlas, 8-byte aligned
    rc = mlgi.lib.mlgi_ism_pullpush(ctx.h, C.byref(st.ism.atlas_desc()), C.byref(st.params), C.byref(aliased))
    assert rc == N.MLGI_ERR_INVALID_ARGUMENT
    fd = mlgi.frame_desc(st, rsm, pts, gb)
    fd.ism.texels = st.ism.softrender.ptr
    bp = C.c_void_p()
    rc = mlgi.lib.mlgi_batch_create(ctx.h, C.byref(fd), 1, C.byref(bp))
    assert rc == N.MLGI_ERR_INVALID_ARGUMENT
    # the context still works afterwards
    st.process(rsm, pts, gb)
    ctx.sync()
