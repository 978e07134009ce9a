"""GPU: a captured frame graph owns its splat work area.

mlgi_graph_capture_frame records the ISM-major splat's kernels with a work
area (candidate masks, spill list, spill counters) that belongs to the
graph. Eager splats on the same context use the context's own area, which
grows (is freed and reallocated) when a larger point set arrives; replays
of an earlier graph must never see that. The capture also works on a fresh
context, before any eager splat has sized the context's area.
"""
import numpy as np
import pytest

import mlgi
import oracle
from mlgi import workloads as W

pytestmark = pytest.mark.gpu


def setup(ctx, wl):
    inp = W.build_inputs(wl)
    st = mlgi.GIStage(ctx, wl.params(), wl.n_vpl, wl.width, wl.height)
    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    return inp, st, rsm, pts, gb


def outputs(st):
    return (st.vpl_processor.host_records(), st.ism.softrender.numpy(), st.ism.pushpull_result.numpy(),
            st.gi_buffer.numpy())


def clobber(st):
    st.ism.softrender.copy_from(np.zeros(st.ism.softrender.shape, np.uint32))
    st.ism.pushpull_result.copy_from(np.zeros(st.ism.pushpull_result.shape, np.uint16))
    st.gi_buffer.copy_from(np.full(st.gi_buffer.shape, -1.0, np.float32))


def test_graph_replays_survive_a_larger_eager_splat():
    small = W.WORKLOADS["tiny"]
    large = small.with_(n_points=400_000)
    with mlgi.Context(0) as ctx:  # fresh: no eager splat has sized its work area
        inp_s, st_s, rsm_s, pts_s, gb_s = setup(ctx, small)
        print(f"[graph] capture {small.name} P={small.n_points} on a fresh context", flush=True)
        g = st_s.capture(rsm_s, pts_s, gb_s)
        g.launch()
        ctx.sync()
        first = outputs(st_s)
        rec, atlas = first[0], first[1]
        ref = oracle.ism_splat(inp_s.points, inp_s.bin_offsets, rec[:, 0:3], rec[:, 4:7], small.atlas_side, 50.0)
        assert np.array_equal(atlas, ref)

        # an eager frame with 20x the points grows the context's work area
        inp_l, st_l, rsm_l, pts_l, gb_l = setup(ctx, large)
        print(f"[graph] eager {large.name} P={large.n_points} on the same context", flush=True)
        st_l.process(rsm_l, pts_l, gb_l)
        ctx.sync()
        rec_l, atlas_l = st_l.vpl_processor.host_records(), st_l.ism.softrender.numpy()
        ref_l = oracle.ism_splat(inp_l.points, inp_l.bin_offsets, rec_l[:, 0:3], rec_l[:, 4:7], large.atlas_side,
                                 50.0)
        assert np.array_equal(atlas_l, ref_l)

        for rep in range(3):
            clobber(st_s)
            g.launch()
            ctx.sync()
            for name, x, y in zip(("vpl", "atlas", "ism", "image"), outputs(st_s), first):
                assert np.array_equal(x, y), (rep, name)
            st_l.process(rsm_l, pts_l, gb_l)  # eager work between replays
            ctx.sync()
        assert np.array_equal(st_l.ism.softrender.numpy(), atlas_l)

        # and the replays equal the eager frame
        clobber(st_s)
        st_s.process(rsm_s, pts_s, gb_s)
        ctx.sync()
        for name, x, y in zip(("vpl", "atlas", "ism", "image"), outputs(st_s), first):
            assert np.array_equal(x, y), name
        g.close()
