"""GPU: the gather's VPL culling is exact.

The culling replaces the reference's clustered light lists
(light_lists.comp:104-119, final_gathering.comp:81-98), whose full form is
the brute-force sum over every VPL of the pixel's interleave class. With
culling switched off (mlgi_params.disable_culling) the gather evaluates
every VPL of the class in the same order with the same arithmetic, so the
culled result must equal it bit for bit: a culled VPL may only be one whose
contribution is exactly +0 (no lit pair, or zero flux).

Checked in both precisions, through the single-frame and the batched
kernels, on the tiny, 512^2 and 1080p workloads, and on a synthetic plane
with VPLs on, just above and just below the pixels' tangent plane and with
nearly tangent normals, where the culling tests' margins decide.
"""
import numpy as np
import pytest

import mlgi
from mlgi import workloads as W

pytestmark = pytest.mark.gpu

PRECISIONS = [("fp32", mlgi.PRECISION_FP32), ("mixed", mlgi.PRECISION_MIXED)]


def gather(ctx, st, gb, precision, disable_culling):
    st.params.precision = precision
    st.params.disable_culling = disable_culling
    try:
        st.gi_buffer.copy_from(np.full((st.height, st.width, 3), -1.0, np.float32))
        st.compute_final_gathering(gb)
        ctx.sync()
        return st.gi_buffer.numpy()
    finally:
        st.params.precision = mlgi.PRECISION_FP32
        st.params.disable_culling = 0


def batched(ctx, st, rsm, pts, gb, precision, disable_culling):
    st.params.precision = precision
    st.params.disable_culling = disable_culling
    try:
        b = mlgi.Batch(ctx, [(st, rsm, pts, gb)])
        st.gi_buffer.copy_from(np.full((st.height, st.width, 3), -1.0, np.float32))
        b.run()
        ctx.sync()
        b.close()
        return st.gi_buffer.numpy()
    finally:
        st.params.precision = mlgi.PRECISION_FP32
        st.params.disable_culling = 0


@pytest.fixture(scope="module", params=["tiny", "small", "1080p"])
def frame(request, ctx):
    wl = W.WORKLOADS[request.param]
    inp = W.build_inputs(wl)
    p = wl.params()
    st = mlgi.GIStage(ctx, p, wl.n_vpl, wl.width, wl.height)
    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
    print(f"[cull] {wl.name} {wl.width}x{wl.height} N={wl.n_vpl} I={wl.interleave}", flush=True)
    st.process(rsm, pts, gb)
    ctx.sync()
    return wl, st, rsm, pts, gb


@pytest.mark.parametrize("name,prec", PRECISIONS)
def test_culled_gather_equals_brute_force(ctx, frame, name, prec):
    wl, st, rsm, pts, gb = frame
    culled = gather(ctx, st, gb, prec, 0)
    full = gather(ctx, st, gb, prec, 1)
    assert (culled > 0).any()
    bad = np.argwhere(culled != full)
    print(f"[cull] {wl.name} {name} single-frame: {len(bad)} differing values")
    assert len(bad) == 0, bad[:8]


@pytest.mark.parametrize("name,prec", PRECISIONS)
def test_culled_batched_gather_equals_brute_force(ctx, frame, name, prec):
    wl, st, rsm, pts, gb = frame
    culled = batched(ctx, st, rsm, pts, gb, prec, 0)
    full = batched(ctx, st, rsm, pts, gb, prec, 1)
    assert (culled > 0).any()
    bad = np.argwhere(culled != full)
    print(f"[cull] {wl.name} {name} batched: {len(bad)} differing values")
    assert len(bad) == 0, bad[:8]


def plane_case(ctx, seed):
    """A 128 x 96 G-buffer on the plane y = 0 (normals +y) and 4,096 VPLs
    on, just above and just below that plane, with normals from straight
    down to nearly tangent, so the culling margins decide for many pairs."""
    rng = np.random.default_rng(seed)
    w, h, n, il = 128, 96, 4096, 4
    m = np.zeros(16, np.float32)  # world = (3 + 2 ndc_x, 0, 4 + 1.5 ndc_y); window depth unused
    m[0], m[6], m[12], m[14], m[15] = 2.0, 1.5, 3.0, 4.0, 1.0
    depth = np.full((h, w), 0.5, np.float32)
    normal = np.broadcast_to(np.array([0.5, 1.0, 0.5], np.float32), (h, w, 3)).copy()
    heights = np.array([0.0, 1e-7, -1e-7, 1e-6, -1e-6, 1e-5, -1e-5, 1e-3, -1e-3, 0.05, 0.5], np.float32)
    pos = np.column_stack([rng.uniform(0.5, 5.5, n), rng.choice(heights, n), rng.uniform(2.0, 6.0, n)])
    pos = pos.astype(np.float32)
    tilt = rng.choice([-1.0, -0.5, -1e-2, -1e-4, -1e-6, 0.0, 1e-6, 1e-4], n)
    phi = rng.uniform(0, 2 * np.pi, n)
    nrm = np.column_stack([np.cos(phi), tilt, np.sin(phi)])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = nrm.astype(np.float32)
    rec = np.zeros((n, mlgi.VPL_RECORD_FLOATS), np.float32)
    rec[:, 0:3] = pos
    rec[:, 3] = nrm[:, 0] * pos[:, 0] + nrm[:, 1] * pos[:, 1] + nrm[:, 2] * pos[:, 2]  # float32, unfused
    rec[:, 4:7] = nrm
    flux = rng.uniform(0.1, 3.0, (n, 3)).astype(np.float32)
    flux[rng.random(n) < 0.1] = 0.0  # zero-flux VPLs (dropped by the culling)
    rec[:, 16:19] = flux
    mlgi.fill_flux_halves(rec)
    p = mlgi.default_params(interleave=il, use_interleaving=1, enable_shadowing=0, atlas_side=1024,
                            vpl_end_index=n, z_far=50.0, tile_w=16, tile_h=16, vpl_clamping_value=10.0)
    st = mlgi.GIStage(ctx, p, n, w, h)
    st.vpl_processor.records.copy_from(rec)
    gb = mlgi.gbuffer_desc(ctx.upload(depth), ctx.upload(normal), m)
    return st, gb


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("name,prec", PRECISIONS)
def test_culling_exact_at_the_tangent_planes(ctx, seed, name, prec):
    st, gb = plane_case(ctx, seed)
    culled = gather(ctx, st, gb, prec, 0)
    full = gather(ctx, st, gb, prec, 1)
    assert (culled > 0).any()
    bad = np.argwhere(culled != full)
    print(f"[cull] plane seed {seed} {name}: {len(bad)} differing values")
    assert len(bad) == 0, bad[:8]
