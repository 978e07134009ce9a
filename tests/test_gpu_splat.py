"""GPU: the ISM-major splat (candidate masks, one LDS tile per ISM, spill
list) against the oracle's splat and the point-major splat, bit for bit.

The points are placed so that projections land exactly on a tile's far
border, which belongs to the neighbouring ISM: every VPL sits at the origin
with normal +z (lookAtRH gives s = -x, u = +y exactly), and points on the -x
axis (normal +x) or the +y axis (normal -y) lie in the VPL's tangent plane,
so their paraboloid coordinate is exactly 1 and ivec2 truncation puts them on
the first column / row of the next tile. A few hundred such points exercise
the spill list; several thousand overflow it, and the point-major splat then
reruns over the tiles. Random points around them splat inside the tiles.
"""
import numpy as np
import pytest

import mlgi
import oracle

pytestmark = pytest.mark.gpu

N_VPL = 64        # ISM grid 8 x 8
ATLAS = 256       # ISM side 32
Z_FAR = 50.0


def unorm4x8(x, y, z, w):
    b = [int(np.clip(np.rint(v * 255.0), 0, 255)) for v in (x, y, z, w)]
    return np.uint32(b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24)).view(np.float32)


def point_set(n_axis, n_random, seed):
    rng = np.random.default_rng(seed)
    pts = []
    for k in range(n_axis):
        t = float(1 + k % 40)           # exact squares: dist == t
        rb = (k * 37) % 256 / 255.0
        if k % 2 == 0:                  # -x axis, normal +x
            pts.append([-t, 0.0, 0.0, unorm4x8(1.0, 0.5, 0.5, rb)])
        else:                           # +y axis, normal -y
            pts.append([0.0, t, 0.0, unorm4x8(0.5, 0.0, 0.5, rb)])
    for _ in range(n_random):
        p = rng.uniform(-5.0, 5.0, 3)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        pts.append([p[0], p[1], p[2], unorm4x8(*(n * 0.5 + 0.5), rng.uniform())])
    pts = np.array(pts, np.float32)
    # random bins, points stored bin by bin
    bins = rng.integers(0, N_VPL, len(pts))
    order = np.argsort(bins, kind="stable")
    pts, bins = pts[order], bins[order]
    off = np.zeros(N_VPL + 1, np.uint32)
    np.add.at(off, bins + 1, 1)
    return pts, np.cumsum(off).astype(np.uint32)


def vpl_records():
    n = np.array([0.0, 0.0, 1.0], np.float32)
    s, u = oracle.look_at_rh(n, np.float32)
    rec = np.zeros((N_VPL, mlgi.VPL_RECORD_FLOATS), np.float32)
    rec[:, 4:7] = n
    rec[:, 8:11] = s
    rec[:, 12:15] = u
    return rec


def run(ctx, pts, off, splat_block):
    vp = mlgi.VPLProcessor(ctx, N_VPL)
    vp.records.copy_from(vpl_records())
    ism = mlgi.ImperfectShadowmap(ctx, ATLAS, N_VPL)
    p = mlgi.default_params(atlas_side=ATLAS, z_far=Z_FAR, splat_block=splat_block)
    ism.render(mlgi.points_desc(ctx.upload(pts), ctx.upload(off)), vp, p)
    ctx.sync()
    stats = ism.spill_stats() if splat_block == 0 else None
    return ism.softrender.numpy(), stats


@pytest.mark.parametrize("n_axis,overflow", [(600, False), (6000, True)])
def test_spills_match_oracle(ctx, n_axis, overflow):
    print(f"[splat] spills n_axis={n_axis}", flush=True)
    pts, off = point_set(n_axis, 4000, seed=n_axis)
    rec = vpl_records()
    ref = oracle.ism_splat(pts, off, rec[:, 0:3], rec[:, 4:7], ATLAS, Z_FAR)
    tiled, (spills, overflowed) = run(ctx, pts, off, 0)
    point_major, _ = run(ctx, pts, off, 256)
    print(f"[splat] spills={spills} overflowed={overflowed}", flush=True)
    assert spills > 0
    assert overflowed == overflow
    assert np.array_equal(point_major, ref)
    assert np.array_equal(tiled, ref)


@pytest.mark.parametrize("name", ["tiny", "small"])
def test_workload_tiled_equals_point_major(ctx, name):
    from mlgi import workloads as W
    wl = W.WORKLOADS[name]
    inp = W.build_inputs(wl)
    p = wl.params()
    st = mlgi.GIStage(ctx, p, wl.n_vpl, wl.width, wl.height)
    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    st.vpl_processor.process(rsm, p.light_intensity, bool(p.shuffle_lights), p)
    st.params.splat_block = 0
    st.ism.render(pts, st.vpl_processor, st.params)
    ctx.sync()
    tiled = st.ism.softrender.numpy()
    spills, overflowed = st.ism.spill_stats()
    print(f"[splat] {name}: spills={spills} overflowed={overflowed}", flush=True)
    st.params.splat_block = 256
    st.ism.render(pts, st.vpl_processor, st.params)
    ctx.sync()
    assert np.array_equal(st.ism.softrender.numpy(), tiled)
    rec = st.vpl_processor.host_records()
    ref = oracle.ism_splat(inp.points, inp.bin_offsets, rec[:, 0:3], rec[:, 4:7], wl.atlas_side, p.z_far)
    assert np.array_equal(tiled, ref)


@pytest.mark.parametrize("n_vpl,atlas", [(1, 64), (4, 16), (9, 64), (16, 64), (64, 128), (64, 16), (256, 64)])
def test_small_vpl_counts_and_ism_sides(ctx, n_vpl, atlas):
    """Candidate lists that wrap ((b + i) mod N repeats VPLs when N < 16),
    ISM sides 2 (point-major only), 4, 8, 16, 64; random VPLs and points."""
    print(f"[splat] N={n_vpl} atlas={atlas}", flush=True)
    rng = np.random.default_rng(n_vpl * 1000 + atlas)
    rec = np.zeros((n_vpl, mlgi.VPL_RECORD_FLOATS), np.float32)
    rec[:, 0:3] = rng.uniform(-2, 2, (n_vpl, 3))
    nrm = rng.normal(size=(n_vpl, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    rec[:, 4:7] = nrm
    for v in range(n_vpl):
        s, u = oracle.look_at_rh(rec[v, 4:7], np.float32)
        rec[v, 8:11], rec[v, 12:15] = s, u
    n_pts = 3000
    pts = np.zeros((n_pts, 4), np.float32)
    pts[:, :3] = rng.uniform(-3, 3, (n_pts, 3))
    for k in range(n_pts):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        pts[k, 3] = unorm4x8(*(n * 0.5 + 0.5), rng.uniform())
    bins = np.sort(rng.integers(0, n_vpl, n_pts))
    off = np.zeros(n_vpl + 1, np.uint32)
    np.add.at(off, bins + 1, 1)
    off = np.cumsum(off).astype(np.uint32)
    ref = oracle.ism_splat(pts, off, rec[:, 0:3], rec[:, 4:7], atlas, Z_FAR)
    outs = []
    for blk in (0, 256):
        vp = mlgi.VPLProcessor(ctx, n_vpl)
        vp.records.copy_from(rec)
        ism = mlgi.ImperfectShadowmap(ctx, atlas, n_vpl)
        p = mlgi.default_params(atlas_side=atlas, z_far=Z_FAR, splat_block=blk)
        ism.render(mlgi.points_desc(ctx.upload(pts), ctx.upload(off)), vp, p)
        ctx.sync()
        outs.append(ism.softrender.numpy())
    assert (ref != 0xFFFFFFFF).any()
    assert np.array_equal(outs[1], ref)
    assert np.array_equal(outs[0], ref)


def test_forced_drain_of_a_long_sparse_bin(ctx):
    """One VPL (ISM side 64) and one bin of 20,000 points of which about 2%
    face it: the ISM-major splat's queue then carries fewer than 64 entries
    from one 128-point step to the next, so the forced drain every 64 steps
    (which keeps the 16-bit queue codes within their bits) runs twice, at
    points 8,192 and 16,384. Bit-exact against the oracle, with the
    point-major splat alongside."""
    rng = np.random.default_rng(77)
    n = 20_000
    rec = np.zeros((1, mlgi.VPL_RECORD_FLOATS), np.float32)
    nrm = np.array([0.0, 0.0, 1.0], np.float32)
    rec[0, 4:7] = nrm
    rec[0, 8:11], rec[0, 12:15] = oracle.look_at_rh(nrm, np.float32)
    lit = rng.random(n) < 0.02
    pts = np.zeros((n, 4), np.float32)
    pts[:, 0:2] = rng.uniform(-4.0, 4.0, (n, 2))
    pts[:, 2] = np.where(lit, rng.uniform(0.5, 4.0, n), rng.uniform(-4.0, -0.5, n))  # z < 0: behind the VPL
    for k in range(n):
        d = -pts[k, :3] / np.linalg.norm(pts[k, :3])  # facing the VPL
        pn = d if lit[k] else -d
        pts[k, 3] = unorm4x8(*(pn * 0.5 + 0.5), rng.uniform())
    off = np.array([0, n], np.uint32)
    ref = oracle.ism_splat(pts, off, rec[:, 0:3], rec[:, 4:7], 64, Z_FAR)
    assert 200 < int((ref != 0xFFFFFFFF).sum()) < 2000  # sparse
    for blk in (0, 256):
        vp = mlgi.VPLProcessor(ctx, 1)
        vp.records.copy_from(rec)
        ism = mlgi.ImperfectShadowmap(ctx, 64, 1)
        p = mlgi.default_params(atlas_side=64, z_far=Z_FAR, splat_block=blk)
        ism.render(mlgi.points_desc(ctx.upload(pts), ctx.upload(off)), vp, p)
        ctx.sync()
        assert np.array_equal(ism.softrender.numpy(), ref), blk


def test_division_form_matches_ieee(ctx):
    # the projection's six divisions are quotients from refined reciprocals
    # of their three divisors (recip_rn, div_rn in mlgi_device.hpp); the
    # device check compares both with IEEE division: every float significand
    # at eleven exponents, four numerators per divisor, two numerator seeds
    for seed in (0, 12345):
        print(f"[division check] seed={seed}", flush=True)
        rec, quo, n = ctx.check_division(seed)
        assert n == (1 << 23) * 11 * 4
        assert rec == 0 and quo == 0, (rec, quo)


def test_points_at_the_vpl_position(ctx):
    # points exactly at the VPLs' position: the projection divides 0 by a
    # distance of 0 (NaN under IEEE division; NaN from the reciprocal form),
    # and the splat must be dropped as the oracle drops it (its texel
    # conversion of NaN lands outside the atlas), in both splat forms
    print("[splat] points at the VPL position", flush=True)
    pts, off = point_set(64, 2000, seed=7)
    rng = np.random.default_rng(8)
    extra = []
    for k in range(32):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        extra.append([0.0, 0.0, 0.0, unorm4x8(*(n * 0.5 + 0.5), rng.uniform())])
    bins = np.repeat(np.arange(N_VPL), np.diff(off))
    bins = np.concatenate([bins, rng.integers(0, N_VPL, len(extra))])
    allp = np.concatenate([pts, np.array(extra, np.float32)])
    order = np.argsort(bins, kind="stable")
    allp, bins = allp[order], bins[order]
    off2 = np.zeros(N_VPL + 1, np.uint32)
    np.add.at(off2, bins + 1, 1)
    off2 = np.cumsum(off2).astype(np.uint32)
    rec = vpl_records()
    ref = oracle.ism_splat(allp, off2, rec[:, 0:3], rec[:, 4:7], ATLAS, Z_FAR)
    base = oracle.ism_splat(pts, off, rec[:, 0:3], rec[:, 4:7], ATLAS, Z_FAR)
    # the extra points change nothing in the oracle's atlas either
    assert np.array_equal(ref, base)
    tiled, _ = run(ctx, allp, off2, 0)
    point_major, _ = run(ctx, allp, off2, 256)
    assert np.array_equal(tiled, ref)
    assert np.array_equal(point_major, ref)
