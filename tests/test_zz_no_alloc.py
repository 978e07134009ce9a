"""GPU: mlgi_ism_splat_noalloc, the splat that allocates nothing.

Called through the C ABI on two consecutive inputs, the second smaller than
the first, on one context: each atlas must equal, bit for bit, the atlas of
mlgi_ism_splat (the entry point it is a variant of) on the same input, and
the FP32 oracle's. The static work area keeps the previous call's contents,
so the second call also shows that nothing stale leaks into a smaller
frame. A point set above MLGI_NOALLOC_MAX_POINTS is rejected before any
launch.
"""
import ctypes as C

import numpy as np
import pytest

import mlgi
import oracle
from mlgi import _native as N
from mlgi import workloads as W

pytestmark = pytest.mark.gpu

lib = N.lib


def splat_both(ctx, wl):
    """(atlas of mlgi_ism_splat, atlas of mlgi_ism_splat_noalloc, oracle atlas) of one workload."""
    inp = W.build_inputs(wl)
    p = wl.params()
    vp = mlgi.VPLProcessor(ctx, wl.n_vpl)
    rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal), ctx.upload(inp.rsm_depth),
                        inp.inv_bvp)
    vp.process(rsm, p.light_intensity, True, p)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    atlases = []
    for fn in (lib.mlgi_ism_splat, lib.mlgi_ism_splat_noalloc):
        buf = ctx.empty((wl.atlas_side, wl.atlas_side), np.uint32)
        buf.copy_from(np.zeros((wl.atlas_side, wl.atlas_side), np.uint32))  # the splat writes every texel
        a = N.Atlas(buf.ptr, wl.atlas_side, 0)
        print(f"[noalloc] {fn.__name__} {wl.name} P={wl.n_points} N={wl.n_vpl} atlas {wl.atlas_side}", flush=True)
        ctx.check(fn(ctx.h, C.byref(pts), C.byref(vp.desc()), C.byref(p), C.byref(a)), fn.__name__)
        ctx.sync()
        atlases.append(buf.numpy())
    rec = vp.host_records()
    ref = oracle.ism_splat(inp.points, inp.bin_offsets, rec[:, 0:3], rec[:, 4:7], wl.atlas_side, p.z_far)
    return atlases[0], atlases[1], ref


def test_noalloc_splat_equals_the_allocating_one_on_two_calls(ctx):
    first = W.WORKLOADS["small"]  # 262,144 points, 1,024 ISMs of 64^2 (16-bit queue codes)
    second = W.WORKLOADS["tiny"].with_(n_points=12_000)  # fewer points, 256 ISMs of 32^2
    assert second.n_points < first.n_points
    for wl in (first, second):
        a, b, ref = splat_both(ctx, wl)
        assert (ref != 0xFFFFFFFF).any()
        assert np.array_equal(b, a), f"{wl.name}: {int((a != b).sum())} texels differ"
        assert np.array_equal(b, ref)


def test_noalloc_splat_rejects_too_many_points(ctx):
    wl = W.WORKLOADS["tiny"]
    inp = W.build_inputs(wl)
    p = wl.params()
    vp = mlgi.VPLProcessor(ctx, wl.n_vpl)
    pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
    pts.count = N.NOALLOC_MAX_POINTS + 1  # rejected by the host check, before any launch
    buf = ctx.empty((wl.atlas_side, wl.atlas_side), np.uint32)
    a = N.Atlas(buf.ptr, wl.atlas_side, 0)
    rc = lib.mlgi_ism_splat_noalloc(ctx.h, C.byref(pts), C.byref(vp.desc()), C.byref(p), C.byref(a))
    assert rc == N.MLGI_ERR_INVALID_ARGUMENT
    assert b"maximum" in lib.mlgi_last_error(ctx.h)
