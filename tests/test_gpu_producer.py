"""GPU: device G-buffer / RSM producer equals the host renderer bit for bit
(same double-precision ray caster, -ffp-contract=off on both sides)."""
import numpy as np
import pytest

import mlgi
from mlgi import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_random,size", [(500, (96, 64)), (40_000, (128, 128))])
def test_device_renders_equal_host(ctx, n_random, size):
    scene = W._scene(1, n_random)
    dev = scene.upload(ctx)
    w, h = size
    cam = mlgi.Camera.default(w / h)
    d_h, n_h = scene.render_gbuffer(cam, w, h)
    dd, dn = ctx.empty((h, w), np.float32), ctx.empty((h, w, 3), np.float32)
    vp_inv = dev.render_gbuffer(cam, dd, dn)
    ctx.sync()
    assert np.array_equal(dd.numpy(), d_h) and np.array_equal(dn.numpy(), n_h)
    assert np.array_equal(vp_inv, cam.matrices()[2])
    for seed in (3, 7):
        light = mlgi.Camera.light(seed)
        dif_h, nrm_h, dep_h, inv_h = scene.render_rsm(light, 64, 64)
        a, b, c = ctx.empty((64, 64, 3), np.float32), ctx.empty((64, 64, 3), np.float32), ctx.empty((64, 64), np.float32)
        inv = dev.render_rsm(light, a, b, c)
        ctx.sync()
        assert np.array_equal(a.numpy(), dif_h) and np.array_equal(b.numpy(), nrm_h)
        assert np.array_equal(c.numpy(), dep_h) and np.array_equal(inv, inv_h)
        assert (dif_h > 0).any()
    dev.close()


def _sorted_bins(rec, off):
    """Records of each bin sorted by their bit patterns (bins are unordered sets)."""
    bits = rec.view(np.uint32)
    out = []
    for b in range(len(off) - 1):
        blk = bits[off[b]:off[b + 1]]
        out.append(blk[np.lexsort(blk.T[::-1])])
    return out


@pytest.mark.parametrize("n_random,n_points,n_bins", [(500, 20_000, 64), (40_000, 300_000, 1024), (4_000, 0, 16)])
def test_device_points_equal_host(ctx, n_random, n_points, n_bins):
    """Device point generation: bin offsets equal the host sampler's; every
    bin holds the same records (order within a bin is free); the splat of
    either set gives the same atlas."""
    import oracle

    print(f"[points] tris={n_random + 12} n={n_points} bins={n_bins}", flush=True)
    scene = W._scene(1, n_random)
    dev = scene.upload(ctx)
    rec_h, off_h = scene.sample_points(n_points, n_bins, n_bins, 11)
    rec_d, off_d = dev.sample_points(n_points, n_bins, n_bins, 11)
    off = off_d.numpy()
    assert np.array_equal(off, off_h)
    rec = rec_d.numpy()[:n_points]
    for a, b in zip(_sorted_bins(rec, off), _sorted_bins(rec_h, off_h)):
        assert np.array_equal(a, b)
    if n_points:
        # order-free consumer: the oracle splat of both sets, one VPL per bin
        rng = np.random.default_rng(5)
        pos = rng.uniform(0.5, 9.5, (n_bins, 3)).astype(np.float32)
        nrm = rng.normal(size=(n_bins, 3))
        nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
        side = mlgi.ism_indices_1d(n_bins) * 16
        a1 = oracle.ism_splat(rec, off, pos, nrm, side, 50.0)
        a2 = oracle.ism_splat(rec_h, off_h, pos, nrm, side, 50.0)
        assert np.array_equal(a1, a2)
    dev.close()
