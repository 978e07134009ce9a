"""GPU: device G-buffer / RSM producer equals the host renderer bit for bit
(same double-precision ray caster, -ffp-contract=off on both sides).

Images of at most 512^2 texels run the cooperative form (four lanes per ray),
larger ones one lane per ray; the 720x400 G-buffer (288,000 texels) covers the
latter, the RSMs and smaller G-buffers the former."""
import numpy as np
import pytest

import mlgi
from mlgi import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_random,size", [(500, (96, 64)), (40_000, (128, 128)), (40_000, (720, 400))])
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


def test_device_surface_equals_host(ctx):
    """Albedo and VSM moments of a view: device == host, bit for bit."""
    scene = W._scene(1, 4_000)
    dev = scene.upload(ctx)
    for cam, (w, h) in [(mlgi.Camera.default(96 / 64), (96, 64)), (mlgi.Camera.light(3), (64, 64))]:
        alb_h, mom_h = scene.render_surface(cam, w, h)
        alb, mom = ctx.empty((h, w, 3), np.float32), ctx.empty((h, w, 2), np.float32)
        dev.render_surface(cam, alb, mom)
        ctx.sync()
        assert np.array_equal(alb.numpy(), alb_h) and np.array_equal(mom.numpy(), mom_h)
        assert (alb_h > 0).all() and (mom_h[..., 0] > 0).all()  # closed room: every ray hits
    dev.close()


def test_composited_frame_of_the_scene(ctx):
    """A whole frame on the device: G-buffer, albedo and the light's VSM from
    the ray caster, a GI buffer, the deferred composite; against the FP64
    oracle on the same buffers (rel-L2; the VSM taps land anywhere, so a few
    pixels sit inside the 0.001-wide smoothstep window where FP32 and FP64
    light distances differ visibly)."""
    import oracle

    w, h, vs = 160, 96, 256
    scene = W._scene(1, 40_000)
    dev = scene.upload(ctx)
    cam = mlgi.Camera.default(w / h)
    light = mlgi.Camera.light(7)
    depth, normal = ctx.empty((h, w), np.float32), ctx.empty((h, w, 3), np.float32)
    vp_inv = dev.render_gbuffer(cam, depth, normal)
    albedo = ctx.empty((h, w, 3), np.float32)
    dev.render_surface(cam, albedo=albedo)
    vsm = ctx.empty((vs, vs, 2), np.float32)
    dev.render_surface(light, moments=vsm)
    dif, nrm, dep = ctx.empty((vs, vs, 3), np.float32), ctx.empty((vs, vs, 3), np.float32), ctx.empty((vs, vs), np.float32)
    inv_bvp = dev.render_rsm(light, dif, nrm, dep)
    blvp = np.linalg.inv(inv_bvp.reshape(4, 4).T.astype(np.float64)).T.ravel().astype(np.float32)
    y, x = np.mgrid[0:h, 0:w]
    gi_h = np.stack([x / w, y / h, 0.5 + 0.0 * x], -1).astype(np.float32)
    gi = ctx.upload(gi_h)
    out = ctx.empty((h, w, 3), np.float32)
    eye = np.array(light.eye)
    ldir = np.array(light.center) - eye
    p = dict(vp_inv=vp_inv, biased_light_vp=blvp, eye=cam.eye, light_pos=light.eye,
             inv_light_dir=-ldir / np.linalg.norm(ldir), light_intensity=2.0, exposure=1.0)
    mlgi.deferred_composite(ctx, depth, normal, albedo, gi, vsm, out, **p)
    ctx.sync()
    got = out.numpy().astype(np.float64)
    ref = oracle.deferred_composite(depth.numpy(), normal.numpy(), albedo.numpy(), gi_h, vsm.numpy(), **p)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    lit = np.mean(ref.mean(-1) > np.median(ref.mean(-1)) * 1.5)
    print(f"[composite] scene frame rel-L2 {rel:.3g}, max abs {np.abs(got - ref).max():.3g}, bright share {lit:.2f}",
          flush=True)
    assert np.isfinite(got).all() and got.min() >= 0 and got.max() <= 1
    assert rel <= 1e-3
    dev.close()
