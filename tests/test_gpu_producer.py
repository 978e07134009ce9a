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
