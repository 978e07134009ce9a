"""GPU: deferred composite (deferredshading.frag:47-81) against the FP64
oracle. FP32 kernel. Where every VSM tap lands exactly on a texel centre in
FP32 too (power-of-two sizes, VSM of the screen's size), the tolerance is
2e-5 absolute on outputs in [0, 1]. Otherwise rel-L2 <= 1e-5: an FP32 tap
coordinate a few 1e-6 off a texel centre blends a neighbour's moments in by
that fraction, and inside the 0.001-wide smoothstep window of the VSM that
moves a pixel by up to ~1e-4 (the reference's own FP32 texture coordinates
do the same)."""
import numpy as np
import pytest

import mlgi
import oracle

pytestmark = pytest.mark.gpu


def inputs(w, h, vsm_scale, seed):
    rng = np.random.default_rng(seed)
    depth = rng.uniform(0.2, 0.8, (h, w)).astype(np.float32)
    n = rng.normal(size=(h, w, 3))
    n[..., 2] = np.abs(n[..., 2]) + 0.5
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    normal = (n * 0.5 + 0.5).astype(np.float32)
    diffuse = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    specular = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    gi = rng.uniform(0, 2, (h, w, 3)).astype(np.float32)
    occ = rng.uniform(0.3, 1, (h, w)).astype(np.float32)
    vw, vh = int(w * vsm_scale), int(h * vsm_scale)
    # light at (0.3, -0.2, 6); world = pixel-centre NDC (identity inverse(P*V))
    y, x = np.mgrid[0:vh, 0:vw]
    wx, wy = (x + 0.5) / vw * 2 - 1, (y + 0.5) / vh * 2 - 1
    if vsm_scale == 1:
        wz = depth.astype(np.float64) * 2 - 1
    else:
        wz = rng.uniform(-0.6, 0.6, (vh, vw))
    dist = np.sqrt((wx - 0.3) ** 2 + (wy + 0.2) ** 2 + (wz - 6.0) ** 2)
    occluded = rng.uniform(size=dist.shape) < 0.4
    mx = np.where(occluded, dist * rng.uniform(0.3, 0.95, dist.shape), dist)
    var = rng.uniform(0.0, 0.05, dist.shape)
    vsm = np.stack([mx, mx * mx + var], -1).astype(np.float32)
    return dict(depth=depth, normal=normal, diffuse=diffuse, specular=specular, gi=gi, occlusion=occ, vsm=vsm)


def params():
    blvp = np.zeros((4, 4), np.float32)
    blvp[0, 0] = blvp[1, 1] = 0.5
    blvp[0, 3] = blvp[1, 3] = 0.5
    blvp[2, 2] = blvp[3, 3] = 1.0
    ld = np.array([-0.3, 0.2, -1.0])
    return dict(vp_inv=np.eye(4, dtype=np.float32).T.ravel(), biased_light_vp=blvp.T.ravel(), eye=(0.1, 0.2, 4.0),
                light_pos=(0.3, -0.2, 6.0), inv_light_dir=-ld / np.linalg.norm(ld), light_intensity=1.7,
                exposure=1.3)


@pytest.mark.parametrize("w,h,scale,with_optional", [(256, 128, 1, True), (128, 64, 1, False), (97, 61, 1, True),
                                                     (160, 90, 1.5, True)])
def test_composite_matches_oracle(ctx, w, h, scale, with_optional):
    print(f"[composite] {w}x{h} vsm x{scale} optional={with_optional}", flush=True)
    i = inputs(w, h, scale, seed=w)
    p = params()
    spec = i["specular"] if with_optional else None
    occ = i["occlusion"] if with_optional else None
    ref = oracle.deferred_composite(i["depth"], i["normal"], i["diffuse"], i["gi"], i["vsm"], specular=spec,
                                    occlusion=occ, **p)
    dev = {k: ctx.upload(v) for k, v in i.items()}
    out = ctx.empty((h, w, 3), np.float32)
    mlgi.deferred_composite(ctx, dev["depth"], dev["normal"], dev["diffuse"], dev["gi"], dev["vsm"], out,
                            specular=dev["specular"] if with_optional else None,
                            occlusion=dev["occlusion"] if with_optional else None, **p)
    ctx.sync()
    got = out.numpy().astype(np.float64)
    assert np.isfinite(got).all()
    err = np.abs(got - ref)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"[composite] max abs {err.max():.3g} rel-L2 {rel:.3g}", flush=True)
    pow2 = (w & (w - 1)) == 0 and (h & (h - 1)) == 0
    if scale == 1 and pow2:
        assert err.max() <= 2e-5
    else:
        assert rel <= 1e-5
