"""Oracle deferred composite (deferredshading.frag:47-81) pinned by cases
whose result is fixed independently of the formula chain: tone-mapping and
gamma fixed points, Lambert at normal incidence, a fully lit and a fully
occluded VSM, the bilinear VSM tap at a texel centre, between texels and
outside the map (border), and the light's back hemisphere (w < 0)."""
import numpy as np

import oracle

W, H = 8, 4
G = float(np.float32(2.2))  # gamma: the GLSL float literal 2.2
IDENT = np.eye(4, dtype=np.float32).T.ravel()  # column-major identity


def blvp_xy():
    """Light transform mapping world (x, y) in [-1, 1] to VSM uv in [0, 1], w = 1."""
    m = np.zeros((4, 4), np.float32)
    m[0, 0] = m[1, 1] = 0.5
    m[0, 3] = m[1, 3] = 0.5
    m[2, 2] = 1.0
    m[3, 3] = 1.0
    return m.T.ravel()  # column-major


def base(depth=0.5):
    return dict(depth=np.full((H, W), depth, np.float32),
                normal=np.tile(np.float32([0.5, 0.5, 1.0]), (H, W, 1)),  # n = +z
                diffuse=np.ones((H, W, 3), np.float32), gi=np.zeros((H, W, 3), np.float32))


def run(inp, vsm, light_intensity=1.0, light_pos=(0.0, 0.0, 10.0), blvp=None, eye=(0.0, 0.0, 5.0), **kw):
    return oracle.deferred_composite(inp["depth"], inp["normal"], inp["diffuse"], inp["gi"], vsm, vp_inv=IDENT,
                                     biased_light_vp=blvp_xy() if blvp is None else blvp, eye=eye,
                                     light_pos=light_pos, inv_light_dir=(0.0, 0.0, 1.0),
                                     light_intensity=light_intensity, **kw)


def light_dist(inp):
    """Each pixel's distance to the light at (0, 0, 10): world = NDC of the pixel centre."""
    y, x = np.mgrid[0:H, 0:W]
    wx, wy = (x + 0.5) / W * 2 - 1, (y + 0.5) / H * 2 - 1
    wz = inp["depth"].astype(np.float64) * 2 - 1
    return np.sqrt(wx * wx + wy * wy + (wz - 10.0) ** 2)


def lit_vsm(inp):
    """Moments equal to each pixel's own light distance (nothing in front)."""
    vsm = np.zeros((H, W, 2), np.float32)
    d = light_dist(inp)
    vsm[..., 0], vsm[..., 1] = d, d * d
    return vsm


def test_ambient_tonemap_fixed_point():
    inp = base()
    inp["gi"][:] = 1.0  # gi * diffuse * occlusion * exposure = 1 -> Reinhard 1/2 -> (1/2)^(1/2.2)
    out = run(inp, np.zeros((H, W, 2), np.float32), light_intensity=0.0)
    assert np.allclose(out, 0.5 ** (1 / G), rtol=0, atol=1e-12)
    inp["gi"][:] = 3.0  # 3 -> 3/4
    out = run(inp, np.zeros((H, W, 2), np.float32), light_intensity=0.0, occlusion=np.ones((H, W), np.float32))
    assert np.allclose(out, 0.75 ** (1 / G), rtol=0, atol=1e-12)
    out = run(inp, np.zeros((H, W, 2), np.float32), light_intensity=0.0, occlusion=np.zeros((H, W), np.float32))
    assert np.all(out == 0.0)
    # diffuse is sRGB: 0.5 is 0.5^2.2 linear, so gi = 0.5^-2.2 gives the fixed point 1 again
    inp["diffuse"][:] = 0.5
    inp["gi"][:] = 0.5 ** -G  # stored as float32: the unit colour to ~6e-8
    out = run(inp, np.zeros((H, W, 2), np.float32), light_intensity=0.0)
    assert np.allclose(out, 0.5 ** (1 / G), rtol=0, atol=1e-7)
    # exposure scales before Reinhard: exposure 3 on a unit colour -> 3/4
    out = run(inp, np.zeros((H, W, 2), np.float32), light_intensity=0.0, exposure=3.0)
    assert np.allclose(out, 0.75 ** (1 / G), rtol=0, atol=1e-7)


def test_lambert_at_normal_incidence_lit_and_occluded():
    inp = base()
    vsm = lit_vsm(inp)
    # N = L = +z, shadow 1, diffuse 1: colour = intensity; intensity 1 -> 1/2 after Reinhard.
    # The half vector of L and V (both +z) is +z: specular needs a specular buffer (absent = 0).
    out = run(inp, vsm, light_intensity=1.0)
    assert np.allclose(out, 0.5 ** (1 / G), rtol=0, atol=1e-9)
    # occluder at half the distance, minimum variance: shadow -> linstep(0.1, 1, 0.001/(0.001+d^2)) = 0
    vsm_occ = vsm.copy()
    vsm_occ[..., 0] *= 0.5
    vsm_occ[..., 1] = vsm_occ[..., 0] ** 2
    assert np.all(run(inp, vsm_occ, light_intensity=1.0) == 0.0)
    # specular: 0.75 * spec * (N.H)^20 with N.H = 1 -> 0.75 for spec 1
    out = run(inp, vsm, light_intensity=0.0, specular=np.ones((H, W, 3), np.float32), eye=(0.0, 0.0, 1e8))
    assert np.allclose(out, (0.75 / 1.75) ** (1 / G), rtol=0, atol=1e-9)


def test_vsm_bilinear_taps_and_border():
    inp = base()
    d = light_dist(inp)
    # a VSM of the screen's size: pixel centres land on texel centres -> exact taps
    vsm = lit_vsm(inp)
    vsm[:, ::2, 0] *= 0.5  # every other column occluded
    vsm[:, ::2, 1] = vsm[:, ::2, 0] ** 2
    out = run(inp, vsm)
    assert np.all(out[:, ::2] == 0.0)
    assert np.allclose(out[:, 1::2], 0.5 ** (1 / G), atol=1e-9)
    # a 2x-wide VSM: pixel centres fall between two texels (weights 1/2 each)
    vsm2 = np.zeros((H, 2 * W, 2), np.float32)
    vsm2[:, 0::2, 0] = vsm2[:, 1::2, 0] = d
    vsm2[:, 0::2, 1] = d * d  # the pair's mean: E[d^2] - E[d]^2 = 8 at mean d
    vsm2[:, 1::2, 1] = d * d + 16.0
    out2 = run(inp, vsm2)
    assert np.allclose(out2, 0.5 ** (1 / G), atol=1e-9)  # p = smoothstep(..., d) = 1 -> lit
    # a light transform that puts every pixel outside the map: border (0, 0) -> compare vs 0 -> dark
    far = blvp_xy().reshape(4, 4).T.copy()
    far[0, 3] = 5.0
    assert np.all(run(inp, vsm, blvp=far.T.ravel()) == 0.0)
    # w < 0 (behind the light's projection): step(0, sign(w)) = 0
    neg = blvp_xy().reshape(4, 4).T.copy()
    neg *= -1.0
    assert np.all(run(inp, vsm, blvp=neg.T.ravel()) == 0.0)
