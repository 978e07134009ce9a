"""The host ray caster (BVH) against brute-force intersection (CPU only).

The G-buffer producer (RasterizationStage's role, RasterizationStage.cpp:131-277)
is the library's BVH ray caster; the device build runs the same code and is
compared with this host build bit for bit on the GPU. Here the host build is
checked independently: for sampled pixels of three scenes (500, 4,000 and
40,000 random triangles in the 10^3 room; camera and light views), the
pixel-centre ray is built in FP64 from the camera with the standard GL
lookAt and perspective matrices, intersected with every triangle by brute
force (Moller-Trumbore), and the nearest hit's window depth and face normal
(turned towards the eye) must be the rendered texel's. Room-wall hits are
also checked analytically: the point lies on the wall plane and the normal
is the wall's axis. The library hands out FP32 copies of its FP64
vertices, so rays that pass within 1e-5 (barycentric) of an edge of the hit
triangle or of a triangle in front of it, or that hit two triangles at
nearly one distance, are skipped: there that rounding decides.
"""
import numpy as np
import pytest

import mlgi

ROOM = 10.0


def brute_nearest(o, d, v0, e1, e2, band=1e-5):
    """Nearest hit of rays o + t d (n x 3) against all triangles
    (Moller-Trumbore, FP64). Returns (t, triangle, barycentric margin of that
    hit, second-nearest t, nearest t over hits widened by `band` beyond the
    edges); t = inf where nothing is hit."""
    n = len(o)
    best_t = np.full(n, np.inf)
    best_i = np.full(n, -1)
    best_m = np.zeros(n)
    second_t = np.full(n, np.inf)
    wide_t = np.full(n, np.inf)
    rows = np.arange(n)
    for s in range(0, len(v0), 8192):
        a0, b1, b2 = v0[s:s + 8192], e1[s:s + 8192], e2[s:s + 8192]
        p = np.cross(d[:, None, :], b2[None, :, :])
        det = np.einsum('tk,rtk->rt', b1, p)
        with np.errstate(divide='ignore', invalid='ignore'):
            inv = 1.0 / det
            tv = o[:, None, :] - a0[None, :, :]
            u = np.einsum('rtk,rtk->rt', tv, p) * inv
            q = np.cross(tv, b1[None, :, :])
            v = np.einsum('rk,rtk->rt', d, q) * inv
            t = np.einsum('tk,rtk->rt', b2, q) * inv
        m = np.minimum(np.minimum(u, v), 1 - u - v)
        front = (np.abs(det) > 1e-14) & (t > 1e-9)
        wide_t = np.minimum(wide_t, np.where(front & (m >= -band), t, np.inf).min(axis=1))
        t = np.where(front & (m >= 0), t, np.inf)
        k = np.argmin(t, axis=1)
        tk = t[rows, k]
        t_two = np.partition(t, 1, axis=1)[:, 1] if t.shape[1] > 1 else np.full(n, np.inf)
        better = tk < best_t
        second_t = np.where(better, np.minimum(best_t, t_two), np.minimum(second_t, tk))
        best_m = np.where(better, m[rows, k], best_m)
        best_i = np.where(better, k + s, best_i)
        best_t = np.minimum(best_t, tk)
    return best_t, best_i, best_m, second_t, wide_t


def view_projection(cam):
    """P V of a camera in FP64: GL lookAt (eye, center, up) and perspective
    (fovy, aspect, near, far); the camera's fields are binary32."""
    eye, ctr, up = (np.array(v, np.float64) for v in (cam.eye, cam.center, cam.up))
    f = ctr - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[:3, 3] = [-s @ eye, -u @ eye, f @ eye]
    fov, asp, n, fa = (float(np.float32(v)) for v in (cam.fovy_deg, cam.aspect, cam.z_near, cam.z_far))
    th = np.tan(np.radians(fov) / 2)
    proj = np.zeros((4, 4))
    proj[0, 0], proj[1, 1] = 1 / (asp * th), 1 / th
    proj[2, 2], proj[2, 3], proj[3, 2] = -(fa + n) / (fa - n), -2 * fa * n / (fa - n), -1.0
    return proj @ view


def check_view(scene, cam, w, h, n_rays, seed):
    depth, normal = scene.render_gbuffer(cam, w, h)
    vp = view_projection(cam)
    inv = np.linalg.inv(vp)
    rng = np.random.default_rng(seed)
    ys, xs = rng.integers(0, h, n_rays), rng.integers(0, w, n_rays)
    nx, ny = (xs + 0.5) / w * 2 - 1, (ys + 0.5) / h * 2 - 1
    ones = np.ones(n_rays)
    a = np.stack([nx, ny, -ones, ones], 1) @ inv.T  # near-plane and far-plane points
    b = np.stack([nx, ny, ones, ones], 1) @ inv.T
    o = a[:, :3] / a[:, 3:4]
    dvec = b[:, :3] / b[:, 3:4] - o
    verts, tri_n, _ = scene.triangles()
    v0 = verts[:, 0].astype(np.float64)
    e1 = (verts[:, 1] - verts[:, 0]).astype(np.float64)
    e2 = (verts[:, 2] - verts[:, 0]).astype(np.float64)
    t, tri, margin, t2, t_wide = brute_nearest(o, dvec, v0, e1, e2)
    assert np.isfinite(t).all()  # the room is closed: every ray hits
    # skip rays where the FP32 copies of the vertices can decide the hit: the
    # hit within the edge band, another triangle within the band in front,
    # or two hits at nearly one distance
    clear = (margin > 1e-5) & (t_wide >= t * (1 - 1e-9)) & (t2 > t * (1 + 1e-9))
    assert clear.mean() > 0.98
    p = o + dvec * t[:, None]
    c = np.concatenate([p, ones[:, None]], 1) @ vp.T
    want_depth = 0.5 * c[:, 2] / c[:, 3] + 0.5
    got_depth = depth[ys, xs].astype(np.float64)
    # The triangles come back as FP32 copies of the library's FP64 vertices
    # (off by up to ~6e-7 in a 10-unit room): allow 2e-6 world units of
    # plane shift, in window depth d(window)/dz = f n / ((f - n) z^2) at the
    # hit's view depth z, plus the FP32 rounding of the stored depth.
    # A shift of a triangle's plane by e moves the hit along the ray by
    # e / |cos| at incidence angle theta, so the allowance grows at grazing
    # incidence (capped at 1e-3).
    z = c[:, 3]  # clip w = view-space distance in front of the eye
    n_, f_ = float(np.float32(cam.z_near)), float(np.float32(cam.z_far))
    tn = tri_n[tri].astype(np.float64)
    cos = np.abs(np.einsum('rk,rk->r', tn, dvec)) / np.linalg.norm(dvec, axis=1)
    world = np.minimum(2e-6 / np.maximum(cos, 1e-9), 1e-3)
    tol = f_ * n_ / ((f_ - n_) * z * z) * world + 2.0 ** -23
    bad = clear & (np.abs(got_depth - want_depth) > tol)
    assert not bad.any(), [(xs[k], ys[k], got_depth[k], want_depth[k], tol[k]) for k in np.flatnonzero(bad)[:5]]
    nrm = tri_n[tri].astype(np.float64)
    nrm = np.where((np.einsum('rk,rk->r', nrm, dvec) > 0)[:, None], -nrm, nrm)
    enc = normal[ys, xs].astype(np.float64) * 2 - 1
    assert np.abs(enc[clear] - nrm[clear]).max() < 1e-6
    # room walls (the scene's first 12 triangles): analytic plane and axis
    wall = clear & (tri < 12)
    for k in np.flatnonzero(wall)[:200]:
        axis = int(np.argmax(np.abs(nrm[k])))
        assert abs(abs(nrm[k][axis]) - 1.0) < 1e-12
        plane = 0.0 if nrm[k][axis] > 0 else ROOM
        assert abs(p[k][axis] - plane) < 1e-9, (p[k], nrm[k])
    return int(clear.sum()), int(wall.sum())


@pytest.mark.parametrize("seed,n_random", [(1, 500), (2, 4_000), (1, 40_000)])
def test_host_ray_caster_equals_brute_force(seed, n_random):
    scene = mlgi.Scene(seed, n_random)
    c1, w1 = check_view(scene, mlgi.Camera.default(1.0), 256, 256, 1800, seed)
    c2, w2 = check_view(scene, mlgi.Camera.light(7 + seed), 128, 128, 1800, seed + 1)
    assert c1 + c2 > 3500
    if n_random <= 4_000:  # 40k triangles hide the walls from both views
        assert w1 + w2 > 0
