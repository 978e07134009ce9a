"""Pins the oracle's stage functions on tiny inputs whose answers are known
(CPU only).

ISM splat (ism.comp:54-139):
  one point in front of one VPL lands on exactly one texel with the packed
  value (uint(d / zFar * 2^24) << 8) | uint(r * 10 / 2); the nearer of two
  points wins a texel; equal depth is decided by the smaller radius byte; a
  point visible to all 16 candidates is splatted into exactly the first four;
  a point behind every candidate hemisphere, or facing away, never lands.
Pull-push (pull.comp, push.comp):
  an empty ISM stays 65535; a fully covered constant-depth ISM returns its
  own quantised depth; a gentle ramp (no texel occluded by a neighbour)
  returns itself; a single sample fills a disc around itself.
Final gathering (final_gathering.comp:63-135):
  one pixel and one VPL with an empty ISM give the closed form
  color * min(cos_v cos_p / d^4, clamp) * gi / N * I^2; a full occluder in
  the ISM gives 0; interleaving with a constant G-buffer averages to the
  non-interleaved result; lit pairs are counted exactly.
FP32 twins agree with the FP64 transcription where rounding allows.
"""
import math

import numpy as np
import pytest

import oracle

ZF = 50.0


def pack_point(pos, normal, radius):
    bits = oracle.pack_unorm4x8((normal[0] * 0.5 + 0.5, normal[1] * 0.5 + 0.5, normal[2] * 0.5 + 0.5,
                                 radius / 25.0))
    rec = np.array([pos[0], pos[1], pos[2], 0.0], np.float32)
    rec[3:4] = np.array([bits], np.uint32).view(np.float32)
    return rec


def quantised_normal(n):
    return oracle.unpack3sn(oracle.pack3sn(n, np.float32), np.float32)


def one_bin(points, n_bins, b=0):
    off = np.zeros(n_bins + 1, np.uint32)
    off[b + 1:] = len(points)
    return np.asarray(points, np.float32).reshape(-1, 4), off


def expected_texel(point, vpos, vn, idx, n1d, side):
    """Closed form of the splat target in float64 (paraboloid, preserveSign)."""
    rel = np.asarray(point, np.float64) - vpos
    d = np.linalg.norm(rel)
    s, u = oracle.look_at_rh(vn, np.float64)
    cz = -np.dot(vn, rel) / d
    x = np.dot(s, rel) / d / (1 - cz)
    y = np.dot(u, rel) / d / (1 - cz)
    S = side // n1d
    tx = (idx % n1d) * S + math.floor((x + 1) / 2 * S)
    ty = (idx // n1d) * S + math.floor((y + 1) / 2 * S)
    return tx, ty, d


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_splat_single_point(dt):
    vpos = np.array([[5.0, 1.0, 5.0]], np.float32)
    vn = quantised_normal((0.0, 1.0, 0.0))[None, :]
    p = (5.3, 4.0, 4.6)
    r_world = 1.0
    pts, off = one_bin([pack_point(p, (0.0, -1.0, 0.0), r_world)], 1)
    atlas = oracle.ism_splat(pts, off, vpos, vn, 64, ZF, dt)
    hit = np.argwhere(atlas != 0xFFFFFFFF)
    assert len(hit) == 1
    tx, ty, d = expected_texel(p, vpos[0].astype(np.float64), vn[0].astype(np.float64), 0, 1, 64)
    assert tuple(hit[0]) == (ty, tx)
    radius = np.float32(round(r_world / 25 * 255)) / np.float32(255) * np.float32(25)
    assert atlas[ty, tx] == (int(d / ZF * 2 ** 24) << 8) | int(radius * 10 / 2)


def test_splat_nearer_wins_and_radius_breaks_ties():
    vpos = np.array([[5.0, 1.0, 5.0]], np.float32)
    vn = quantised_normal((0.0, 1.0, 0.0))[None, :]
    far = pack_point((5.0, 6.0, 5.0), (0.0, -1.0, 0.0), 1.0)
    near = pack_point((5.0, 3.0, 5.0), (0.0, -1.0, 0.0), 1.0)
    pts, off = one_bin([far, near], 1)
    atlas = oracle.ism_splat(pts, off, vpos, vn, 32, ZF)
    v = atlas[atlas != 0xFFFFFFFF]
    assert len(v) == 1 and (v[0] >> 8) == int(np.float32(2.0) / np.float32(ZF) * 2 ** 24)
    a = pack_point((5.0, 3.0, 5.0), (0.0, -1.0, 0.0), 2.0)
    b = pack_point((5.0, 3.0, 5.0), (0.0, -1.0, 0.0), 0.5)
    pts, off = one_bin([a, b], 1)
    atlas = oracle.ism_splat(pts, off, vpos, vn, 32, ZF)
    v = atlas[atlas != 0xFFFFFFFF]
    assert len(v) == 1 and (v[0] & 0xFF) == int(np.float32(round(0.5 / 25 * 255)) / 255 * 25 * 10 / 2)


def test_splat_at_most_four_of_sixteen():
    n = 20
    rng = np.random.default_rng(0)
    vpos = np.column_stack([rng.uniform(2, 8, n), np.full(n, 0.5), rng.uniform(2, 8, n)]).astype(np.float32)
    vn = np.array([quantised_normal((0.0, 1.0, 0.0))] * n, np.float32)
    b = 7
    pts, off = one_bin([pack_point((5.0, 8.0, 5.0), (0.0, -1.0, 0.0), 0.5)], n, b)
    atlas = oracle.ism_splat(pts, off, vpos, vn, 5 * 8 * 4, ZF)  # grid 5x5 ISMs of 32
    n1d, S = oracle.ism_indices_1d(n), (5 * 8 * 4) // oracle.ism_indices_1d(n)
    hits = np.argwhere(atlas != 0xFFFFFFFF)
    isms = sorted({(int(y) // S) * n1d + int(x) // S for y, x in hits})
    assert isms == [(b + i) % n for i in range(4)]


def test_splat_culled_points_never_land():
    vpos = np.array([[5.0, 1.0, 5.0]] * 16, np.float32)
    vn = np.array([quantised_normal((0.0, 1.0, 0.0))] * 16, np.float32)
    behind = pack_point((5.0, 0.5, 5.0), (0.0, -1.0, 0.0), 1.0)   # below every VPL plane
    away = pack_point((5.0, 4.0, 5.0), (0.0, 1.0, 0.0), 1.0)      # faces away from the VPLs
    pts, off = one_bin([behind, away], 16)
    atlas = oracle.ism_splat(pts, off, vpos, vn, 64, ZF)
    assert np.all(atlas == 0xFFFFFFFF)


def test_splat_empty_bins():
    vpos = np.zeros((16, 3), np.float32)
    vn = np.array([quantised_normal((0.0, 1.0, 0.0))] * 16, np.float32)
    atlas = oracle.ism_splat(np.zeros((0, 4), np.float32), np.zeros(17, np.uint32), vpos, vn, 64, ZF)
    assert np.all(atlas == 0xFFFFFFFF)


# ----------------------------------------------------------- pull-push
def packed(depth, radius_byte):
    d24 = np.asarray(depth * 2 ** 24).astype(np.uint64)
    return ((d24 << np.uint64(8)) | np.asarray(radius_byte, np.uint64)).astype(np.uint32)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_pullpush_empty_stays_far(dt):
    atlas = np.full((128, 128), 0xFFFFFFFF, np.uint32)
    out = oracle.pullpush(atlas, 32, 4, ZF, dt)
    r16 = out if dt == np.float32 else out[0]
    assert np.all(r16 == 65535)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_pullpush_covered_constant_is_identity(dt):
    d = 0.25
    atlas = np.full((64, 64), packed(d, 5), np.uint32)
    out = oracle.pullpush(atlas, 32, 5, ZF, dt)
    r16 = out if dt == np.float32 else out[0]
    assert np.all(r16 == round(float(np.float32(d)) * 65535))


def test_pullpush_gentle_ramp_is_identity():
    y, x = np.mgrid[0:64, 0:64]
    d = (0.2 + 0.0004 * x + 0.0003 * y)
    atlas = packed(d, 8)
    r16 = oracle.pullpush(atlas, 64, 4, ZF)
    dq = (atlas >> 8).astype(np.float64) / 2 ** 24
    assert np.all(r16 == np.rint(dq * 65535))


def test_pullpush_single_sample_fills_a_disc():
    atlas = np.full((64, 64), 0xFFFFFFFF, np.uint32)
    d, rb = 0.2, 30
    atlas[32, 32] = packed(d, rb)
    r16 = oracle.pullpush(atlas, 64, 6, ZF)
    filled = np.argwhere(r16 < 65535)
    # pull.comp:57-65: radius in pixels = r / (d zFar) / 3.14 * S * 1.3, at most 15
    r_px = min(rb / 10 / (d * ZF) / 3.14 * 64 * 1.3, 15.0)
    dist = np.hypot(filled[:, 0] - 32, filled[:, 1] - 32)
    assert len(filled) > 9 and dist.max() <= r_px + 2.0
    assert np.all(r16[31:34, 31:34] < 65535)
    assert abs(int(r16[32, 32]) - round(d * 65535)) <= 1


def test_pullpush_fp32_twin_vs_fp64():
    rng = np.random.default_rng(1)
    atlas = np.full((256, 256), 0xFFFFFFFF, np.uint32)
    idx = rng.integers(0, 256, size=(3000, 2))
    atlas[idx[:, 0], idx[:, 1]] = packed(rng.uniform(0.05, 0.3, 3000), rng.integers(1, 60, 3000))
    a = oracle.pullpush(atlas, 32, 5, ZF, np.float32)
    b, _ = oracle.pullpush(atlas, 32, 5, ZF, np.float64)
    diff = np.abs(a.astype(np.int64) - b.astype(np.int64))
    assert (diff > 0).mean() < 0.01 and np.quantile(diff, 0.999) <= 1


# ------------------------------------------------------ final gathering
def single_pixel_gbuffer(world, normal):
    """1x1 G-buffer whose pixel reconstructs to `world`: a VP inverse that
    ignores uv and maps ndc z to a fixed point (w = 1)."""
    m = np.zeros(16, np.float32)
    m[12:15] = world  # constant column
    m[15] = 1.0
    depth = np.zeros((1, 1), np.float32)
    nrm = (np.asarray(normal, np.float32) * 0.5 + 0.5).reshape(1, 1, 3)
    return depth, nrm, m


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_gather_closed_form_one_pair(dt):
    world = np.array([2.0, 0.0, 1.0])
    depth, nrm, m = single_pixel_gbuffer(world, (0.0, 1.0, 0.0))
    for vp, clamp in [((2.5, 3.0, 0.8), 1e-3), ((2.2, 1.0, 1.4), 1e-3), ((3.0, 0.5, 3.0), 10.0)]:
        vpos = np.array([vp], np.float32)
        vn = quantised_normal(-(np.array(vp) - world) / np.linalg.norm(np.array(vp) - world))[None, :]
        col = np.array([[2.0, 1.0, 0.5]], np.float32)
        ism = np.full((4, 4), 65535, np.uint16)
        out, lit = oracle.final_gather(depth, nrm, m, vpos, vn, col, ism, 4, ZF, 3000.0, clamp, 1,
                                       dtype=dt)
        diff = world - vpos[0].astype(np.float64)
        d = np.linalg.norm(diff)
        cv = max(0.0, vn[0].astype(np.float64) @ (diff / d))
        cp = max(0.0, np.array([0.0, 1.0, 0.0]) @ (-diff / d))
        # clamp and gi reach the shader as binary32 uniforms
        want = col[0].astype(np.float64) * min(cv * cp / d ** 4, float(np.float32(clamp))) * 3000.0
        assert lit == 1
        assert np.allclose(out[0, 0], want, rtol=1e-5 if dt == np.float32 else 1e-12)


def test_gather_full_occluder_blocks():
    world = np.array([2.0, 0.0, 1.0])
    depth, nrm, m = single_pixel_gbuffer(world, (0.0, 1.0, 0.0))
    vp = np.array([[2.0, 3.0, 1.0]], np.float32)
    vn = quantised_normal((0.0, -1.0, 0.0))[None, :]
    col = np.ones((1, 3), np.float32)
    free, _ = oracle.final_gather(depth, nrm, m, vp, vn, col, np.full((2, 2), 65535, np.uint16), 2, ZF,
                                  3000.0, 1e-3, 1)
    blocked, _ = oracle.final_gather(depth, nrm, m, vp, vn, col, np.zeros((2, 2), np.uint16), 2, ZF,
                                     3000.0, 1e-3, 1)
    no_shadow, _ = oracle.final_gather(depth, nrm, m, vp, vn, col, None, 2, ZF, 3000.0, 1e-3, 1,
                                       shadowing=False)
    assert free[0, 0, 0] > 0 and np.all(blocked == 0.0)
    assert np.array_equal(free, no_shadow)


def test_gather_interleave_average_equals_full_sum():
    """A G-buffer whose every pixel reconstructs to one point: the mean over an
    I x I block of the interleaved result is the non-interleaved sum."""
    rng = np.random.default_rng(2)
    I, n = 4, 64
    world = np.array([5.0, 0.0, 5.0])
    m = np.zeros(16, np.float32)
    m[12:15] = world
    m[15] = 1.0
    depth = np.zeros((8, 8), np.float32)
    nrm = np.broadcast_to(np.array([0.5, 1.0, 0.5], np.float32), (8, 8, 3)).copy()
    vpos = np.column_stack([rng.uniform(1, 9, n), rng.uniform(3, 9, n), rng.uniform(1, 9, n)]).astype(np.float32)
    vn = np.array([quantised_normal((0.0, -1.0, 0.0))] * n, np.float32)
    col = rng.uniform(0.2, 1.0, (n, 3)).astype(np.float32)
    ism = np.full((64, 64), 65535, np.uint16)
    inter, lit_i = oracle.final_gather(depth, nrm, m, vpos, vn, col, ism, 64, ZF, 3000.0, 10.0, I)
    full, lit_f = oracle.final_gather(depth, nrm, m, vpos, vn, col, ism, 64, ZF, 3000.0, 10.0, 1)
    assert np.allclose(inter[:I, :I].reshape(-1, 3).mean(0), full[0, 0], rtol=1e-12)
    # every pair is lit; interleaving gives each pixel 1/I^2 of the VPLs
    assert lit_f == 64 * n and lit_i == 64 * n // (I * I)


def test_gather_fp32_twin_vs_fp64():
    rng = np.random.default_rng(3)
    h, w, n = 16, 16, 64
    m = np.zeros(16, np.float32)
    m[0], m[5], m[12], m[13], m[14], m[15] = 4.0, 4.0, 5.0, 5.0, 0.0, 1.0  # plane z = 0 region
    depth = np.zeros((h, w), np.float32)
    nrm = np.broadcast_to(np.array([0.5, 0.5, 1.0], np.float32), (h, w, 3)).copy()
    vpos = np.column_stack([rng.uniform(1, 9, n), rng.uniform(1, 9, n), rng.uniform(1, 5, n)]).astype(np.float32)
    vn = np.array([quantised_normal((0.0, 0.0, -1.0))] * n, np.float32)
    col = rng.uniform(0.2, 1.0, (n, 3)).astype(np.float32)
    ism = rng.integers(60000, 65536, (64, 64)).astype(np.uint16)
    a, la = oracle.final_gather(depth, nrm, m, vpos, vn, col, ism, 64, ZF, 3000.0, 1e-3, 2)
    b, lb = oracle.final_gather(depth, nrm, m, vpos, vn, col, ism, 64, ZF, 3000.0, 1e-3, 2, dtype=np.float32)
    assert la == lb > 0
    assert np.linalg.norm(b - a) / np.linalg.norm(a) < 1e-5


def test_threaded_splat_and_pullpush_match_single_thread():
    # the CPU baseline runs these on many threads; the atlas is a minimum and
    # the pull-push rows are independent, so the bits must not change
    rng = np.random.default_rng(5)
    n = 256
    vpos = np.column_stack([rng.uniform(1, 9, n), rng.uniform(1, 9, n), rng.uniform(1, 9, n)]).astype(np.float32)
    vn = np.array([quantised_normal(v / np.linalg.norm(v)) for v in rng.normal(size=(n, 3))], np.float32)
    pts = [pack_point(rng.uniform(1, 9, 3), rng.normal(size=3) / 1.8, rng.uniform(0.05, 0.5)) for _ in range(30000)]
    pts = np.array(pts, np.float32)
    bins = rng.integers(0, n, len(pts))
    order = np.argsort(bins, kind="stable")
    off = np.concatenate([[0], np.cumsum(np.bincount(bins, minlength=n))]).astype(np.uint32)
    a1 = oracle.ism_splat(pts[order], off, vpos, vn, 512, ZF)
    a8 = oracle.ism_splat(pts[order], off, vpos, vn, 512, ZF, threads=8)
    assert (a1 != 0xFFFFFFFF).sum() > 1000
    assert np.array_equal(a1, a8)
    assert np.array_equal(oracle.pullpush(a1, 32, 4, ZF), oracle.pullpush(a1, 32, 4, ZF, threads=8))
