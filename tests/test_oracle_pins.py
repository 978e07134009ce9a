"""Known-answer pins of the oracle's VPL extraction and pull-push.

Each expected value here comes from something other than the oracle's own
code: hand-listed texel coordinates, bilinear weights derived from pixel
centres, projection round trips, and geometry (ISM borders). Each test is
chosen so that a plausible transcription mistake fails it:

  * VPL extraction (vpl_processor.comp:46-63): x and y of the sampling
    grid swapped, the scatter to shuffle[i] dropped, the flux scale
    dropped, the matrix read row-major, the texel centre used instead of
    the corner.
  * push (push.comp:48-86): the neighbour order of the 9/3/3/1 weights
    changed (e.g. to the pull's order), the ISM-border shift wrong.
"""
import numpy as np
import pytest

import oracle

ZF = 50.0
DTYPES = [np.float32, np.float64]


# ------------------------------------------------------------ VPL extraction
def encoded_rsm(w, h, depth=0.5, normal=(0.0, 1.0, 0.0)):
    """RSM whose diffuse colour encodes the texel: ((x+1)/16, (y+1)/16, 1/4)."""
    y, x = np.mgrid[0:h, 0:w]
    dif = np.stack([(x + 1) / 16.0, (y + 1) / 16.0, np.full_like(x, 4) / 16.0], -1).astype(np.float32)
    nrm = np.broadcast_to(np.asarray(normal, np.float32) * 0.5 + 0.5, (h, w, 3)).copy()
    dep = np.full((h, w), depth, np.float32)
    return dif, nrm, dep


def decode_texel(col, intensity):
    """Texel (x, y) from a VPL colour of an encoded RSM; None for black."""
    d = np.asarray(col, np.float64) / intensity
    if not d.any():
        return None
    return int(round(d[0] * 16 - 1)), int(round(d[1] * 16 - 1))


AFFINE = np.array([  # column-major: world = M (u, v, depth, 1); w stays 1
    3.0, 0.5, -1.0, 0.0,
    -0.25, 2.0, 0.75, 0.0,
    1.5, -2.5, 4.0, 0.0,
    0.125, 1.0, -3.0, 1.0], np.float32)


@pytest.mark.parametrize("dt", DTYPES)
def test_vpl_grid_is_row_major_on_a_wide_rsm(dt):
    # N = 16 on 8 x 4: s = ceil(sqrt 16) = 4 samples per row, 16/4 = 4 rows;
    # x = (i mod 4) * 8/4, y = (i div 4) * 4/4, x fastest.
    want = [(0, 0), (2, 0), (4, 0), (6, 0), (0, 1), (2, 1), (4, 1), (6, 1),
            (0, 2), (2, 2), (4, 2), (6, 2), (0, 3), (2, 3), (4, 3), (6, 3)]
    dif, nrm, dep = encoded_rsm(8, 4)
    _, _, _, col = oracle.vpl_extract(dif, nrm, dep, AFFINE, 2.0, None, 16, dt)
    assert [decode_texel(c, 2.0) for c in col] == want


@pytest.mark.parametrize("dt", DTYPES)
def test_vpl_grid_of_a_non_square_count(dt):
    # N = 8 on 6 x 4: s = 3, 8 div 3 = 2 rows of stride 4/2 = 2; the third row
    # (i = 6, 7) lands on y = 4, outside the RSM: black VPLs.
    want = [(0, 0), (2, 0), (4, 0), (0, 2), (2, 2), (4, 2), None, None]
    dif, nrm, dep = encoded_rsm(6, 4)
    _, _, _, col = oracle.vpl_extract(dif, nrm, dep, AFFINE, 1.0, None, 8, dt)
    assert [decode_texel(c, 1.0) for c in col] == want


@pytest.mark.parametrize("dt", DTYPES)
def test_vpl_scatter_to_shuffled_slot_and_flux(dt):
    n = 16
    perm = np.array([11, 3, 14, 0, 7, 9, 1, 15, 4, 12, 6, 2, 13, 8, 10, 5], np.int32)
    dif, nrm, dep = encoded_rsm(8, 4)
    _, _, _, plain = oracle.vpl_extract(dif, nrm, dep, AFFINE, 2.0, None, n, dt)
    pos_s, _, pk_s, col_s = oracle.vpl_extract(dif, nrm, dep, AFFINE, 2.0, perm, n, dt)
    pos_p, _, pk_p, _ = oracle.vpl_extract(dif, nrm, dep, AFFINE, 2.0, None, n, dt)
    # resultIndex = shuffledIndicesBuffer[i]: sample i is written to slot perm[i]
    for i in range(n):
        assert np.array_equal(col_s[perm[i]], plain[i])
        assert np.array_equal(pos_s[perm[i]], pos_p[i])
    # flux = diffuse * lightIntensity, exact for these binary fractions
    grid = [(0, 0), (2, 0), (4, 0), (6, 0), (0, 1), (2, 1), (4, 1), (6, 1),
            (0, 2), (2, 2), (4, 2), (6, 2), (0, 3), (2, 3), (4, 3), (6, 3)]
    for i, (x, y) in enumerate(grid):
        assert np.array_equal(plain[i], np.array([(x + 1) / 8.0, (y + 1) / 8.0, 0.5], dt))


@pytest.mark.parametrize("dt", DTYPES)
def test_vpl_affine_unprojection_at_the_texel_corner(dt):
    # worldcoords = M * vec4(texel / size, depth, 1): the texel's lower-left
    # corner, not its centre; the matrix is column-major.
    w, h, depth = 8, 4, 0.375
    dif, nrm, dep = encoded_rsm(w, h, depth)
    pos, _, _, _ = oracle.vpl_extract(dif, nrm, dep, AFFINE, 1.0, None, 16, dt)
    m = AFFINE.astype(np.float64).reshape(4, 4).T  # row-major view of the column-major matrix
    for i in range(16):
        x, y = (i % 4) * 2, i // 4
        want = m @ np.array([x / w, y / h, depth, 1.0])
        assert np.allclose(pos[i], want[:3], rtol=0, atol=2e-6), (i, pos[i], want)


def perspective(fovy_deg, aspect, near, far):
    f = 1.0 / np.tan(np.radians(fovy_deg) / 2)
    return np.array([[f / aspect, 0, 0, 0], [0, f, 0, 0],
                     [0, 0, (far + near) / (near - far), 2 * far * near / (near - far)], [0, 0, -1, 0]])


def look_at(eye, center, up):
    f = np.asarray(center, float) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    v = np.eye(4)
    v[0, :3], v[1, :3], v[2, :3] = s, u, -f
    v[:3, 3] = -v[:3, :3] @ eye
    return v


@pytest.mark.parametrize("dt", DTYPES)
def test_vpl_perspective_unprojection_round_trip(dt):
    # A spot light above the floor plane y = 0, looking down: RSM depth from
    # the forward projection of the floor point seen through each texel
    # corner; unprojecting must land on the floor at that point.
    eye = np.array([1.0, 6.0, 2.0])
    bias = np.array([[0.5, 0, 0, 0.5], [0, 0.5, 0, 0.5], [0, 0, 0.5, 0.5], [0, 0, 0, 1]])
    bvp = bias @ perspective(45.0, 1.0, 0.5, 30.0) @ look_at(eye, np.array([1.2, 0.0, 2.1]), np.array([0, 0, 1.0]))
    inv = np.linalg.inv(bvp)
    n_side, size = 8, 32
    dep = np.empty((size, size), np.float32)
    floor = {}
    for ty in range(size):
        for tx in range(size):
            # ray through window point (u, v) at two depths, intersected with y = 0
            a = inv @ np.array([tx / size, ty / size, 0.0, 1.0])
            b = inv @ np.array([tx / size, ty / size, 1.0, 1.0])
            a, b = a[:3] / a[3], b[:3] / b[3]
            t = a[1] / (a[1] - b[1])
            p = a + t * (b - a)
            q = bvp @ np.append(p, 1.0)
            dep[ty, tx] = q[2] / q[3]
            floor[(tx, ty)] = p
    dif = np.full((size, size, 3), 0.5, np.float32)
    nrm = np.broadcast_to(np.array([0.5, 1.0, 0.5], np.float32), (size, size, 3)).copy()
    m16 = inv.T.astype(np.float32).reshape(16)  # column-major
    pos, _, _, _ = oracle.vpl_extract(dif, nrm, dep, m16, 1.0, None, n_side * n_side, dt)
    for i in range(n_side * n_side):
        tx, ty = (i % n_side) * size // n_side, (i // n_side) * size // n_side
        # the FP32 depth and matrix carry ~1e-7 relative error, amplified by
        # the perspective division near the far plane
        assert np.allclose(pos[i], floor[(tx, ty)], rtol=0, atol=5e-4), (i, pos[i], floor[(tx, ty)])


# ----------------------------------------------------------------- push
def packed(depth, radius_byte):
    d24 = np.asarray(np.asarray(depth, np.float64) * 2 ** 24).astype(np.uint64)
    return ((d24 << np.uint64(8)) | np.asarray(radius_byte, np.uint64)).astype(np.uint32)


def bilinear_weights(px, py):
    """The four coarse texels around fine pixel (px, py) and their bilinear
    weights at the fine pixel's centre; coarse texel c is centred at fine
    coordinate 2c + 1."""
    out = {}
    cx0, cy0 = (px + 1) // 2 - 1, (py + 1) // 2 - 1
    for cy in (cy0, cy0 + 1):
        for cx in (cx0, cx0 + 1):
            wx = 1.0 - abs((px + 0.5) - (2 * cx + 1)) / 2.0
            wy = 1.0 - abs((py + 0.5) - (2 * cy + 1)) / 2.0
            out[(cx, cy)] = wx * wy
    return out


def test_bilinear_weights_are_9_3_3_1():
    for p in [(2, 2), (5, 2), (2, 5), (5, 5)]:
        w = sorted(bilinear_weights(*p).values())
        assert w == [1 / 16, 3 / 16, 3 / 16, 9 / 16]


@pytest.mark.parametrize("dt", DTYPES)
def test_push_fills_holes_with_9_3_3_1_bilinear_weights(dt):
    # One 8x8 ISM, one pull level. Every 2x2 block has one depth (k/256), so
    # the level-1 pull of a block is that depth exactly, also for the blocks
    # with a hole (the empty texel decodes to a far depth that the depth-range
    # test removes). A hole is then the 9/3/3/1 mean of the four coarse texels
    # around it; every other texel keeps its own depth.
    S = 8
    blocks = np.array([[64, 66, 71, 65], [69, 77, 64, 73], [67, 75, 70, 79], [72, 68, 78, 66]], np.float64) / 256.0
    depth = np.kron(blocks, np.ones((2, 2)))
    atlas = packed(depth, 255)
    holes = [(2, 2), (5, 2), (2, 5), (5, 5)]  # (x, y): every parity of the quad
    for x, y in holes:
        atlas[y, x] = 0xFFFFFFFF
    out = oracle.pullpush(atlas, S, 1, ZF, dt)
    r16 = out if dt == np.float32 else out[0]
    want = np.rint(depth * 65535)
    for x, y in holes:
        fill = sum(wt * blocks[cy, cx] for (cx, cy), wt in bilinear_weights(x, y).items())
        frac = fill * 65535 - np.floor(fill * 65535)
        assert abs(frac - 0.5) > 1e-3  # no rounding tie in the expected value
        want[y, x] = np.rint(fill * 65535)
    assert np.array_equal(r16, want.astype(np.uint16)), np.argwhere(r16 != want)


def one_sample_atlas(S, n1d, x, y, depth=0.25, radius_byte=255):
    atlas = np.full((S * n1d, S * n1d), 0xFFFFFFFF, np.uint32)
    atlas[y, x] = packed(depth, radius_byte)
    return atlas


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("S,levels", [(8, 3), (16, 3), (16, 4)])
def test_push_never_fills_across_an_ism_border(dt, S, levels):
    # A sample on the border column / row of one ISM, with a large radius
    # (about 7 px at S = 8). Its fill must stay inside its own ISM
    # (push.comp:77-86, the border at the ISM side, not a fixed 64 px).
    n1d = 2
    cases = [(S, 3, (1, 0)), (S - 1, 3, (0, 0)), (S + 3, S, (1, 1)), (3, S - 1, (0, 0))]
    for x, y, own in cases:
        atlas = one_sample_atlas(S, n1d, x, y)
        out = oracle.pullpush(atlas, S, levels, ZF, dt)
        r16 = out if dt == np.float32 else out[0]
        for iy in range(n1d):
            for ix in range(n1d):
                tile = r16[iy * S:(iy + 1) * S, ix * S:(ix + 1) * S]
                if (ix, iy) == own:
                    assert (tile < 65535).sum() > 4, (x, y)  # the sample spreads inside its ISM
                else:
                    assert np.all(tile == 65535), (x, y, ix, iy)


@pytest.mark.parametrize("dt", DTYPES)
def test_push_keeps_a_valid_unoccluded_sample_and_replaces_an_occluded_one(dt):
    # push.comp:150-161: a valid texel keeps its own depth unless it lies
    # behind the max depth of a contributing coarse texel. Radius byte 50:
    # r = 5 world units, maxDepth = d + 2 r / zFar * 0.6 = d + 0.12, and
    # r_px = r / (d zFar) / 3.14 * 8 * 1.3 = 1.10 .. 1.32 px (pull.comp:55-65).
    S = 8
    blocks = np.array([[64, 66, 71, 65], [69, 77, 64, 73], [67, 75, 70, 79], [72, 68, 78, 66]], np.float64) / 256.0
    depth = np.kron(blocks, np.ones((2, 2)))
    depth[3, 3] = 224 / 256.0  # behind its block's max depth: dropped by the pull's depth range
    atlas = packed(depth, 50)
    out = oracle.pullpush(atlas, S, 1, ZF, dt)
    r16 = out if dt == np.float32 else out[0]
    # (3, 3) is occluded by its own block (depth 77/256, max depth 77/256 +
    # 0.12); the other three coarse texels around it lie 1.58 and 2.12 px
    # away, beyond their radii, so the fill is its block's depth alone
    want = np.rint(depth * 65535)
    want[3, 3] = np.rint(blocks[1, 1] * 65535)
    assert np.array_equal(r16, want.astype(np.uint16)), np.argwhere(r16 != want)
