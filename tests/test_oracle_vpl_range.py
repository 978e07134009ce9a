"""Known answers for the oracle's VPL-range forms (CPU only).

The reference's scaleISMs and pointsOnlyIntoScaledISMs act on the VPL range
[vplStartIndex, vplEndIndex) (ism.comp:31-40,56-66,124; ism_utils.glsl:33-37;
final_gathering.comp:40-45): with scaleISMs the atlas is laid out for the
range's VPL count and VPL v splats into, and is gathered from, ISM v - start;
with pointsOnlyIntoScaledISMs the points are binned over the range's count
and bin b takes its 16 candidates ((b + i) mod N) mod count + start. ISM
indices outside the grid follow int() (truncation toward zero), GL's invalid
image atomics (no effect) and the sampler's GL_REPEAT wrap.
"""
import numpy as np
import pytest

import oracle

Z_FAR = 50.0
A = 64  # atlas side


def point(pos, normal, radius=0.4):
    b = [int(np.clip(np.rint((c * 0.5 + 0.5) * 255.0), 0, 255)) for c in normal]
    b.append(int(np.rint(radius / 25.0 * 255.0)))
    w = np.uint32(b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24)).view(np.float32)
    return [pos[0], pos[1], pos[2], w]


def facing(rel):
    d = -np.asarray(rel, np.float64)
    return d / np.linalg.norm(d)


def one_bin(points, n_bins=1):
    """All points in bin 0 of n_bins."""
    off = np.zeros(n_bins + 1, np.uint32)
    off[1:] = len(points)
    return np.array(points, np.float32), off


def vpls_at_origin(n, up):
    """n VPLs at the origin; VPLs in `up` face +z, the others -z."""
    pos = np.zeros((n, 3), np.float32)
    nrm = np.zeros((n, 3), np.float32)
    nrm[:, 2] = [1.0 if v in up else -1.0 for v in range(n)]
    return pos, nrm


def written(atlas):
    return sorted(map(tuple, np.argwhere(atlas != 0xFFFFFFFF)[:, ::-1].tolist()))  # (x, y)


def test_full_range_equals_the_plain_splat():
    rng = np.random.default_rng(3)
    n = 16
    pos = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    pts = [point(rng.uniform(-3, 3, 3), rng.normal(size=3) / 1.7) for _ in range(2000)]
    pts = np.array(pts, np.float32)
    bins = np.sort(rng.integers(0, n, len(pts)))
    off = np.concatenate([[0], np.cumsum(np.bincount(bins, minlength=n))]).astype(np.uint32)
    plain = oracle.ism_splat(pts, off, pos, nrm, A, Z_FAR)
    assert (plain != 0xFFFFFFFF).sum() > 100
    for scale in (False, True):
        for only in (False, True):
            got = oracle.ism_splat(pts, off, pos, nrm, A, Z_FAR, vpl_range=(0, n), scale_isms=scale,
                                   points_only=only)
            assert np.array_equal(got, plain), (scale, only)


def test_scale_isms_lays_out_the_range():
    # 8 VPLs, VPL 5 faces the point straight ahead (tile centre), the others
    # face away; range [3, 7): 4 ISMs in a 2 x 2 grid of 32^2, VPL 5 in ISM 2
    pos, nrm = vpls_at_origin(8, up={5})
    pts, off = one_bin([point((0, 0, 1), facing((0, 0, 1)))], n_bins=8)
    off[1:] = 1  # the point in bin 0 (candidates 0..7 twice)
    scaled = oracle.ism_splat(pts, off, pos, nrm, A, Z_FAR, vpl_range=(3, 7), scale_isms=True)
    assert written(scaled) == [(16, 48)]  # tile (0, 1) of side 32, centre
    plain = oracle.ism_splat(pts, off, pos, nrm, A, Z_FAR)
    assert written(plain) == [(24, 24)]   # 4 x 4 grid of 16^2: VPL 5 in tile (1, 1)


@pytest.mark.parametrize("delta,expect", [(0.02, [(0, 16)]), (0.2, [])])
def test_an_index_before_the_range_truncates_into_column_0_or_writes_nothing(delta, expect):
    # VPL 2 before the range [3, 7): ISM index -1, tile (int(-1/2), -1) = (-1, 0).
    # A point nearly along the basis row s = -x lands at local x = 1 - delta/2,
    # i.e. atlas x = (local - 1) * 32: above -1 truncates to column 0 (ivec2),
    # below it is an invalid atomic
    pos, nrm = vpls_at_origin(8, up={2})
    rel = np.array([-1.0, 0.0, delta]) / np.hypot(1.0, delta) * 2.0
    pts, off = one_bin([point(rel, facing(rel))], n_bins=8)
    off[1:] = 1
    got = oracle.ism_splat(pts, off, pos, nrm, A, Z_FAR, vpl_range=(3, 7), scale_isms=True)
    assert written(got) == expect


def test_points_only_takes_candidates_from_the_range():
    # range [3, 5): bin 0 of 2 takes VPLs 3, 4, 3, 4, ...; VPL 3 faces away, VPL 4
    # faces the point, and so do VPLs 0-2 and 5-7, which are no candidates then
    pos, nrm = vpls_at_origin(8, up={0, 1, 2, 4, 5, 6, 7})
    pts, off = one_bin([point((0, 0, 1), facing((0, 0, 1)))], n_bins=2)
    only = oracle.ism_splat(pts, off, pos, nrm, A, Z_FAR, vpl_range=(3, 5), points_only=True)
    assert written(only) == [(8, 24)]  # VPL 4's tile (0, 1) of the 4 x 4 grid of 16^2
    both = oracle.ism_splat(pts, off, pos, nrm, A, Z_FAR, vpl_range=(3, 5), points_only=True, scale_isms=True)
    assert written(both) == [(A // 2 + A // 4, A // 4)]  # 2 ISMs: grid 2 x 2 of 32^2, VPL 4 in ISM 1


def plane_frame():
    """A 2 x 2 G-buffer on the plane y = 0 around (3, 0, 4), normals +y."""
    m = np.zeros(16, np.float32)  # world = (3 + 0.1 ndc_x, 0, 4 + 0.1 ndc_y)
    m[0], m[6], m[12], m[14], m[15] = 0.1, 0.1, 3.0, 4.0, 1.0
    depth = np.full((2, 2), 0.5, np.float32)
    normal = np.broadcast_to(np.array([0.5, 1.0, 0.5], np.float32), (2, 2, 3)).copy()
    return depth, normal, m


def test_gather_samples_ism_index_minus_the_range_start():
    # VPL 5 lights the pixels from above; the others face away from them.
    # The ISM has one occluder tile: ISM 2 of the range [3, 7)'s 2 x 2 grid.
    depth, normal, m = plane_frame()
    n = 8
    pos = np.tile(np.array([3.0, 1.0, 4.0], np.float32), (n, 1))
    nrm = np.tile(np.array([0.0, 1.0, 0.0], np.float32), (n, 1))
    nrm[5] = [0.0, -1.0, 0.0]
    col = np.ones((n, 3), np.float32)
    ism = np.full((A, A), 65535, np.uint16)
    ism[32:64, 0:32] = 0  # tile (0, 1) of side 32
    args = dict(atlas_side=A, z_far=Z_FAR, gi_intensity=1.0, clamp_value=10.0, interleave=1, vpl_start=3,
                vpl_end=7)
    scaled, lit = oracle.final_gather(depth, normal, m, pos, nrm, col, ism, scale_isms=True, **args)
    plain, lit2 = oracle.final_gather(depth, normal, m, pos, nrm, col, ism, **args)
    assert lit == lit2 == 4  # VPL 5 and the four pixels
    assert np.all(scaled == 0.0)  # shadowed: ISM 2 is the occluder tile
    assert np.all(plain > 0.0)    # VPL 5's own tile (1, 1) of the 4 x 4 grid is empty
