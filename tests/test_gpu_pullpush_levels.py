"""GPU: pull-push of every ISM side and level count against the oracle's FP32
twin, bit for bit, on random atlases.

The workloads cover sides 32 and 64 at three and four levels; this sweep adds
sides 8 and 16 and every level count from 1 (the level-0 push reads the pull
directly) to log2(side). Texels are empty (0xFFFFFFFF) with a density that
differs per ISM (from nearly full to entirely empty), otherwise a random
24-bit depth and radius byte, so that pulls, occlusion tests and fills all
take both branches. One more case runs the largest atlas accepted, 16384^2,
whose byte offsets reach 2^30.
"""
import numpy as np
import pytest

import mlgi
import oracle

pytestmark = pytest.mark.gpu

N_VPL = 16   # 4 x 4 ISMs
Z_FAR = 50.0


def random_atlas(side, ism_side, seed):
    rng = np.random.default_rng(seed)
    n1d = side // ism_side
    empty_p = np.repeat(np.repeat(rng.choice([0.1, 0.5, 0.8, 0.97, 1.0], (n1d, n1d)), ism_side, 0), ism_side, 1)
    depth = rng.integers(1 << 18, (1 << 24) - 1, (side, side), dtype=np.uint32)
    # a smooth depth surface under the noise, so that neighbours often agree
    yy, xx = np.mgrid[0:side, 0:side]
    base = ((np.sin(xx * 0.3) + np.cos(yy * 0.2) + 2.0) * 0.25 * (1 << 23)).astype(np.uint32)
    depth = np.where(rng.random((side, side)) < 0.5, base + (depth & 0xFFFF), depth)
    radius = rng.integers(0, 256, (side, side), dtype=np.uint32)
    texel = (depth << 8) | radius
    return np.where(rng.random((side, side)) < empty_p, np.uint32(0xFFFFFFFF), texel).astype(np.uint32)


CASES = [(s, l) for s in (8, 16, 32, 64) for l in range(1, int(np.log2(s)) + 1)]


@pytest.mark.parametrize("ism_side,levels", CASES)
def test_pullpush_levels_bit_exact(ctx, ism_side, levels):
    side = 4 * ism_side
    print(f"[pullpush] S={ism_side} L={levels} atlas={side}", flush=True)
    atlas = random_atlas(side, ism_side, seed=ism_side * 10 + levels)
    ism = mlgi.ImperfectShadowmap(ctx, side, N_VPL)
    assert ism.ism_side == ism_side
    ism.softrender.copy_from(atlas)
    p = mlgi.default_params(atlas_side=side, z_far=Z_FAR, pull_push_levels=levels)
    ism.pullpush(p)
    ctx.sync()
    got = ism.pushpull_result.numpy()
    ref = oracle.pullpush(atlas, ism_side, levels, Z_FAR)
    assert len(np.unique(ref)) > 16
    assert np.array_equal(got, ref), int((got != ref).sum())



def test_pullpush_largest_atlas_bit_exact(ctx):
    # the largest atlas the entry point accepts, 16384^2 (1 GiB of u32
    # texels): the kernel addresses atlas and result with 32-bit byte offsets
    # of buffers sized exactly to them, which reach 2^30 here.
    # Sparse: every 37th row of ISMs and the last row carry samples.
    side, ism_side, levels = 16384, 64, 4
    print(f"[pullpush] S={ism_side} L={levels} atlas={side}", flush=True)
    rng = np.random.default_rng(16384)
    atlas = np.full((side, side), 0xFFFFFFFF, np.uint32)
    rows = list(range(0, side // ism_side, 37)) + [side // ism_side - 1]
    for r in rows:
        band = atlas[r * ism_side:(r + 1) * ism_side]
        depth = rng.integers(1 << 18, (1 << 24) - 1, band.shape, dtype=np.uint32)
        radius = rng.integers(0, 256, band.shape, dtype=np.uint32)
        band[...] = np.where(rng.random(band.shape) < 0.3, (depth << 8) | radius, np.uint32(0xFFFFFFFF))
    ism = mlgi.ImperfectShadowmap(ctx, side, (side // ism_side) ** 2)
    assert ism.ism_side == ism_side
    ism.softrender.copy_from(atlas)
    p = mlgi.default_params(atlas_side=side, z_far=Z_FAR, pull_push_levels=levels)
    ism.pullpush(p)
    ctx.sync()
    got = ism.pushpull_result.numpy()
    ref = oracle.pullpush(atlas, ism_side, levels, Z_FAR, threads=16)
    assert np.array_equal(got, ref), int((got != ref).sum())
    # the last row of ISMs (the largest offsets) holds filled samples
    assert len(np.unique(ref[-ism_side:])) > 16


def border_atlas(ism_side, seed):
    # 4 x 4 ISMs whose samples lie only along their borders: one edge, one
    # corner, two opposite edges, none or all, with the radius byte at its
    # maximum (fills reach as far as the reference allows) or random. A push
    # quad pairs pixels on one side of an ISM with neighbours on the other
    # side only if its border handling is wrong, and then these atlases fill
    # texels the oracle leaves untouched.
    rng = np.random.default_rng(seed)
    s = ism_side
    side = 4 * s
    atlas = np.full((side, side), 0xFFFFFFFF, np.uint32)
    masks = []
    for k in range(16):
        m = np.zeros((s, s), bool)
        if k == 0: m[:, 0] = True
        elif k == 1: m[:, -1] = True
        elif k == 2: m[0, :] = True
        elif k == 3: m[-1, :] = True
        elif k == 4: m[0, 0] = True
        elif k == 5: m[0, -1] = True
        elif k == 6: m[-1, 0] = True
        elif k == 7: m[-1, -1] = True
        elif k == 8: m[:, [0, -1]] = True
        elif k == 9: m[[0, -1], :] = True
        elif k == 11: m[...] = True
        elif k >= 12: m[:, :2] = m[:2, :] = True  # two-texel bands along two edges
        masks.append(m)
    for k, m in enumerate(masks):
        y0, x0 = (k // 4) * s, (k % 4) * s
        depth = rng.integers(1 << 18, (1 << 24) - 1, (s, s), dtype=np.uint32)
        radius = np.full((s, s), 255, np.uint32) if k < 12 else rng.integers(0, 256, (s, s), dtype=np.uint32)
        atlas[y0:y0 + s, x0:x0 + s] = np.where(m, (depth << 8) | radius, np.uint32(0xFFFFFFFF))
    return atlas


@pytest.mark.parametrize("ism_side,levels", [(s, l) for s in (8, 16, 32, 64)
                                             for l in sorted({1, 2, int(np.log2(s))})])
def test_pullpush_border_samples_bit_exact(ctx, ism_side, levels):
    side = 4 * ism_side
    print(f"[pullpush border] S={ism_side} L={levels}", flush=True)
    atlas = border_atlas(ism_side, seed=ism_side + levels)
    ism = mlgi.ImperfectShadowmap(ctx, side, N_VPL)
    ism.softrender.copy_from(atlas)
    p = mlgi.default_params(atlas_side=side, z_far=Z_FAR, pull_push_levels=levels)
    ism.pullpush(p)
    ctx.sync()
    got = ism.pushpull_result.numpy()
    ref = oracle.pullpush(atlas, ism_side, levels, Z_FAR)
    # the border samples fill part of their ISMs (not all of them: the
    # opposite border stays empty in the one-edge ISMs at one level)
    assert (ref < 65535).sum() > (atlas != 0xFFFFFFFF).sum()
    assert np.array_equal(got, ref), int((got != ref).sum())
