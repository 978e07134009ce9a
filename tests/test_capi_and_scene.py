"""C ABI surface and the host-side scene tooling (CPU only).

* libmlgi.so loads and exports every function include/*.h declares;
* mlgi_params_default matches GIStage::initialize (GIStage.cpp:215-232) and
  the ImperfectShadowmap constants (ImperfectShadowmap.cpp:34-35);
* the VPL permutation equals the committed fixtures written by
  tests/golden/make_shuffle.cpp (std::shuffle + std::mt19937(1979982));
* the scene generator is deterministic, its points carry the reference's
  record (packUnorm4x8 normal + radius) and bin hash, and the ray-cast
  G-buffer / RSM hold valid window depths and unit normals;
* without a device, context creation fails cleanly (no crash).
"""
import math
import re
from pathlib import Path

import numpy as np
import pytest

import mlgi
import oracle
from mlgi import _native as N
from mlgi import workloads as W

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = Path(__file__).resolve().parent / "golden"


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(mlgi_\w+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(N.lib, n)]
    assert not missing, missing
    # and the Python binding covers exactly the declared surface
    assert set(N.SIGNATURES) == set(names)


def test_params_defaults_match_reference():
    p = mlgi.default_params()
    assert p.light_intensity == 5.0
    assert p.gi_intensity_factor == 3000.0
    assert p.vpl_clamping_value == np.float32(0.001)
    assert (p.vpl_start_index, p.scale_isms, p.points_only_into_scaled_isms) == (0, 0, 0)
    assert p.tess_level_factor == 2.0
    assert (p.use_push_pull, p.enable_shadowing, p.use_interleaving, p.shuffle_lights) == (1, 1, 1, 1)
    assert p.interleave == 4
    assert (p.atlas_side, p.pull_push_levels) == (2048, 6)


@pytest.mark.parametrize("n", [256, 1024, 2048, 4096, 16384])
def test_shuffle_matches_fixture(n):
    want = np.fromfile(GOLDEN / f"shuffle_{n}.i32", np.int32)
    got = mlgi.shuffle_indices(n)
    assert np.array_equal(got, want)
    assert np.array_equal(np.sort(got), np.arange(n))


def test_ism_grid_side_matches_oracle():
    for n in list(range(1, 300)) + [1000, 1024, 1025, 2048, 4096, 16384, 16385]:
        assert mlgi.ism_indices_1d(n) == oracle.ism_indices_1d(n)


def test_no_device_fails_cleanly():
    if mlgi.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(mlgi.MlgiError) as e:
        mlgi.Context(0)
    assert e.value.code == N.MLGI_ERR_NO_DEVICE


# ------------------------------------------------------------------ scene
def test_scene_deterministic_and_shaped():
    a, b = mlgi.Scene(3, 200), mlgi.Scene(3, 200)
    va, na, aa = a.triangles()
    vb, nb, ab = b.triangles()
    assert a.triangle_count == 212
    assert np.array_equal(va, vb) and np.array_equal(na, nb) and np.array_equal(aa, ab)
    assert np.all((va >= 0) & (va <= 10))
    assert np.all((aa >= 0.2) & (aa <= 0.8))
    assert np.allclose(np.linalg.norm(na, axis=1), 1.0, atol=1e-6)
    edges = np.linalg.norm(va[12:, 1] - va[12:, 0], axis=1)
    assert edges.min() >= 0.1 - 1e-6 and edges.max() <= 0.5 + 1e-6
    area = 0.5 * np.linalg.norm(np.cross(va[:, 1] - va[:, 0], va[:, 2] - va[:, 0]), axis=1).sum()
    assert abs(area - a.total_area) < 1e-3
    assert mlgi.Scene(4, 200).triangles()[0][20, 0, 0] != va[20, 0, 0]


def test_points_record_and_bins():
    scene = mlgi.Scene(1, 300)
    n_bins, ism_count, P = 256, 256, 20000
    pts, off = scene.sample_points(P, n_bins, ism_count, seed=9)
    assert off[0] == 0 and off[-1] == P and np.all(np.diff(off.astype(np.int64)) >= 0)
    bits = pts[:, 3].view(np.uint32)
    r_world = math.sqrt(scene.total_area / (P * math.pi)) * math.sqrt(ism_count)
    assert np.all((bits >> 24) == round(r_world / 25 * 255))
    tv, tn, _ = scene.triangles()
    for b in (0, 17, 255):
        for i in range(off[b], min(off[b + 1], off[b] + 20)):
            x = pts[i, :3]
            # int(random(vec3(pos)) * bins) (random.glsl:62-66)
            assert int(oracle.random3(*map(float, x)) * n_bins) == b
            u = oracle.unpack_unorm4x8(int(bits[i]))
            n = u[:3] * 2 - 1
            assert np.min(np.linalg.norm(tn - n, axis=1)) < 0.012  # one of the triangle normals
    # area-proportional: the room (600 of the area) holds its share of points
    inside_walls = np.isclose(pts[:, :3], 0.0, atol=1e-5) | np.isclose(pts[:, :3], 10.0, atol=1e-5)
    frac = inside_walls.any(axis=1).mean()
    assert abs(frac - 600.0 / scene.total_area) < 0.02


def test_gbuffer_and_rsm_are_valid():
    wl = W.WORKLOADS["tiny"]
    inp = W.build_inputs(wl)
    assert inp.gb_depth.shape == (wl.height, wl.width)
    assert np.all((inp.gb_depth > 0.0) & (inp.gb_depth < 1.0))
    n = inp.gb_normal * 2 - 1
    assert np.allclose(np.linalg.norm(n, axis=2), 1.0, atol=1e-5)
    assert np.all((inp.rsm_depth > 0.0) & (inp.rsm_depth < 1.0))
    # spot cone: RSM corners lie outside the 45 degree cone
    r = wl.rsm
    for y, x in [(0, 0), (0, r - 1), (r - 1, 0), (r - 1, r - 1)]:
        assert np.all(inp.rsm_diffuse[y, x] == 0.0)
    assert np.all(inp.rsm_diffuse[r // 2, r // 2] > 0.0)
    # VPLs unprojected from the RSM lie in the room, up to the half-texel
    # offset of unprojecting the texel's corner uv with its centre's depth
    # (vpl_processor.comp:50-57): 2 tan(22.5 deg) * 14.2 / 32 / 2 < 0.2
    pos, _, _, _ = oracle.vpl_extract(inp.rsm_diffuse, inp.rsm_normal, inp.rsm_depth, inp.inv_bvp, 5.0,
                                      None, wl.n_vpl)
    assert np.all((pos > -0.2) & (pos < 10.2))


def test_camera_matrices_invert():
    cam = mlgi.Camera.default(16 / 9)
    v, p, inv = cam.matrices()
    vp = p.reshape(4, 4).T.astype(np.float64) @ v.reshape(4, 4).T.astype(np.float64)
    assert np.allclose(inv.reshape(4, 4).T.astype(np.float64) @ vp, np.eye(4), atol=1e-4)
