"""Named workloads and their synthetic inputs.

Shapes follow the benchmark configurations of the many-lights GI path:

  small   512x512 G-buffer, 4,000 random triangles, 64^2 RSM -> 1,024 VPLs,
          interleave 4x4, 256k points into 1,024 ISMs of 64^2 (2048^2 atlas),
          3 pull-push levels, 16x16 tiles
  1080p   1920x1080, 40,000 triangles, 128^2 RSM -> 4,096 VPLs, interleave
          8x8, 1M points into 4,096 ISMs of 64^2 (4096^2 atlas), 4 levels
  4k      3840x2160, 160,000 triangles, 128^2 RSM (stride 1) -> 16,384 VPLs,
          interleave 8x8, 8M points into 16,384 ISMs of 32^2 (4096^2 atlas),
          5 levels
  tiny    test-sized: 96x64, 500 triangles, 32^2 RSM -> 256 VPLs, 4x4,
          20k points into 256 ISMs of 32^2 (512^2 atlas), 3 levels

All inputs are generated on the host from a seed by the library's scene
tooling (procedural room, BVH ray caster, area-proportional sampler).
"""
from __future__ import annotations

import functools
from dataclasses import dataclass, replace

import numpy as np

from . import Camera, Scene, default_params, ism_indices_1d, Params


@dataclass(frozen=True)
class Workload:
    name: str
    n_random: int
    width: int
    height: int
    rsm: int
    n_vpl: int
    interleave: int
    atlas_side: int
    n_points: int
    levels: int
    tile: tuple = (16, 16)
    seed: int = 1
    light_seed: int = 7

    @property
    def ism_side(self) -> int:
        return self.atlas_side // ism_indices_1d(self.n_vpl)

    def params(self, **kw) -> Params:
        p = default_params(
            interleave=self.interleave, use_interleaving=1, atlas_side=self.atlas_side,
            pull_push_levels=self.levels, vpl_end_index=self.n_vpl, z_far=50.0,
            tile_w=self.tile[0], tile_h=self.tile[1])
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def with_(self, **kw) -> "Workload":
        return replace(self, **kw)


WORKLOADS = {
    "tiny": Workload("tiny", 500, 96, 64, 32, 256, 4, 512, 20_000, 3, (8, 8)),
    "small": Workload("small", 4_000, 512, 512, 64, 1024, 4, 2048, 262_144, 3),
    "1080p": Workload("1080p", 40_000, 1920, 1080, 128, 4096, 8, 4096, 1_000_000, 4),
    "4k": Workload("4k", 160_000, 3840, 2160, 128, 16384, 8, 4096, 8_000_000, 5),
}


@dataclass
class Inputs:
    wl: Workload
    gb_depth: np.ndarray       # H x W
    gb_normal: np.ndarray      # H x W x 3
    vp_inv: np.ndarray         # 16
    rsm_diffuse: np.ndarray    # R x R x 3
    rsm_normal: np.ndarray     # R x R x 3
    rsm_depth: np.ndarray      # R x R
    inv_bvp: np.ndarray        # 16
    points: np.ndarray         # P x 4
    bin_offsets: np.ndarray    # N + 1
    camera: Camera
    light: Camera


@functools.lru_cache(maxsize=8)
def _scene(seed: int, n_random: int) -> Scene:
    return Scene(seed, n_random)


@functools.lru_cache(maxsize=8)
def build_inputs(wl: Workload, threads: int = 0) -> Inputs:
    scene = _scene(wl.seed, wl.n_random)
    cam = Camera.default(wl.width / wl.height)
    depth, normal = scene.render_gbuffer(cam, wl.width, wl.height, threads)
    _, _, vp_inv = cam.matrices()
    light = Camera.light(wl.light_seed)
    dif, nrm, dep, inv_bvp = scene.render_rsm(light, wl.rsm, wl.rsm, threads)
    pts, off = scene.sample_points(wl.n_points, wl.n_vpl, wl.n_vpl, wl.seed * 7919 + 13, threads)
    return Inputs(wl, depth, normal, vp_inv, dif, nrm, dep, inv_bvp, pts, off, cam, light)
