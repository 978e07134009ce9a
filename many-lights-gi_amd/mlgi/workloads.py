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


# ------------------------------------------------------------------ batch64
BATCH64_SIZES = (128, 256, 512, 1024)
BATCH64_VPLS = (256, 1024, 2048, 4096)


@dataclass
class BatchFrameSpec:
    size: int        # square G-buffer side
    n_vpl: int
    ism_side: int    # 32 for <= 1,024 VPLs, else 64
    levels: int      # pull-push levels (3 at 32^2 ISMs, 4 at 64^2)
    light_seed: int


def batch64_specs(seed: int = 2024, frames: int = 64) -> list:
    """The batched workload's frames: resolutions drawn from 128^2..1024^2 and
    VPL counts from 256..4,096 (ISMs of 32^2 up to 1,024 VPLs, else 64^2),
    each with its own light pose."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(frames):
        s = int(rng.choice(BATCH64_SIZES))
        n = int(rng.choice(BATCH64_VPLS))
        ism = 32 if n <= 1024 else 64
        out.append(BatchFrameSpec(s, n, ism, 3 if ism == 32 else 4, 5000 + i))
    return out


class Batch64:
    """The batched workload on one context: 64 frames of the 40k-triangle
    room (random square resolutions and VPL counts, a light pose each),
    G-buffers and RSMs ray-cast on the device, 1M points per VPL count,
    mixed precision, interleave and gather tile picked per shape by the
    autotuner. `frames` holds (GIStage, rsm, points, gbuffer) per frame; the
    host inputs of frame i (for the oracle) come from host_inputs(i)."""

    SCENE = (1, 40_000)
    POINTS = 1_000_000
    POINT_SEED = 11
    RSM = 128

    def __init__(self, ctx, seed: int = 2024, frames: int = 64, precision=None):
        from . import PRECISION_MIXED, GIStage, autotune, gbuffer_desc, points_desc, rsm_desc
        self.ctx = ctx
        self.specs = batch64_specs(seed, frames)
        self.scene = _scene(*self.SCENE)
        self.dev = self.scene.upload(ctx)
        self.points_host, self.points, self.gbufs = {}, {}, {}
        self.frames, self.tune = [], []
        prec = PRECISION_MIXED if precision is None else precision
        for i, sp in enumerate(self.specs):
            n, s = sp.n_vpl, sp.size
            if n not in self.points:
                rec, off = self.scene.sample_points(self.POINTS, n, n, self.POINT_SEED)
                self.points_host[n] = (rec, off)
                self.points[n] = points_desc(ctx.upload(rec), ctx.upload(off))
            if s not in self.gbufs:
                d, nm = ctx.empty((s, s), np.float32), ctx.empty((s, s, 3), np.float32)
                inv = self.dev.render_gbuffer(Camera.default(1.0), d, nm)
                self.gbufs[s] = (d, nm, gbuffer_desc(d, nm, inv))
            p = default_params(atlas_side=ism_indices_1d(n) * sp.ism_side, pull_push_levels=sp.levels,
                               vpl_end_index=n, z_far=50.0, precision=prec)
            st = GIStage(ctx, p, n, s, s)
            r = self.RSM
            dif, nrm, dep = (ctx.empty((r, r, 3), np.float32), ctx.empty((r, r, 3), np.float32),
                             ctx.empty((r, r), np.float32))
            inv = self.dev.render_rsm(Camera.light(sp.light_seed), dif, nrm, dep)
            rsm = rsm_desc(dif, nrm, dep, inv)
            st.rsm_buffers = (dif, nrm, dep)  # keep the RSM alive with its stage
            self.tune.append(autotune(ctx, st, rsm, self.points[n], self.gbufs[s][2]))
            self.frames.append((st, rsm, self.points[n], self.gbufs[s][2]))

    def host_inputs(self, i: int) -> dict:
        """Frame i's inputs rebuilt on the host (the device renders equal the
        host renderer's bit for bit)."""
        sp = self.specs[i]
        cam = Camera.default(1.0)
        depth, normal = self.scene.render_gbuffer(cam, sp.size, sp.size)
        dif, nrm, dep, inv = self.scene.render_rsm(Camera.light(sp.light_seed), self.RSM, self.RSM)
        rec, off = self.points_host[sp.n_vpl]
        return {"gb_depth": depth, "gb_normal": normal, "vp_inv": cam.matrices()[2], "rsm": (dif, nrm, dep, inv),
                "points": rec, "bin_offsets": off}

    def close(self) -> None:
        self.dev.close()
