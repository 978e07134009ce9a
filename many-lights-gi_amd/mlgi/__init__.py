"""Python mirror of the mlgi C ABI (libmlgi.so).

Device memory and streams live behind the C ABI; this module only moves
numpy arrays in and out and names things the way the reference does:

    VPLProcessor.process(rsm, light_intensity, shuffle_lights)
        VPLProcessor.h:22, VPLProcessor.cpp:73-101
    ImperfectShadowmap.process(points, vpl_processor, vpl_start_index,
        vpl_end_index, scale_isms, points_only_into_scaled_isms,
        tess_level_factor, use_push_pull, z_far)
        ImperfectShadowmap.h:30-39, ImperfectShadowmap.cpp:184-192
    GIStage.compute_final_gathering() / GIStage.process()
        GIStage.cpp:261-306, :346-418

Errors raise MlgiError with the library's message; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from ._native import MlgiError, Params, PRECISION_FP32, PRECISION_MIXED, VPL_RECORD_FLOATS

__all__ = [
    "MlgiError", "Params", "PRECISION_FP32", "PRECISION_MIXED", "VPL_RECORD_FLOATS",
    "Context", "DeviceArray", "Scene", "Camera", "VPLProcessor", "ImperfectShadowmap",
    "GIStage", "default_params", "shuffle_indices", "ism_indices_1d", "device_count",
    "vpl_fields", "fill_flux_halves", "ism_count",
]

lib = N.lib


def vpl_fields(rec: np.ndarray) -> dict:
    """Named views of a host copy of VPL records (count x VPL_RECORD_FLOATS)."""
    return {
        "position": rec[:, 0:3],
        "plane": rec[:, 3],
        "normal": rec[:, 4:7],
        "packed_normal": rec[:, 19].view(np.uint32),
        "zero_flux_flags": rec[:, 7].view(np.uint32),
        "s": rec[:, 8:11],
        "u": rec[:, 12:15],
        "flux": rec[:, 16:19],
        "flux_halves": rec[:, [11, 15]].view(np.uint32),
    }


def fill_flux_halves(rec: np.ndarray) -> np.ndarray:
    """Writes the record slots VPL extraction derives from the flux (slots
    16..18): 11 and 15, the flux as binary16 bits (round to nearest even:
    r | g << 16 and b), and 7, the zero-flux flags (bit 0: the floats are all
    +-0; bit 1: the halves are all zero). Records built by hand need them for
    the gather."""
    h = rec[:, 16:19].astype(np.float16).view(np.uint16).astype(np.uint32)
    rg, b = h[:, 0] | (h[:, 1] << 16), h[:, 2]
    rec[:, 11] = rg.view(np.float32)
    rec[:, 15] = b.view(np.float32)
    fz = ((np.ascontiguousarray(rec[:, 16:19]).view(np.uint32) & 0x7FFFFFFF) == 0).all(axis=1)
    rec[:, 7] = (fz.astype(np.uint32) | (((rg | b) == 0).astype(np.uint32) << 1)).view(np.float32)
    return rec


def default_params(**kw) -> Params:
    p = Params()
    lib.mlgi_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def shuffle_indices(n: int, seed: int = 1979982) -> np.ndarray:
    out = np.empty(n, np.int32)
    rc = lib.mlgi_shuffle_indices(n, seed, out.ctypes.data)
    if rc:
        raise MlgiError(rc, "mlgi_shuffle_indices")
    return out


def ism_indices_1d(n: int) -> int:
    return int(lib.mlgi_ism_indices_1d(n))


def device_count() -> int:
    n = C.c_int32(0)
    lib.mlgi_device_count(C.byref(n))
    return n.value


class Context:
    """One HIP device + stream (mlgi_context)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib.mlgi_context_create(device, C.byref(h))
        if rc:
            raise MlgiError(rc, f"mlgi_context_create(device={device}) failed")
        self.h = h
        self.device = device
        self._live: set[int] = set()

    def check(self, rc: int, what: str = "") -> None:
        if rc:
            msg = lib.mlgi_last_error(self.h)
            raise MlgiError(rc, f"{what}: {msg.decode() if msg else ''}")

    def close(self) -> None:
        if getattr(self, "h", None):  # None also when __init__ raised
            for ptr in list(self._live):
                lib.mlgi_free(self.h, C.c_void_p(ptr))
            self._live.clear()
            lib.mlgi_context_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- memory
    def empty(self, shape, dtype) -> "DeviceArray":
        return DeviceArray(self, shape, dtype)

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        d.copy_from(a)
        return d

    def sync(self) -> None:
        self.check(lib.mlgi_sync(self.h), "mlgi_sync")

    def wait_for(self, other: "Context") -> None:
        """Order this context's stream after the work enqueued so far on
        `other`'s (mlgi_context_wait; the host does not block)."""
        self.check(lib.mlgi_context_wait(self.h, other.h), "mlgi_context_wait")

    def probe_valu(self) -> float:
        """Measured FP32 vector FMA rate (lane FMAs per second)."""
        v = C.c_double(0.0)
        self.check(lib.mlgi_probe_valu(self.h, C.byref(v)), "mlgi_probe_valu")
        return float(v.value)

    def check_division(self, seed: int = 0) -> tuple:
        """(reciprocal mismatches, quotient mismatches, quotients tested) of
        the splat's division form against IEEE division on this device
        (mlgi_check_division); both counts are 0 where the atlas is exact."""
        r, q, n = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        self.check(lib.mlgi_check_division(self.h, seed & 0xFFFFFFFF, C.byref(r), C.byref(q), C.byref(n)),
                   "mlgi_check_division")
        return int(r.value), int(q.value), int(n.value)

    # -- profiling (PerfCounter analogue)
    def profile(self, on: bool = True) -> None:
        self.check(lib.mlgi_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self) -> None:
        self.check(lib.mlgi_profile_reset(self.h))

    def profile_read(self) -> dict:
        t = N.StageTimes()
        self.check(lib.mlgi_profile_read(self.h, C.byref(t)), "mlgi_profile_read")
        return {k: getattr(t, k) for k, _ in N.StageTimes._fields_}


class DeviceArray:
    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        h = C.c_void_p()
        ctx.check(lib.mlgi_malloc(ctx.h, max(self.nbytes, 1), C.byref(h)), "mlgi_malloc")
        self.ptr = h.value
        ctx._live.add(self.ptr)

    def copy_from(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"size mismatch {a.nbytes} != {self.nbytes}")
        self.ctx.check(lib.mlgi_memcpy_htod(self.ctx.h, C.c_void_p(self.ptr), a.ctypes.data, a.nbytes),
                       "mlgi_memcpy_htod")

    def copy_from_device(self, src: "DeviceArray") -> None:
        """Stream-ordered device-to-device copy of src (same byte size)."""
        if src.nbytes != self.nbytes:
            raise ValueError(f"size mismatch {src.nbytes} != {self.nbytes}")
        self.ctx.check(lib.mlgi_memcpy_dtod(self.ctx.h, C.c_void_p(self.ptr), C.c_void_p(src.ptr), self.nbytes),
                       "mlgi_memcpy_dtod")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        self.ctx.check(lib.mlgi_memcpy_dtoh(self.ctx.h, out.ctypes.data, C.c_void_p(self.ptr), self.nbytes),
                       "mlgi_memcpy_dtoh")
        return out

    def free(self) -> None:
        if self.ptr and self.ctx.h:
            self.ctx._live.discard(self.ptr)
            lib.mlgi_free(self.ctx.h, C.c_void_p(self.ptr))
        self.ptr = None


# ------------------------------------------------------------------ scene
@dataclass
class Camera:
    eye: tuple
    center: tuple
    up: tuple
    fovy_deg: float
    aspect: float
    z_near: float
    z_far: float

    def to_c(self) -> N.Camera:
        c = N.Camera()
        c.eye[:] = self.eye
        c.center[:] = self.center
        c.up[:] = self.up
        c.fovy_deg, c.aspect, c.z_near, c.z_far = self.fovy_deg, self.aspect, self.z_near, self.z_far
        return c

    @staticmethod
    def from_c(c: N.Camera) -> "Camera":
        return Camera(tuple(c.eye), tuple(c.center), tuple(c.up), c.fovy_deg, c.aspect, c.z_near, c.z_far)

    @staticmethod
    def default(aspect: float) -> "Camera":
        c = N.Camera()
        lib.mlgi_scene_default_camera(aspect, C.byref(c))
        return Camera.from_c(c)

    @staticmethod
    def light(seed: int) -> "Camera":
        c = N.Camera()
        lib.mlgi_scene_light_pose(seed, C.byref(c))
        return Camera.from_c(c)

    def matrices(self):
        v = np.empty(16, np.float32)
        p = np.empty(16, np.float32)
        inv = np.empty(16, np.float32)
        rc = lib.mlgi_camera_matrices(C.byref(self.to_c()), v.ctypes.data, p.ctypes.data, inv.ctypes.data)
        if rc:
            raise MlgiError(rc, "mlgi_camera_matrices")
        return v, p, inv


class Scene:
    """Seeded procedural scene (host): room + random triangles."""

    def __init__(self, seed: int, n_random: int):
        h = C.c_void_p()
        rc = lib.mlgi_scene_create(seed, n_random, C.byref(h))
        if rc:
            raise MlgiError(rc, "mlgi_scene_create")
        self.h = h
        self.seed = seed
        self.n_random = n_random

    def __del__(self):
        if getattr(self, "h", None):
            lib.mlgi_scene_destroy(self.h)
            self.h = None

    @property
    def triangle_count(self) -> int:
        return int(lib.mlgi_scene_triangle_count(self.h))

    @property
    def total_area(self) -> float:
        return float(lib.mlgi_scene_total_area(self.h))

    def triangles(self):
        n = self.triangle_count
        v = np.empty((n, 3, 3), np.float32)
        nr = np.empty((n, 3), np.float32)
        a = np.empty((n, 3), np.float32)
        lib.mlgi_scene_triangles(self.h, v.ctypes.data, nr.ctypes.data, a.ctypes.data)
        return v, nr, a

    def render_gbuffer(self, cam: Camera, width: int, height: int, threads: int = 0):
        depth = np.empty((height, width), np.float32)
        normal = np.empty((height, width, 3), np.float32)
        rc = lib.mlgi_scene_render_gbuffer(self.h, C.byref(cam.to_c()), width, height, threads,
                                           depth.ctypes.data, normal.ctypes.data)
        if rc:
            raise MlgiError(rc, "mlgi_scene_render_gbuffer")
        return depth, normal

    def render_rsm(self, light: Camera, width: int, height: int, threads: int = 0):
        diffuse = np.empty((height, width, 3), np.float32)
        normal = np.empty((height, width, 3), np.float32)
        depth = np.empty((height, width), np.float32)
        inv = np.empty(16, np.float32)
        rc = lib.mlgi_scene_render_rsm(self.h, C.byref(light.to_c()), width, height, threads,
                                       diffuse.ctypes.data, normal.ctypes.data, depth.ctypes.data,
                                       inv.ctypes.data)
        if rc:
            raise MlgiError(rc, "mlgi_scene_render_rsm")
        return diffuse, normal, depth, inv

    def render_surface(self, cam: Camera, width: int, height: int, threads: int = 0):
        """(albedo (h, w, 3), VSM moments (h, w, 2)) of the view (mlgi_scene_render_surface)."""
        albedo = np.empty((height, width, 3), np.float32)
        moments = np.empty((height, width, 2), np.float32)
        rc = lib.mlgi_scene_render_surface(self.h, C.byref(cam.to_c()), width, height, threads, albedo.ctypes.data,
                                           moments.ctypes.data)
        if rc:
            raise MlgiError(rc, "mlgi_scene_render_surface")
        return albedo, moments

    def upload(self, ctx: "Context") -> "DeviceScene":
        """The scene's BVH on the device (mlgi_scene_upload)."""
        return DeviceScene(ctx, self)

    def sample_points(self, n_points: int, n_bins: int, ism_count: int, seed: int, threads: int = 0):
        rec = np.empty((n_points, 4), np.float32)
        off = np.empty(n_bins + 1, np.uint32)
        rc = lib.mlgi_scene_sample_points(self.h, n_points, n_bins, ism_count, seed, threads,
                                          rec.ctypes.data, off.ctypes.data)
        if rc:
            raise MlgiError(rc, "mlgi_scene_sample_points")
        return rec, off


class DeviceScene:
    """Device G-buffer / RSM producer: the host renderer's ray caster on the GPU."""

    def __init__(self, ctx: "Context", scene: Scene):
        self.ctx, self.scene = ctx, scene
        h = C.c_void_p()
        ctx.check(lib.mlgi_scene_upload(ctx.h, scene.h, C.byref(h)), "mlgi_scene_upload")
        self.h = h

    def render_gbuffer(self, cam: Camera, depth: "DeviceArray", normal: "DeviceArray"):
        """Into device arrays (h, w) and (h, w, 3); returns inverse(P*V) (host)."""
        height, width = depth.shape
        inv = np.empty(16, np.float32)
        self.ctx.check(lib.mlgi_scene_device_render_gbuffer(self.ctx.h, self.h, C.byref(cam.to_c()), width, height,
                                                            C.c_void_p(depth.ptr), C.c_void_p(normal.ptr),
                                                            inv.ctypes.data), "mlgi_scene_device_render_gbuffer")
        return inv

    def sample_points(self, n_points: int, n_bins: int, ism_count: int, seed: int):
        """Point set on the device: (records (n, 4) f32, bin_offsets (n_bins + 1) u32) device arrays."""
        rec = self.ctx.empty((max(n_points, 1), 4), np.float32)
        off = self.ctx.empty((n_bins + 1,), np.uint32)
        self.ctx.check(lib.mlgi_scene_device_sample_points(self.ctx.h, self.h, n_points, n_bins, ism_count, seed,
                                                           C.c_void_p(rec.ptr), C.c_void_p(off.ptr)),
                       "mlgi_scene_device_sample_points")
        return rec, off

    def render_surface(self, cam: Camera, albedo: "DeviceArray" = None, moments: "DeviceArray" = None):
        """Albedo (h, w, 3) and / or VSM moments (h, w, 2) into device arrays."""
        ref = albedo if albedo is not None else moments
        height, width = ref.shape[:2]
        self.ctx.check(lib.mlgi_scene_device_render_surface(
            self.ctx.h, self.h, C.byref(cam.to_c()), width, height,
            C.c_void_p(albedo.ptr) if albedo is not None else None,
            C.c_void_p(moments.ptr) if moments is not None else None), "mlgi_scene_device_render_surface")

    def render_rsm(self, light: Camera, diffuse: "DeviceArray", normal: "DeviceArray", depth: "DeviceArray"):
        """Into device arrays; returns inverse(bias*P*V) (host)."""
        height, width = depth.shape
        inv = np.empty(16, np.float32)
        self.ctx.check(lib.mlgi_scene_device_render_rsm(self.ctx.h, self.h, C.byref(light.to_c()), width, height,
                                                        C.c_void_p(diffuse.ptr), C.c_void_p(normal.ptr),
                                                        C.c_void_p(depth.ptr), inv.ctypes.data),
                       "mlgi_scene_device_render_rsm")
        return inv

    def close(self) -> None:
        if getattr(self, "h", None):  # None also when __init__ raised
            lib.mlgi_scene_device_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


# ------------------------------------------------------------ descriptors
def rsm_desc(diffuse: DeviceArray, normal: DeviceArray, depth: DeviceArray, inv_bvp) -> N.Rsm:
    r = N.Rsm()
    r.diffuse, r.face_normal, r.depth = diffuse.ptr, normal.ptr, depth.ptr
    r.height, r.width = depth.shape
    r.inv_biased_view_projection[:] = [float(x) for x in np.asarray(inv_bvp, np.float32)]
    return r


def gbuffer_desc(depth: DeviceArray, normal: DeviceArray, vp_inv) -> N.GBuffer:
    g = N.GBuffer()
    g.depth, g.face_normal = depth.ptr, normal.ptr
    g.height, g.width = depth.shape
    g.view_projection_inverse[:] = [float(x) for x in np.asarray(vp_inv, np.float32)]
    return g


def points_desc(records: DeviceArray, bin_offsets: DeviceArray) -> N.Points:
    p = N.Points()
    p.records, p.bin_offsets = records.ptr, bin_offsets.ptr
    p.count = records.shape[0]
    p.n_bins = bin_offsets.shape[0] - 1
    return p


# --------------------------------------------------------- stage classes
class VPLProcessor:
    """Owns the VPL buffer (vplBuffer + packedVplBuffer of the reference)."""

    def __init__(self, ctx: Context, vpl_count: int):
        self.ctx = ctx
        self.count = vpl_count
        self.records = ctx.empty((vpl_count, VPL_RECORD_FLOATS), np.float32)

    def desc(self) -> N.Vpls:
        v = N.Vpls()
        v.records, v.count = self.records.ptr, self.count
        return v

    def process(self, rsm: N.Rsm, light_intensity: float, shuffle_lights: bool, params: Params | None = None):
        p = params if params is not None else default_params()
        p.light_intensity = light_intensity
        p.shuffle_lights = 1 if shuffle_lights else 0
        self.ctx.check(lib.mlgi_vpl_extract(self.ctx.h, C.byref(rsm), C.byref(p), C.byref(self.desc())),
                       "mlgi_vpl_extract")

    def host_records(self) -> np.ndarray:
        return self.records.numpy()


def ism_count(params: Params, vpl_count: int) -> int:
    """ISMs in the atlas (ImperfectShadowmap.cpp:188): one per VPL, or with
    scaleISMs one per VPL of [vplStartIndex, vplEndIndex)."""
    if params.scale_isms:
        end = params.vpl_end_index if params.vpl_end_index > 0 else vpl_count
        return end - params.vpl_start_index
    return vpl_count


class ImperfectShadowmap:
    """Owns the atlas ("ISM softrender") and the pull-push result."""

    def __init__(self, ctx: Context, atlas_side: int, vpl_count: int):
        self.ctx = ctx
        self.side = atlas_side
        self.vpl_count = vpl_count
        self.ism_side = atlas_side // ism_indices_1d(vpl_count)
        self.softrender = ctx.empty((atlas_side, atlas_side), np.uint32)
        self.pushpull_result = ctx.empty((atlas_side, atlas_side), np.uint16)

    def layout(self, params: Params) -> None:
        """ISM side for the ISM count params give (scaleISMs sizes the ISMs
        for the VPL range)."""
        self.ism_side = self.side // ism_indices_1d(ism_count(params, self.vpl_count))

    def atlas_desc(self) -> N.Atlas:
        a = N.Atlas()
        a.texels, a.side, a.ism_side = self.softrender.ptr, self.side, self.ism_side
        return a

    def depth_desc(self) -> N.IsmDepth:
        d = N.IsmDepth()
        d.texels, d.side, d.ism_side = self.pushpull_result.ptr, self.side, self.ism_side
        return d

    def render(self, points: N.Points, vpl_processor: VPLProcessor, params: Params) -> None:
        self.layout(params)
        self.ctx.check(lib.mlgi_ism_splat(self.ctx.h, C.byref(points), C.byref(vpl_processor.desc()),
                                          C.byref(params), C.byref(self.atlas_desc())), "mlgi_ism_splat")

    def spill_stats(self) -> tuple[int, bool]:
        """(spills, overflowed) of the context's last ISM-major splat."""
        n, o = C.c_uint32(0), C.c_int32(0)
        self.ctx.check(lib.mlgi_splat_spill_stats(self.ctx.h, C.byref(n), C.byref(o)), "mlgi_splat_spill_stats")
        return int(n.value), bool(o.value)

    def splat_count(self) -> int:
        """(point, ISM) splats of the context's last ISM-major splat (mlgi_splat_count)."""
        n = C.c_uint64(0)
        self.ctx.check(lib.mlgi_splat_count(self.ctx.h, C.byref(n)), "mlgi_splat_count")
        return int(n.value)

    def pullpush(self, params: Params) -> None:
        self.ctx.check(lib.mlgi_ism_pullpush(self.ctx.h, C.byref(self.atlas_desc()), C.byref(params),
                                             C.byref(self.depth_desc())), "mlgi_ism_pullpush")

    def process(self, points: N.Points, vpl_processor: VPLProcessor, vpl_start_index: int,
                vpl_end_index: int, scale_isms: bool, points_only_into_scaled_isms: bool,
                tess_level_factor: float, use_push_pull: bool, z_far: float,
                params: Params | None = None) -> None:
        p = params if params is not None else default_params()
        p.vpl_start_index, p.vpl_end_index = vpl_start_index, vpl_end_index
        p.scale_isms = int(scale_isms)
        p.points_only_into_scaled_isms = int(points_only_into_scaled_isms)
        p.tess_level_factor = tess_level_factor
        p.use_push_pull = int(use_push_pull)
        p.z_far = z_far
        p.atlas_side = self.side
        self.layout(p)
        self.ctx.check(lib.mlgi_ism_process(self.ctx.h, C.byref(points), C.byref(vpl_processor.desc()),
                                            C.byref(p), C.byref(self.atlas_desc()),
                                            C.byref(self.depth_desc())), "mlgi_ism_process")


class GIStage:
    """Per-frame GI driver: VPLP -> ISM (splat + pull-push) -> FG."""

    def __init__(self, ctx: Context, params: Params, vpl_count: int, width: int, height: int):
        self.ctx = ctx
        self.params = params
        self.vpl_processor = VPLProcessor(ctx, vpl_count)
        self.ism = ImperfectShadowmap(ctx, params.atlas_side, vpl_count)
        self.gi_buffer = ctx.empty((height, width, 3), np.float32)
        self.width, self.height = width, height

    def image_desc(self) -> N.Image:
        im = N.Image()
        im.rgb, im.width, im.height = self.gi_buffer.ptr, self.width, self.height
        return im

    def compute_final_gathering(self, gbuffer: N.GBuffer) -> None:
        self.ctx.check(lib.mlgi_final_gather(self.ctx.h, C.byref(gbuffer), C.byref(self.vpl_processor.desc()),
                                             C.byref(self.ism.depth_desc()), C.byref(self.params),
                                             C.byref(self.image_desc())), "mlgi_final_gather")

    def process(self, rsm: N.Rsm, points: N.Points, gbuffer: N.GBuffer) -> None:
        self.ism.layout(self.params)
        v, a, d, im = self.vpl_processor.desc(), self.ism.atlas_desc(), self.ism.depth_desc(), self.image_desc()
        self.ctx.check(lib.mlgi_frame(self.ctx.h, C.byref(rsm), C.byref(points), C.byref(gbuffer),
                                      C.byref(self.params), C.byref(v), C.byref(a), C.byref(d),
                                      C.byref(im)), "mlgi_frame")

    def capture(self, rsm: N.Rsm, points: N.Points, gbuffer: N.GBuffer) -> "FrameGraph":
        """Capture one frame (process()) into a replayable HIP graph."""
        v, a, d, im = self.vpl_processor.desc(), self.ism.atlas_desc(), self.ism.depth_desc(), self.image_desc()
        h = C.c_void_p()
        self.ctx.check(lib.mlgi_graph_capture_frame(self.ctx.h, C.byref(rsm), C.byref(points), C.byref(gbuffer),
                                                    C.byref(self.params), C.byref(v), C.byref(a), C.byref(d),
                                                    C.byref(im), C.byref(h)), "mlgi_graph_capture_frame")
        return FrameGraph(self.ctx, h)

    def blur(self, gbuffer: N.GBuffer, projection) -> None:
        """GIStage::blur: geometry-aware x/y blur of gi_buffer into gi_blur_final."""
        if getattr(self, "gi_blur_final", None) is None:
            self.gi_blur_final = self.ctx.empty((self.height, self.width, 3), np.float32)
        out = N.Image()
        out.rgb, out.width, out.height = self.gi_blur_final.ptr, self.width, self.height
        self._projection = np.ascontiguousarray(projection, np.float32)
        self.ctx.check(lib.mlgi_gi_blur(self.ctx.h, C.byref(gbuffer), self._projection.ctypes.data,
                                        C.byref(self.image_desc()), C.byref(out)), "mlgi_gi_blur")

    def count_lit_pairs(self, gbuffer: N.GBuffer) -> int:
        out = C.c_uint64(0)
        self.ctx.check(lib.mlgi_count_lit_pairs(self.ctx.h, C.byref(gbuffer), C.byref(self.vpl_processor.desc()),
                                                C.byref(self.params), C.byref(out)), "mlgi_count_lit_pairs")
        return int(out.value)


def deferred_composite(ctx: "Context", depth, normal, diffuse, gi, vsm, out, *, vp_inv, biased_light_vp, eye,
                       light_pos, inv_light_dir, light_intensity: float, exposure: float = 1.0, specular=None,
                       occlusion=None) -> None:
    """DeferredShadingStage::process on device arrays: depth (h, w), normal / diffuse / gi / out (h, w, 3),
    vsm (vh, vw, 2), optional specular (h, w, 3) and occlusion (h, w)."""
    h, w = depth.shape
    vh, vw = vsm.shape[:2]
    i = N.CompositeInputs()
    i.depth, i.normal, i.diffuse, i.gi, i.vsm = depth.ptr, normal.ptr, diffuse.ptr, gi.ptr, vsm.ptr
    i.specular = specular.ptr if specular is not None else None
    i.occlusion = occlusion.ptr if occlusion is not None else None
    i.width, i.height, i.vsm_width, i.vsm_height = w, h, vw, vh
    p = N.CompositeParams()
    p.view_projection_inverse[:] = [float(x) for x in np.ravel(vp_inv)]
    p.biased_light_view_projection[:] = [float(x) for x in np.ravel(biased_light_vp)]
    p.camera_eye[:] = [float(x) for x in eye]
    p.light_position[:] = [float(x) for x in light_pos]
    p.inverse_light_direction[:] = [float(x) for x in inv_light_dir]
    p.light_intensity, p.exposure = light_intensity, exposure
    o = N.Image()
    o.rgb, o.width, o.height = out.ptr, w, h
    ctx.check(lib.mlgi_deferred_composite(ctx.h, C.byref(i), C.byref(p), C.byref(o)), "mlgi_deferred_composite")


class Batch:
    """Many GI frames as one set of batched launches, one per stage.

    `frames` is a list of (GIStage, rsm, points, gbuffer): each stage owns its
    parameters and its VPL / atlas / ISM / image buffers.
    """

    def __init__(self, ctx: Context, frames):
        self.ctx = ctx
        self.frames = list(frames)
        arr = (N.FrameDesc * len(self.frames))()
        for i, (st, rsm, pts, gb) in enumerate(self.frames):
            arr[i] = frame_desc(st, rsm, pts, gb)
        self._descs = arr
        h = C.c_void_p()
        ctx.check(lib.mlgi_batch_create(ctx.h, arr, len(self.frames), C.byref(h)), "mlgi_batch_create")
        self.h = h

    def run(self) -> None:
        self.ctx.check(lib.mlgi_batch_run(self.ctx.h, self.h), "mlgi_batch_run")

    def update_light(self, frame: int, inv_bvp) -> None:
        """New inverse(bias*P*V) for frame `frame` (stream-ordered)."""
        m = np.ascontiguousarray(inv_bvp, np.float32)
        self._light_staging = m  # the copy is asynchronous: keep the source alive
        self.ctx.check(lib.mlgi_batch_update_light(self.ctx.h, self.h, frame, m.ctypes.data),
                       "mlgi_batch_update_light")

    def capture(self) -> "FrameGraph":
        g = C.c_void_p()
        self.ctx.check(lib.mlgi_batch_capture(self.ctx.h, self.h, C.byref(g)), "mlgi_batch_capture")
        return FrameGraph(self.ctx, g)

    def close(self) -> None:
        if getattr(self, "h", None):  # None also when __init__ raised
            lib.mlgi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class RsmPipeline:
    """Frames whose light moves every frame, with each frame's RSM ray-cast on
    the device. The RSM of frame i + 1 renders on a second context (its own
    stream) while frame i's GI runs on `ctx`, into the other of two RSM
    buffers; mlgi_context_wait orders the streams: frame i's GI after its RSM
    render, the render of frame i + 1 after frame i - 1's GI (the last reader
    of its buffer). The work and the results are those of rendering and then
    running each frame in turn. The stage's frame runs as the batched kernels
    replayed from a HIP graph, one graph per RSM buffer.

    `light_of(i)` is frame i's light Camera. After `frame()` returned i and
    `ctx` has synchronised, frame i's results are in the stage's buffers.
    """

    def __init__(self, ctx: Context, scene: Scene, stage: "GIStage", points: N.Points, gbuffer: N.GBuffer,
                 rsm_size: int, light_of):
        self.ctx, self.light_of = ctx, light_of
        self.render_ctx, self.dev, self.batches, self.graphs = None, None, [], []
        try:  # on failure, release what was made here before the error leaves
            self.render_ctx = Context(ctx.device)
            self.dev = scene.upload(self.render_ctx)
            r = rsm_size
            self.bufs = [(ctx.empty((r, r, 3), np.float32), ctx.empty((r, r, 3), np.float32),
                          ctx.empty((r, r), np.float32)) for _ in range(2)]
            eye = np.eye(4, dtype=np.float32).ravel()  # replaced per frame by update_light
            for b in self.bufs:
                self.batches.append(Batch(ctx, [(stage, rsm_desc(*b, eye), points, gbuffer)]))
            for b in self.batches:
                self.graphs.append(b.capture())
            self.next = 0
            self._inv = self._render(0)
        except BaseException:
            self.close()
            raise

    def _render(self, i: int):
        dif, nrm, dep = self.bufs[i % 2]
        return self.dev.render_rsm(self.light_of(i), dif, nrm, dep)

    def frame(self) -> int:
        """Enqueue the next frame; returns its index."""
        i = self.next
        self.ctx.wait_for(self.render_ctx)   # frame i's RSM render, the last work on the render stream
        self.render_ctx.wait_for(self.ctx)   # frame i - 1's GI, the last work on ctx
        inv_next = self._render(i + 1)
        self.batches[i % 2].update_light(0, self._inv)
        self.graphs[i % 2].launch()
        self._inv, self.next = inv_next, i + 1
        return i

    def sync(self) -> None:
        self.ctx.sync()
        if self.render_ctx is not None:
            self.render_ctx.sync()

    def close(self) -> None:
        if getattr(self, "render_ctx", None) is None:
            return
        try:
            self.sync()
        finally:
            for g in self.graphs:
                g.close()
            for b in self.batches:
                b.close()
            if self.dev is not None:
                self.dev.close()
            self.render_ctx.close()
            self.graphs, self.batches, self.dev, self.render_ctx = [], [], None, None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def frame_desc(stage: "GIStage", rsm: N.Rsm, points: N.Points, gbuffer: N.GBuffer) -> N.FrameDesc:
    d = N.FrameDesc()
    d.rsm, d.points, d.gbuffer, d.params = rsm, points, gbuffer, stage.params
    d.vpls, d.atlas, d.ism, d.image = (stage.vpl_processor.desc(), stage.ism.atlas_desc(), stage.ism.depth_desc(),
                                       stage.image_desc())
    return d


def autotune(ctx: Context, stage: "GIStage", rsm: N.Rsm, points: N.Points, gbuffer: N.GBuffer,
             min_vpls_per_pixel: int = 32) -> dict:
    """Pick interleave and gather tile for the stage's shape (mlgi_autotune);
    updates stage.params and returns the choice. Runs the frame."""
    d = frame_desc(stage, rsm, points, gbuffer)
    r = N.TuneResult()
    ctx.check(lib.mlgi_autotune(ctx.h, C.byref(d), min_vpls_per_pixel, C.byref(r)), "mlgi_autotune")
    stage.params.interleave, stage.params.use_interleaving = d.params.interleave, d.params.use_interleaving
    stage.params.tile_w, stage.params.tile_h = d.params.tile_w, d.params.tile_h
    return {"interleave": r.interleave, "tile": (r.tile_w, r.tile_h), "gather_ms": list(r.gather_ms),
            "cached": bool(r.cached)}


class Server:
    """GI request server (mlgi_server_*): resident scenes, HBM-budget
    admission, continuous batching on the device."""

    def __init__(self, ctx: Context, scenes, vpl_counts, resolutions, hbm_budget: int, max_inflight: int = 64,
                 max_batch: int = 32, max_queued_batches: int = 3, rsm_size: int = 128,
                 precision: int = PRECISION_MIXED, min_vpls_per_pixel: int = 32, streams: int = 0):
        """scenes: list of (Scene, n_points, point_seed); resolutions: list of (w, h);
        streams: device streams the queued batches are spread over (0: one per queued batch)."""
        self.ctx = ctx
        self._scenes = [sc for sc, _, _ in scenes]  # keep the host scenes alive
        arr = (N.ServerScene * len(scenes))()
        for i, (sc, npts, seed) in enumerate(scenes):
            arr[i].scene, arr[i].n_points, arr[i].point_seed = sc.h, npts, seed
        vc = (C.c_int32 * len(vpl_counts))(*vpl_counts)
        rs = (C.c_int32 * (2 * len(resolutions)))(*[v for wh in resolutions for v in wh])
        cfg = N.ServerConfig(hbm_budget, max_inflight, max_batch, max_queued_batches, rsm_size, precision,
                             min_vpls_per_pixel, vc, len(vpl_counts), rs, len(resolutions), streams)
        h = C.c_void_p()
        ctx.check(lib.mlgi_server_create(ctx.h, C.byref(cfg), arr, len(scenes), C.byref(h)), "mlgi_server_create")
        self.h = h

    def run(self, requests, outputs=None):
        """requests: list of (scene, w, h, n_vpl, light_seed, arrival_s), sorted by arrival.
        outputs: optional list of float32 (h, w, 3) arrays (or None) receiving the results.
        Returns (stats dict, latencies ms array)."""
        arr = (N.Request * len(requests))()
        for i, (sc, w, h, nv, seed, t) in enumerate(requests):
            out = outputs[i] if outputs is not None else None
            arr[i] = N.Request(sc, w, h, nv, seed, t, out.ctypes.data if out is not None else None)
        st = N.ServerStats()
        lat = np.empty(len(requests), np.float64)
        self.ctx.check(lib.mlgi_server_run(self.h, arr, len(requests), C.byref(st),
                                           lat.ctypes.data_as(C.POINTER(C.c_double))), "mlgi_server_run")
        stats = {k: getattr(st, k) for k, _ in N.ServerStats._fields_}
        return stats, lat

    def close(self) -> None:
        if getattr(self, "h", None):  # None also when __init__ raised
            lib.mlgi_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class FrameGraph:
    """An instantiated HIP graph of one GI frame (or one batch)."""

    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h

    def launch(self) -> None:
        self.ctx.check(lib.mlgi_graph_launch(self.ctx.h, self.h), "mlgi_graph_launch")

    def close(self) -> None:
        if getattr(self, "h", None):  # None also when __init__ raised
            lib.mlgi_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
