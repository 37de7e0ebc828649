"""Python bindings of libpsy_gpu.so over ctypes, with torch for device memory.

Ray and hit streams are torch tensors on the GPU: world rays as a
(20, n) float32 tensor, object rays as (15, n) (plane layout of
include/psy_gpu.h).  The library is the only compute path; importing this
module on a machine without the built library raises immediately.
"""
import ctypes as C
import os

import numpy as np
import torch

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpsy_gpu.so")
_lib = None


class PsyError(RuntimeError):
    pass


def lib():
    """Load libpsy_gpu.so (built in-tree by psychopath_cpp_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PsyError(f"{LIB_PATH} is missing: run `python -m psychopath_cpp_amd.build` first")
    L = C.CDLL(LIB_PATH)
    vp, i64, u32, i32 = C.c_void_p, C.c_int64, C.c_uint32, C.c_int32
    L.psy_last_error.restype = C.c_char_p
    L.psy_abi_version.restype = C.c_int
    L.psy_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.psy_ctx_destroy.argtypes = [vp]
    L.psy_ctx_set_trace_mode.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.psy_ctx_retrace_count.argtypes = [vp, C.POINTER(i64)]
    L.psy_ctx_trace_pass_counts.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    L.psy_ctx_set_render_chunk.argtypes = [vp, i64]
    L.psy_ctx_set_render_streams.argtypes = [vp, C.c_int]
    L.psy_ctx_set_render_shade_waves.argtypes = [vp, C.c_int]
    L.psy_ctx_set_render_mode.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.psy_scene_create.argtypes = [vp, C.POINTER(abi.SceneDesc), C.c_int, C.POINTER(vp)]
    L.psy_scene_destroy.argtypes = [vp]
    L.psy_scene_get_info.argtypes = [vp, C.POINTER(abi.SceneInfo)]
    L.psy_scene_instance_ids.argtypes = [vp, vp, i64]
    L.psy_subdivision_patches.argtypes = [vp, i32, vp, i32, vp, vp, i64, C.POINTER(i64)]
    L.psy_scene_set_dice_depth.argtypes = [vp, C.c_int]
    L.psy_trace_world.argtypes = [vp, vp, vp, i64, C.POINTER(abi.Hits), u32, vp]
    L.psy_trace_rays.argtypes = [vp, vp, vp, i64, C.POINTER(abi.Hits), u32, vp]
    L.psy_trace_rays_static.argtypes = [vp, vp, vp, i64, C.POINTER(abi.Hits), u32, vp]
    L.psy_intersections.argtypes = [vp, vp, vp, i64, C.POINTER(abi.Hits), vp, vp]
    L.psy_world_to_rays.argtypes = [vp, vp, i64, vp, vp]
    L.psy_halton.argtypes = [vp, u32, vp, i64, vp, vp]
    L.psy_ctx_set_trace_timing.argtypes = [vp, C.c_int]
    L.psy_ctx_set_trace_stack.argtypes = [vp, C.c_int]
    L.psy_ctx_set_second_pass.argtypes = [vp, u32, i32]
    L.psy_ctx_set_continuation.argtypes = [vp, C.c_int]
    L.psy_bucket_size.restype = C.c_int32
    L.psy_bucket_size.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.psy_render_adaptive.argtypes = [vp, vp, C.POINTER(abi.RenderParams), C.c_int32, C.c_float, C.c_int32, vp, vp,
                                      vp, C.POINTER(i64), vp]
    L.psy_film_adaptive.argtypes = [vp, vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_int32, vp, vp,
                                    vp, vp]
    L.psy_ctx_set_stream_steps.argtypes = [vp, C.c_int]
    L.psy_render_autotune.argtypes = [vp, vp, C.POINTER(abi.RenderParams), i64, C.POINTER(abi.RenderTuning)]
    L.psy_ctx_trace_timing.argtypes = [vp, vp, vp, vp]
    L.psy_sobol_set_matrices.argtypes = [vp, vp, u32]
    L.psy_sobol.argtypes = [vp, u32, vp, i64, C.c_uint64, vp, vp]
    L.psy_parse_scene_count.argtypes = [C.c_char_p, i64, C.POINTER(i32)]
    L.psy_parse_scene.argtypes = [C.c_char_p, i64, i32, C.POINTER(vp)]
    L.psy_parsed_scene_desc.restype = C.POINTER(abi.SceneDesc)
    L.psy_parsed_scene_desc.argtypes = [vp]
    L.psy_parsed_scene_settings.argtypes = [vp, C.POINTER(abi.RenderSettings)]
    L.psy_parsed_scene_warnings.restype = C.c_char_p
    L.psy_parsed_scene_warnings.argtypes = [vp]
    L.psy_parsed_scene_destroy.argtypes = [vp]
    L.psy_parse_dump_tree.argtypes = [C.c_char_p, i64, C.c_char_p, i64, C.POINTER(i64)]
    L.psy_film_to_srgb8.argtypes = [vp, vp, i32, i32, u32, vp]
    L.psy_write_png.argtypes = [C.c_char_p, vp, i32, i32]
    L.psy_write_pfm.argtypes = [C.c_char_p, vp, i32, i32]
    L.psy_spectrum_set_grid.argtypes = [vp, vp, C.c_int32, C.c_int32, vp, C.c_int32, C.c_int32, C.c_float,
                                        C.c_float, vp]
    L.psy_color_to_spectrum.argtypes = [vp, vp, vp, i64, vp, vp]
    L.psy_pixel_samples.argtypes = [vp, i32, i32, i32, i32, i32, u32, u32, i32, vp, vp]
    L.psy_camera_rays.argtypes = [vp, vp, i32, i32, i32, u32, vp, vp, vp]
    L.psy_bounce_rays.argtypes = [vp, vp, vp, C.POINTER(abi.Hits), i64, i32, i32, u32, u32, u32, vp, vp, vp]
    L.psy_render.argtypes = [vp, vp, C.POINTER(abi.RenderParams), vp, vp, C.POINTER(i64), vp]
    L.psy_scene_param_counts.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    L.psy_scene_get_params.argtypes = [vp, vp, vp]
    L.psy_scene_set_params.argtypes = [vp, vp, i64, vp, i64]
    L.psy_adjoint.argtypes = [vp, vp, C.POINTER(abi.RenderParams), vp, vp, vp, C.POINTER(i64), vp]
    if L.psy_abi_version() != 3:
        raise PsyError("libpsy_gpu.so ABI version mismatch")
    _lib = L
    return L


def subdivision_patches(verts, face_vert_counts, face_vert_indices):
    """SubdivisionSurface::finalize (object/subdivision_surface.cpp:85-233):
    the Catmull-Clark limit surface of a control mesh as (n, 16, 3) bicubic
    Bezier control points (rows along v).  Host only; no GPU is needed."""
    L = lib()
    v = np.ascontiguousarray(np.asarray(verts, np.float32).reshape(-1, 3))
    cnt = np.ascontiguousarray(np.asarray(face_vert_counts, np.int32).reshape(-1))
    idx = np.ascontiguousarray(np.asarray(face_vert_indices, np.int32).reshape(-1))
    if int(cnt.sum()) != idx.size:
        raise PsyError("face_vert_indices must hold sum(face_vert_counts) entries")
    n = C.c_int64(0)
    args = (v.ctypes.data, v.shape[0], cnt.ctypes.data, cnt.size, idx.ctypes.data)
    check(L.psy_subdivision_patches(*args, None, 0, C.byref(n)), "psy_subdivision_patches")
    out = np.zeros((max(1, n.value), 16, 3), np.float32)
    check(L.psy_subdivision_patches(*args, out.ctypes.data, n.value, C.byref(n)), "psy_subdivision_patches")
    return out[:n.value]


def check(rc, what):
    if rc != abi.PSY_OK:
        msg = lib().psy_last_error().decode(errors="replace")
        raise PsyError(f"{what} failed (status {rc}): {msg}")


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


class Context:
    """psy_ctx_create / psy_ctx_destroy."""

    def __init__(self, device=0):
        self.device = device
        h = C.c_void_p()
        check(lib().psy_ctx_create(device, C.byref(h)), "psy_ctx_create")
        self.handle = h

    def set_trace_mode(self, persistent=True, refill_min=16, blocks_per_cu=0, coherent=False):
        """psy_ctx_set_trace_mode: persistent main pass with lane refill (coherent: one kind of step per wave
        step), or one thread per ray."""
        mode = abi.PSY_TRACE_MODE_PLAIN_COHERENT if coherent else abi.PSY_TRACE_MODE_PLAIN
        if persistent:
            mode = abi.PSY_TRACE_MODE_COHERENT if coherent else abi.PSY_TRACE_MODE_PERSISTENT
        check(lib().psy_ctx_set_trace_mode(self.handle, mode, refill_min, blocks_per_cu), "psy_ctx_set_trace_mode")

    def set_trace_stack(self, entries):
        """psy_ctx_set_trace_stack: main-pass LDS stack entries (16 or 32)."""
        check(lib().psy_ctx_set_trace_stack(self.handle, int(entries)), "psy_ctx_set_trace_stack")

    def set_second_pass(self, wide_min=4096, budget=512):
        """psy_ctx_set_second_pass: queue length from which the second pass's wide build walks the queue
        (0: always wide, 0xffffffff: always narrow) and its node budget per ray before the third pass."""
        check(lib().psy_ctx_set_second_pass(self.handle, int(wide_min) & 0xFFFFFFFF, int(budget)),
              "psy_ctx_set_second_pass")

    def set_continuation(self, below=24):
        """psy_ctx_set_continuation: a main-pass wave hands its rays to a follow-on launch once fewer than
        `below` of them are still tracing (0: never)."""
        check(lib().psy_ctx_set_continuation(self.handle, int(below)), "psy_ctx_set_continuation")

    def set_trace_timing(self, enable=True):
        """psy_ctx_set_trace_timing: HIP events around each trace call's passes."""
        check(lib().psy_ctx_set_trace_timing(self.handle, 1 if enable else 0), "psy_ctx_set_trace_timing")

    def trace_timing(self):
        """psy_ctx_trace_timing: (main-pass ms, second-pass ms, calls) summed
        over the trace calls since the previous query."""
        a, b, n = C.c_double(), C.c_double(), C.c_int64()
        check(lib().psy_ctx_trace_timing(self.handle, C.byref(a), C.byref(b), C.byref(n)), "psy_ctx_trace_timing")
        return a.value, b.value, n.value

    def retrace_count(self):
        """psy_ctx_retrace_count: rays of the latest trace that took the second pass."""
        return self._retrace_count()

    def pass_counts(self):
        """psy_ctx_trace_pass_counts: (second-pass rays, third-pass rays) of the latest trace."""
        a, b = C.c_int64(), C.c_int64()
        check(lib().psy_ctx_trace_pass_counts(self.handle, C.byref(a), C.byref(b)), "psy_ctx_trace_pass_counts")
        return a.value, b.value

    def _retrace_count(self):
        v = C.c_int64()
        check(lib().psy_ctx_retrace_count(self.handle, C.byref(v)), "psy_ctx_retrace_count")
        return v.value

    def close(self):
        if self.handle:
            lib().psy_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuScene:
    """psy_scene_create: flatten + BVH4 + dice cache resident in HBM."""

    def __init__(self, ctx, scene, dice_depth=-1):
        self.ctx = ctx
        self._desc = scene.to_desc() if hasattr(scene, "to_desc") else scene
        h = C.c_void_p()
        check(lib().psy_scene_create(ctx.handle, self._desc.ptr(), int(dice_depth), C.byref(h)),
              "psy_scene_create")
        self.handle = h
        self.refresh_info()

    def refresh_info(self):
        info = abi.SceneInfo()
        check(lib().psy_scene_get_info(self.handle, C.byref(info)), "psy_scene_get_info")
        self.info = info

    def set_dice_depth(self, depth):
        """psy_scene_set_dice_depth: rebuild the dice cache at `depth` (< 0:
        the memory budget's depth)."""
        check(lib().psy_scene_set_dice_depth(self.handle, int(depth)), "psy_scene_set_dice_depth")
        self.refresh_info()

    def param_counts(self):
        """psy_scene_param_counts: (shader albedos, light emission scales)."""
        a, e = C.c_int64(), C.c_int64()
        check(lib().psy_scene_param_counts(self.handle, C.byref(a), C.byref(e)), "psy_scene_param_counts")
        return a.value, e.value

    def get_params(self):
        """psy_scene_get_params: (albedo, emission_scale) float32 arrays."""
        na, ne = self.param_counts()
        a = np.zeros(max(1, na), np.float32)
        e = np.zeros(max(1, ne), np.float32)
        check(lib().psy_scene_get_params(self.handle, a.ctypes.data_as(C.c_void_p), e.ctypes.data_as(C.c_void_p)),
              "psy_scene_get_params")
        return a[:na], e[:ne]

    def set_params(self, albedo=None, emission_scale=None):
        """psy_scene_set_params from host arrays (None keeps a set)."""
        a = None if albedo is None else np.ascontiguousarray(
            albedo.detach().cpu().numpy() if torch.is_tensor(albedo) else albedo, np.float32)
        e = None if emission_scale is None else np.ascontiguousarray(
            emission_scale.detach().cpu().numpy() if torch.is_tensor(emission_scale) else emission_scale, np.float32)
        check(lib().psy_scene_set_params(self.handle, None if a is None else a.ctypes.data_as(C.c_void_p),
                                         0 if a is None else a.size,
                                         None if e is None else e.ctypes.data_as(C.c_void_p),
                                         0 if e is None else e.size), "psy_scene_set_params")

    def instance_ids(self):
        n = int(self.info.n_leaf_instances)
        out = np.zeros(max(1, n), np.uint64)
        check(lib().psy_scene_instance_ids(self.handle, out.ctypes.data_as(C.c_void_p), out.size),
              "psy_scene_instance_ids")
        return out[:n]

    def close(self):
        if self.handle:
            lib().psy_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HitBuffers:
    def __init__(self, n, device, aux=True):
        self.t = torch.empty(n, dtype=torch.float32, device=device)
        self.inst = torch.empty(n, dtype=torch.int32, device=device)
        self.u = torch.empty(n, dtype=torch.float32, device=device)
        self.v = torch.empty(n, dtype=torch.float32, device=device)
        self.aux = torch.empty(n, dtype=torch.float32, device=device) if aux else None

    def struct(self):
        return abi.Hits(self.t.data_ptr(), self.inst.data_ptr(), self.u.data_ptr(), self.v.data_ptr(),
                        self.aux.data_ptr() if self.aux is not None else None)

    def inst_u32(self):
        return self.inst.view(torch.int32).cpu().numpy().view(np.uint32)


def _check_planes(t, planes, name):
    if not (t.is_cuda and t.dtype == torch.float32 and t.dim() == 2 and t.shape[0] == planes and t.is_contiguous()):
        raise PsyError(f"{name} must be a contiguous CUDA float32 tensor of shape ({planes}, n)")


def trace_world(ctx, scene, wrays, occlusion_only=False, hits=None):
    """Tracer::trace over world rays (20, n)."""
    _check_planes(wrays, abi.PSY_WRAY_PLANES, "world rays")
    n = wrays.shape[1]
    hits = hits or HitBuffers(n, wrays.device)
    hs = hits.struct()
    flags = abi.PSY_TRACE_OCCLUSION_ONLY if occlusion_only else 0
    check(lib().psy_trace_world(ctx.handle, scene.handle, _ptr(wrays), n, C.byref(hs), flags, _stream()),
          "psy_trace_world")
    return hits


def trace_rays(ctx, scene, rays, occlusion_only=False, hits=None):
    """Tracer::trace over rays with derived footprints (15, n)."""
    _check_planes(rays, abi.PSY_RAY_PLANES, "rays")
    n = rays.shape[1]
    hits = hits or HitBuffers(n, rays.device)
    hs = hits.struct()
    flags = abi.PSY_TRACE_OCCLUSION_ONLY if occlusion_only else 0
    check(lib().psy_trace_rays(ctx.handle, scene.handle, _ptr(rays), n, C.byref(hs), flags, _stream()),
          "psy_trace_rays")
    return hits


def trace_rays_static(ctx, scene, rays, occlusion_only=False, hits=None):
    """psy_trace_rays_static: trace_rays through the library's fixed scratch (no device allocation in the
    call); at most abi.PSY_STATIC_MAX_RAYS rays."""
    _check_planes(rays, abi.PSY_RAY_PLANES, "rays")
    n = rays.shape[1]
    hits = hits or HitBuffers(n, rays.device)
    hs = hits.struct()
    flags = abi.PSY_TRACE_OCCLUSION_ONLY if occlusion_only else 0
    check(lib().psy_trace_rays_static(ctx.handle, scene.handle, _ptr(rays), n, C.byref(hs), flags, _stream()),
          "psy_trace_rays_static")
    return hits


def intersections(ctx, scene, wrays, hits):
    """psy_intersections: the Intersection fields of the hits of trace_world(wrays) as a
    (PSY_ISECT_PLANES, n) float32 tensor (see include/psy_gpu.h for the planes)."""
    _check_planes(wrays, abi.PSY_WRAY_PLANES, "world rays")
    n = wrays.shape[1]
    out = torch.empty((abi.PSY_ISECT_PLANES, n), dtype=torch.float32, device=wrays.device)
    hs = hits.struct()
    check(lib().psy_intersections(ctx.handle, scene.handle, _ptr(wrays), n, C.byref(hs), _ptr(out), _stream()),
          "psy_intersections")
    return out


def world_to_rays(ctx, wrays):
    _check_planes(wrays, abi.PSY_WRAY_PLANES, "world rays")
    n = wrays.shape[1]
    out = torch.empty((abi.PSY_RAY_PLANES, n), dtype=torch.float32, device=wrays.device)
    check(lib().psy_world_to_rays(ctx.handle, _ptr(wrays), n, _ptr(out), _stream()), "psy_world_to_rays")
    return out


def halton(ctx, dim, index):
    if not (index.is_cuda and index.dtype == torch.int32 and index.is_contiguous()):
        raise PsyError("index must be a contiguous CUDA int32 tensor (uint32 bit patterns)")
    out = torch.empty(index.shape, dtype=torch.float32, device=index.device)
    check(lib().psy_halton(ctx.handle, dim, _ptr(index), index.numel(), _ptr(out), _stream()), "psy_halton")
    return out


def sobol_set_matrices(ctx, matrices):
    """Generator matrices (numpy uint64, 52 columns per dimension), e.g. the
    reference's Sobol::Matrices::matrices."""
    m = np.ascontiguousarray(matrices, np.uint64).reshape(-1)
    if m.size == 0 or m.size % 52:
        raise PsyError("matrices must hold 52 columns per dimension")
    check(lib().psy_sobol_set_matrices(ctx.handle, m.ctypes.data_as(C.c_void_p), m.size // 52),
          "psy_sobol_set_matrices")


def sobol(ctx, dim, index, scramble=0):
    """Sobol::sample(dim, index[i], scramble) for a CUDA int64 tensor of
    uint64 bit patterns; float64 results."""
    if not (index.is_cuda and index.dtype == torch.int64 and index.is_contiguous()):
        raise PsyError("index must be a contiguous CUDA int64 tensor (uint64 bit patterns)")
    out = torch.empty(index.shape, dtype=torch.float64, device=index.device)
    check(lib().psy_sobol(ctx.handle, dim, _ptr(index), index.numel(), int(scramble) & (2**64 - 1), _ptr(out),
                          _stream()), "psy_sobol")
    return out


def spectrum_set_grid(ctx, cells, w, h, points, n_samples, sample_min, sample_max, mat_xy_to_uv):
    """The spectrum grid of colour upsampling (color/spectrum_grid.h) as
    arrays: cells (w*h, 8) int32 = inside, num_points, idx[6]; points
    (n, 4 + n_samples) float32 = xystar, uv, spectrum; the xy -> uv matrix
    (6 floats).  cells None clears it (grey colours only)."""
    if cells is None:
        check(lib().psy_spectrum_set_grid(ctx.handle, None, 0, 0, None, 0, 0, 0.0, 0.0, None),
              "psy_spectrum_set_grid")
        return
    c = np.ascontiguousarray(cells, np.int32).reshape(-1)
    p = np.ascontiguousarray(points, np.float32).reshape(-1)
    m = np.ascontiguousarray(mat_xy_to_uv, np.float32).reshape(-1)
    if c.size != 8 * w * h or p.size % (4 + n_samples) or m.size != 6:
        raise PsyError("spectrum grid arrays do not match their sizes")
    check(lib().psy_spectrum_set_grid(ctx.handle, c.ctypes.data_as(C.c_void_p), w, h, p.ctypes.data_as(C.c_void_p),
                                      p.size // (4 + n_samples), n_samples, sample_min, sample_max,
                                      m.ctypes.data_as(C.c_void_p)), "psy_spectrum_set_grid")


def color_to_spectrum(ctx, rgb, hero):
    """Color_to_SpectralSample for CUDA float32 tensors rgb (n, 3) and hero
    wavelengths (n,): (n, 4) spectral samples."""
    for t in (rgb, hero):
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
            raise PsyError("rgb and hero must be contiguous CUDA float32 tensors")
    n = hero.numel()
    if rgb.numel() != 3 * n:
        raise PsyError("rgb must hold 3 floats per wavelength")
    out = torch.empty((n, 4), dtype=torch.float32, device=hero.device)
    check(lib().psy_color_to_spectrum(ctx.handle, _ptr(rgb), _ptr(hero), n, _ptr(out), _stream()),
          "psy_color_to_spectrum")
    return out


def pixel_samples(ctx, x0, y0, w, h, spp, seed, dim0, n_dims, device="cuda"):
    out = torch.empty((w * h * spp, n_dims), dtype=torch.float32, device=device)
    check(lib().psy_pixel_samples(ctx.handle, x0, y0, w, h, spp, seed, dim0, n_dims, _ptr(out), _stream()),
          "psy_pixel_samples")
    return out


def camera_rays(ctx, scene, width, height, spp, seed, device="cuda"):
    n = width * height * spp
    w = torch.empty((abi.PSY_WRAY_PLANES, n), dtype=torch.float32, device=device)
    wl = torch.empty(n, dtype=torch.float32, device=device)
    check(lib().psy_camera_rays(ctx.handle, scene.handle, width, height, spp, seed, _ptr(w), _ptr(wl), _stream()),
          "psy_camera_rays")
    return w, wl


def bounce_rays(ctx, scene, wrays, hits, width, spp, seed, dim_u=9, dim_v=10):
    _check_planes(wrays, abi.PSY_WRAY_PLANES, "world rays")
    n = wrays.shape[1]
    out = torch.empty_like(wrays)
    valid = torch.empty(n, dtype=torch.int32, device=wrays.device)
    hs = hits.struct()
    check(lib().psy_bounce_rays(ctx.handle, scene.handle, _ptr(wrays), C.byref(hs), n, width, spp, seed, dim_u,
                                dim_v, _ptr(out), _ptr(valid), _stream()), "psy_bounce_rays")
    return out, valid


def render(ctx, scene, width, height, spp, seed=1, path_length=4, film=True, samples=False, sobol=False):
    """psy_render: the path-traced image (PathTraceIntegrator::integrate).

    Returns (film, samples, rays_traced): film is a (height, width, 3) CIE XYZ
    tensor (per-pixel mean), samples a (height, width, spp, 3) tensor of every
    sample's film contribution; either is None when not requested.  sobol:
    PSY_RENDER_SOBOL (the context's Sobol matrices instead of Halton)."""
    dev = torch.device("cuda", ctx.device)
    f = torch.empty((height, width, 3), dtype=torch.float32, device=dev) if film else None
    s = torch.empty((height, width, spp, 3), dtype=torch.float32, device=dev) if samples else None
    prm = abi.RenderParams(width, height, spp, seed, path_length, abi.PSY_RENDER_SOBOL if sobol else 0)
    rays = C.c_int64()
    check(lib().psy_render(ctx.handle, scene.handle, C.byref(prm), _ptr(f), _ptr(s), C.byref(rays), _stream()),
          "psy_render")
    return f, s, rays.value


def bucket_size(width, height, spp, thread_count):
    """psy_bucket_size: PathTraceIntegrator::integrate's bucket size."""
    return int(lib().psy_bucket_size(int(width), int(height), int(spp), int(thread_count)))


def render_adaptive(ctx, scene, width, height, spp, spp_max, variance_max, thread_count=1, seed=1, path_length=4):
    """psy_render_adaptive: (film (H, W, 3), variance (H, W, 3), passes per bucket, rays_traced) of the
    reference's adaptive sampling loop with spp_max and variance_max."""
    dev = torch.device("cuda", ctx.device)
    f = torch.empty((height, width, 3), dtype=torch.float32, device=dev)
    v = torch.empty((height, width, 3), dtype=torch.float32, device=dev)
    bs = bucket_size(width, height, spp, thread_count)
    nb = ((width + bs - 1) // bs) * ((height + bs - 1) // bs)
    p = torch.empty(nb, dtype=torch.int32, device=dev)
    prm = abi.RenderParams(width, height, spp, seed, path_length, 0)
    rays = C.c_int64()
    check(lib().psy_render_adaptive(ctx.handle, scene.handle, C.byref(prm), int(spp_max), C.c_float(variance_max),
                                    int(thread_count), _ptr(f), _ptr(v), _ptr(p), C.byref(rays), _stream()),
          "psy_render_adaptive")
    return f, v, p, rays.value


def film_adaptive(ctx, samples, spp_max, variance_max, bucket_size_):
    """psy_film_adaptive over a (H, W, spp, 3) CUDA float32 tensor of samples: (film, variance, passes)."""
    samples = samples.contiguous()
    H, W, spp = int(samples.shape[0]), int(samples.shape[1]), int(samples.shape[2])
    f = torch.empty((H, W, 3), dtype=torch.float32, device=samples.device)
    v = torch.empty((H, W, 3), dtype=torch.float32, device=samples.device)
    nb = ((W + bucket_size_ - 1) // bucket_size_) * ((H + bucket_size_ - 1) // bucket_size_)
    p = torch.empty(nb, dtype=torch.int32, device=samples.device)
    check(lib().psy_film_adaptive(ctx.handle, _ptr(samples), W, H, spp, int(spp_max), C.c_float(variance_max),
                                  int(bucket_size_), _ptr(f), _ptr(v), _ptr(p), _stream()), "psy_film_adaptive")
    return f, v, p


def adjoint(ctx, scene, d_film, spp, seed=1, path_length=4):
    """psy_adjoint: (d_albedo, d_emission, rays_traced) for the film gradient
    d_film, a (height, width, 3) CUDA float32 tensor."""
    if not (d_film.is_cuda and d_film.dtype == torch.float32 and d_film.dim() == 3 and d_film.shape[2] == 3):
        raise PsyError("d_film must be a (height, width, 3) CUDA float32 tensor")
    d_film = d_film.contiguous()
    height, width = int(d_film.shape[0]), int(d_film.shape[1])
    na, ne = scene.param_counts()
    ga = torch.zeros(na, dtype=torch.float32, device=d_film.device)
    ge = torch.zeros(ne, dtype=torch.float32, device=d_film.device)
    prm = abi.RenderParams(width, height, spp, seed, path_length, 0)
    rays = C.c_int64()
    check(lib().psy_adjoint(ctx.handle, scene.handle, C.byref(prm), _ptr(d_film), _ptr(ga) if na else None,
                            _ptr(ge) if ne else None, C.byref(rays), _stream()), "psy_adjoint")
    return ga, ge, rays.value


def set_render_chunk(ctx, paths):
    """psy_ctx_set_render_chunk: paths per wavefront chunk of psy_render."""
    check(lib().psy_ctx_set_render_chunk(ctx.handle, int(paths)), "psy_ctx_set_render_chunk")


def set_render_streams(ctx, streams):
    """psy_ctx_set_render_streams: wavefront chunks in flight (1..4)."""
    check(lib().psy_ctx_set_render_streams(ctx.handle, int(streams)), "psy_ctx_set_render_streams")


def set_render_shade_waves(ctx, waves):
    """psy_ctx_set_render_shade_waves: 0 (compiler's registers) or 4 (the
    hit step's shading held to 4 waves per SIMD)."""
    check(lib().psy_ctx_set_render_shade_waves(ctx.handle, int(waves)), "psy_ctx_set_render_shade_waves")


def set_render_mode(ctx, fused=False, refill_min=16, blocks_per_cu=0, stream=False):
    """psy_ctx_set_render_mode: wavefront (default), fused persistent paths,
    or the stream form (stream=True)."""
    mode = abi.PSY_RENDER_STREAM if stream else (abi.PSY_RENDER_FUSED if fused else abi.PSY_RENDER_WAVEFRONT)
    check(lib().psy_ctx_set_render_mode(ctx.handle, mode, refill_min, blocks_per_cu), "psy_ctx_set_render_mode")


def set_stream_steps(ctx, steps):
    """psy_ctx_set_stream_steps: steps per captured graph (even, 2..64)."""
    check(lib().psy_ctx_set_stream_steps(ctx.handle, int(steps)), "psy_ctx_set_stream_steps")


def autotune_render(ctx, scene, width, height, spp, seed=1, path_length=4, max_samples=0):
    """psy_render_autotune: leaves the fastest render setting on ctx (and
    the fastest dicing depth on the scene) and returns it as a dict."""
    prm = abi.RenderParams(width, height, spp, seed, path_length, 0)
    out = abi.RenderTuning()
    check(lib().psy_render_autotune(ctx.handle, scene.handle, C.byref(prm), int(max_samples), C.byref(out)),
          "psy_render_autotune")
    scene.refresh_info()
    return {f: getattr(out, f) for f, _ in abi.RenderTuning._fields_}
