"""TEST INFRASTRUCTURE: builds and loads libpsy_hostbuild.so, the device
code of libpsy_gpu.so compiled for the host CPU (psy_host_build.cpp), so
CPU tests can run the kernels' per-ray and per-path logic against the
oracle.  Never imported by the product."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "psy_host_build.cpp")
LIB = os.path.join(HERE, "libpsy_hostbuild.so")
_lib = None


def build():
    from psychopath_cpp_amd import build as pb
    prod = os.path.join(ROOT, "psychopath_cpp_amd", "libpsy_gpu.so")
    if not os.path.exists(prod):
        pb.build(verbose=False)
    flags = [f for f in pb.HIPCC_FLAGS if f not in ("-shared", "--offload-arch=gfx950")]
    cmd = ([pb.hipcc(), "-x", "hip", "--offload-host-only", "-shared"] + flags +
           [SRC, "-o", LIB + ".tmp", "-L" + os.path.dirname(prod), "-lpsy_gpu",
            "-Wl,-rpath," + os.path.dirname(prod), "-lpthread"])
    r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build failed:\n" + r.stderr[-4000:])
    os.replace(LIB + ".tmp", LIB)
    return LIB


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(
            os.path.getmtime(os.path.join(ROOT, "psychopath_cpp_amd", "libpsy_gpu.so")), os.path.getmtime(SRC)):
        build()
    from psychopath_cpp_amd import abi
    L = C.CDLL(LIB)
    vp, i64 = C.c_void_p, C.c_int64
    L.hb_trace.argtypes = [C.POINTER(abi.SceneDesc), C.c_int, C.c_int, vp, i64, C.c_uint32, vp, vp, vp, vp, vp,
                           C.POINTER(i64), C.c_int]
    L.hb_render.argtypes = [C.POINTER(abi.SceneDesc), C.c_int, C.POINTER(abi.RenderParams), vp, C.POINTER(i64),
                            C.c_int]
    L.hb_render_params.argtypes = [C.POINTER(abi.SceneDesc), C.c_int, C.POINTER(abi.RenderParams), vp, C.c_int, vp,
                                   C.c_int, vp, vp, vp, C.POINTER(i64), C.c_int]
    L.hb_set_spectrum_grid.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int, C.c_int, C.c_float, C.c_float, vp]
    L.hb_scene_stack_need.argtypes = [C.POINTER(abi.SceneDesc), C.c_int, C.POINTER(C.c_int)]
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def trace(scene_py, rays, world=True, dice_depth=10, occlusion_only=False, threads=8):
    """Hits of the (20 or 15, n) float32 ray planes, as the kernels compute them."""
    desc = scene_py.to_desc()
    rays = np.ascontiguousarray(rays, np.float32)
    n = rays.shape[1]
    t = np.zeros(n, np.float32)
    inst = np.zeros(n, np.uint32)
    u = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    aux = np.zeros(n, np.float32)
    rt = C.c_int64()
    rc = lib().hb_trace(desc.ptr(), dice_depth, 1 if world else 0, _p(rays), n, 1 if occlusion_only else 0,
                        _p(t), _p(inst), _p(u), _p(v), _p(aux), C.byref(rt), threads)
    if rc != 0:
        raise RuntimeError(f"hb_trace failed with status {rc}")
    return {"t": t, "inst": inst, "u": u, "v": v, "aux": aux, "retraced": rt.value}


def scene_stack_need(scene_py, dice_depth=0):
    """(status, traversal stack need) of psy_scene_create's host preparation."""
    desc = scene_py.to_desc()
    need = C.c_int()
    rc = lib().hb_scene_stack_need(desc.ptr(), dice_depth, C.byref(need))
    return rc, need.value


def render(scene_py, width, height, spp, seed=1, path_length=4, dice_depth=10, threads=8):
    """Per-sample film values (height, width, spp, 3) and rays traced."""
    from psychopath_cpp_amd import abi
    desc = scene_py.to_desc()
    out = np.zeros((height, width, spp, 3), np.float32)
    prm = abi.RenderParams(width, height, spp, seed, path_length, 0)
    rays = C.c_int64()
    rc = lib().hb_render(desc.ptr(), dice_depth, C.byref(prm), _p(out), C.byref(rays), threads)
    if rc != 0:
        raise RuntimeError(f"hb_render failed with status {rc}")
    return out, rays.value


def render_params(scene_py, width, height, spp, albedo=None, escale=None, d_film=None, seed=1, path_length=4,
                  dice_depth=10, threads=8):
    """psy_render (d_film None: per-sample film values) or psy_adjoint
    (d_film (height, width, 3): the gradient vector, albedos then emission
    scales) with the given parameters, as the kernels compute them."""
    from psychopath_cpp_amd import abi
    desc = scene_py.to_desc()
    prm = abi.RenderParams(width, height, spp, seed, path_length, 0)
    a = None if albedo is None else np.ascontiguousarray(albedo, np.float32)
    e = None if escale is None else np.ascontiguousarray(escale, np.float32)
    n_params = param_counts(desc)
    out = np.zeros((height, width, spp, 3), np.float32) if d_film is None else None
    df = None if d_film is None else np.ascontiguousarray(d_film, np.float32)
    grad = np.zeros(max(1, sum(n_params)), np.float32)
    rays = C.c_int64()
    rc = lib().hb_render_params(desc.ptr(), dice_depth, C.byref(prm), None if a is None else _p(a),
                                0 if a is None else a.size, None if e is None else _p(e), 0 if e is None else e.size,
                                None if out is None else _p(out), None if df is None else _p(df), _p(grad),
                                C.byref(rays), threads)
    if rc != 0:
        raise RuntimeError(f"hb_render_params failed with status {rc}")
    return (out if d_film is None else grad[:sum(n_params)]), rays.value


def set_spectrum_grid(grid):
    """Hand the host-built integrator a spectrum grid (an object with the
    arrays of oracle.ora.SpectrumGrid), or None to clear it."""
    L = lib()
    if grid is None:
        L.hb_set_spectrum_grid(None, 0, 0, None, 0, 0, 0.0, 0.0, None)
    else:
        L.hb_set_spectrum_grid(_p(grid.cells), grid.w, grid.h, _p(grid.points), grid.n_points, grid.n_samples,
                               grid.smin, grid.smax, _p(grid.mat))


def param_counts(desc):
    """(shaders, light objects) of a scene description, in the order of
    psy_scene_param_counts."""
    from psychopath_cpp_amd import abi
    d = desc.desc if hasattr(desc, "desc") else desc.ptr().contents
    n_light = 0
    for a in range(d.n_assemblies):
        ad = d.assemblies[a]
        for i in range(ad.n_objects):
            if ad.objects[i].kind in (abi.PSY_OBJ_SPHERE_LIGHT, abi.PSY_OBJ_RECTANGLE_LIGHT):
                n_light += 1
    return d.n_shaders, n_light
