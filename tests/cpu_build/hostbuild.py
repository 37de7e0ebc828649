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
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
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
