"""Scene construction mirroring the reference's Assembly / Scene API.

    Assembly.add_object(name, obj)              scene/assembly.hpp:136-141
    Assembly.add_assembly(name, asm)            scene/assembly.hpp:151-155
    Assembly.create_object_instance(...)        scene/assembly.hpp:160-167
    Assembly.create_assembly_instance(...)      scene/assembly.hpp:172-179
    Scene.add_surface_shader(name, shader)      scene/assembly.hpp:105-109
    Scene(camera, root, background)             scene/scene.hpp:33-53

Names resolve to indices; `Scene.to_desc()` produces the psy_scene_desc
(include/psy_gpu.h) consumed by the GPU library, keeping every buffer alive
for as long as the returned handle lives.  Matrices are 4x4 row-major in
the reference's row-vector convention (p' = p * M).
"""
import ctypes as C

import numpy as np

from . import abi


class Object:
    """A geometric object or light with one or more motion samples."""

    def __init__(self, kind, samples):
        nf = abi.FLOATS_PER_KIND[kind]
        arr = np.asarray(samples, dtype=np.float32).reshape(-1, nf)
        if arr.shape[0] < 1:
            raise ValueError("object needs at least one time sample")
        self.kind = kind
        self.data = np.ascontiguousarray(arr)

    @property
    def time_samples(self):
        return self.data.shape[0]


def Sphere(center, radius):
    """object/sphere.hpp; center (ts,3) or (3,), radius (ts,) or scalar."""
    c = np.atleast_2d(np.asarray(center, np.float32))
    r = np.atleast_1d(np.asarray(radius, np.float32))
    return Object(abi.PSY_OBJ_SPHERE, np.concatenate([c, r[:, None]], axis=1))


def Bilinear(*verts_per_sample):
    """object/bilinear.hpp: v1 v2 v3 v4 per sample (u along v1->v2, v along v1->v3)."""
    return Object(abi.PSY_OBJ_BILINEAR, [np.asarray(v, np.float32).reshape(12) for v in verts_per_sample])


def Bicubic(*points_per_sample):
    """object/bicubic.hpp: 16 Bezier control points per sample, row-major."""
    return Object(abi.PSY_OBJ_BICUBIC, [np.asarray(p, np.float32).reshape(48) for p in points_per_sample])


class SubdivisionSurface:
    """object/subdivision_surface.hpp:15-71: a Catmull-Clark control mesh
    (vertices (n,3), per-face vertex counts, concatenated face vertex
    indices) with one time sample."""

    kind = abi.PSY_OBJ_SUBDIVISION

    def __init__(self, verts, face_vert_counts, face_vert_indices):
        self.data = np.ascontiguousarray(np.asarray(verts, np.float32).reshape(-1, 3))
        self.face_vert_counts = np.ascontiguousarray(np.asarray(face_vert_counts, np.int32).reshape(-1))
        self.face_vert_indices = np.ascontiguousarray(np.asarray(face_vert_indices, np.int32).reshape(-1))
        if int(self.face_vert_counts.sum()) != self.face_vert_indices.size:
            raise ValueError("face_vert_indices must hold sum(face_vert_counts) entries")

    @property
    def time_samples(self):
        return 1


def SphereLight(center, radius, color):
    """lights/sphere_light.hpp: positions, radii, colours per sample."""
    c = np.atleast_2d(np.asarray(center, np.float32))
    r = np.atleast_1d(np.asarray(radius, np.float32))
    col = np.atleast_2d(np.asarray(color, np.float32))
    return Object(abi.PSY_OBJ_SPHERE_LIGHT, np.concatenate([c, r[:, None], col], axis=1))


def RectangleLight(width, height, color):
    """lights/rectangle_light.hpp: dimensions and colour per sample."""
    w = np.atleast_1d(np.asarray(width, np.float32))
    h = np.atleast_1d(np.asarray(height, np.float32))
    col = np.atleast_2d(np.asarray(color, np.float32))
    return Object(abi.PSY_OBJ_RECTANGLE_LIGHT, np.concatenate([w[:, None], h[:, None], col], axis=1))


class Assembly:
    def __init__(self):
        self.objects = []
        self.object_map = {}
        self.assemblies = []
        self.assembly_map = {}
        self.instances = []   # (type, data_index, transform_index, transform_count, shader_name)
        self.xforms = []

    def add_object(self, name, obj):
        self.objects.append(obj)
        self.object_map[name] = len(self.objects) - 1
        return True

    def add_assembly(self, name, asm):
        self.assemblies.append(asm)
        self.assembly_map[name] = len(self.assemblies) - 1
        return True

    def _instance(self, kind, index, transforms, shader):
        ts = [np.asarray(t, np.float32).reshape(4, 4) for t in (transforms or [])]
        self.instances.append((kind, index, len(self.xforms), len(ts), shader))
        self.xforms.extend(ts)
        return True

    def create_object_instance(self, name, transforms=None, shader=None):
        return self._instance(abi.PSY_INSTANCE_OBJECT, self.object_map[name], transforms, shader)

    def create_assembly_instance(self, name, transforms=None, shader=None):
        return self._instance(abi.PSY_INSTANCE_ASSEMBLY, self.assembly_map[name], transforms, shader)


class Camera:
    """basics/camera.hpp:14-69; transforms map camera space to world."""

    def __init__(self, transforms, fovs, aperture_radii=(), focus_distances=()):
        self.transforms = [np.asarray(t, np.float32).reshape(4, 4) for t in transforms]
        self.fovs = [float(f) for f in fovs]
        self.aperture_radii = [float(a) for a in aperture_radii]
        self.focus_distances = [float(f) for f in focus_distances]


class SceneHandle:
    """A psy_scene_desc plus the buffers it points into."""

    def __init__(self, desc, keep):
        self.desc = desc
        self._keep = keep

    def ptr(self):
        return C.byref(self.desc)


def _fptr(arr):
    return arr.ctypes.data_as(C.POINTER(C.c_float))


class Scene:
    def __init__(self, camera, root, background=(0.0, 0.0, 0.0), dice_rate=0.25, min_upoly_size=0.00001,
                 rect_light_fix=False):
        self.camera = camera
        self.root = root
        self.background = tuple(float(b) for b in background)
        self.dice_rate = float(dice_rate)
        self.min_upoly_size = float(min_upoly_size)
        # PSY_SCENE_RECT_LIGHT_FIX: rectangle lights test y against their
        # height (the reference tests it against the width)
        self.rect_light_fix = bool(rect_light_fix)
        self.shaders = []
        self.shader_map = {}

    def add_surface_shader(self, name, kind, color, roughness=0.1, tail_shape=2.0, fresnel=0.25):
        """A LambertShader, EmitShader or GTRShader (shading/surface_shader.hpp:28-71); the last three
        arguments are the GTR shader's (parser/parser.cpp:511-546 defaults)."""
        self.shaders.append((int(kind), tuple(float(c) for c in color), float(roughness), float(tail_shape),
                             float(fresnel)))
        self.shader_map[name] = len(self.shaders) - 1
        return True

    def assemblies(self):
        """All assemblies, root first, in the order used by the description."""
        order, seen = [], {}

        def visit(a):
            if id(a) in seen:
                return
            seen[id(a)] = len(order)
            order.append(a)
            for s in a.assemblies:
                visit(s)

        visit(self.root)
        return order, seen

    def to_desc(self):
        keep = []
        order, index_of = self.assemblies()
        adescs = (abi.AssemblyDesc * len(order))()
        for ai, a in enumerate(order):
            objs = (abi.ObjectDesc * max(1, len(a.objects)))()
            for i, o in enumerate(a.objects):
                keep.append(o.data)
                if o.kind == abi.PSY_OBJ_SUBDIVISION:
                    keep.extend([o.face_vert_counts, o.face_vert_indices])
                    ip = C.POINTER(C.c_int32)
                    objs[i] = abi.ObjectDesc(o.kind, 1, _fptr(o.data), o.data.shape[0], o.face_vert_counts.size,
                                             o.face_vert_counts.ctypes.data_as(ip),
                                             o.face_vert_indices.ctypes.data_as(ip))
                else:
                    objs[i] = abi.ObjectDesc(o.kind, o.time_samples, _fptr(o.data))
            insts = (abi.InstanceDesc * max(1, len(a.instances)))()
            for i, (kind, idx, ti, tc, sh) in enumerate(a.instances):
                shi = -1 if sh is None else self.shader_map[sh]
                insts[i] = abi.InstanceDesc(kind, idx, ti, tc, shi)
            xf = np.ascontiguousarray(np.array(a.xforms, np.float32).reshape(-1, 16)) if a.xforms \
                else np.zeros((1, 16), np.float32)
            subs = np.array([index_of[id(s)] for s in a.assemblies] or [0], np.int32)
            keep.extend([objs, insts, xf, subs])
            adescs[ai] = abi.AssemblyDesc(objs, len(a.objects), insts, len(a.instances), _fptr(xf),
                                          len(a.xforms), subs.ctypes.data_as(C.POINTER(C.c_int32)),
                                          len(a.assemblies))
        shaders = (abi.ShaderDesc * max(1, len(self.shaders)))()
        for i, (kind, col, rough, tail, fres) in enumerate(self.shaders):
            shaders[i] = abi.ShaderDesc(kind, (C.c_float * 3)(*col), rough, tail, fres)
        cam = self.camera
        cxf = np.ascontiguousarray(np.array(cam.transforms, np.float32).reshape(-1, 16))
        fovs = np.array(cam.fovs or [0.0], np.float32)
        aps = np.array(cam.aperture_radii or [0.0], np.float32)
        fds = np.array(cam.focus_distances or [1.0], np.float32)
        keep.extend([adescs, shaders, cxf, fovs, aps, fds])
        cdesc = abi.CameraDesc(_fptr(cxf), len(cam.transforms), _fptr(fovs), len(cam.fovs), _fptr(aps),
                               len(cam.aperture_radii), _fptr(fds), len(cam.focus_distances))
        desc = abi.SceneDesc(adescs, len(order), shaders, len(self.shaders), cdesc,
                             (C.c_float * 3)(*self.background), self.dice_rate, self.min_upoly_size,
                             abi.PSY_SCENE_RECT_LIGHT_FIX if self.rect_light_fix else 0)
        return SceneHandle(desc, keep)


# ---- matrices (row-vector convention) --------------------------------------
def identity():
    return np.eye(4, dtype=np.float32)


def translate(x, y, z):
    """World -> local transform of an instance placed at (x, y, z)."""
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = (-x, -y, -z)
    return m


def look_at(eye, target, up=(0.0, 0.0, 1.0)):
    """Camera-to-world matrix: camera +z looks from eye to target, +y is up."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    m = np.eye(4)
    m[0, :3] = r
    m[1, :3] = u
    m[2, :3] = f
    m[3, :3] = eye
    return m.astype(np.float32)
