"""Synthetic benchmark scenes (the workloads the measurements are quoted on).

Scene generation uses the reference's JKISS generator (utils/rng.hpp:27-162)
restated here, so a seed fully determines a scene.

  sphere_scene      64 spheres r~U[0.2,1.0], centres in [-10,10]^2 x [0,5],
                    one 40x40 bilinear ground quad, two sphere lights r=0.5
                    colour (50,50,50), per-object grey Lambert albedo
                    U[0.2,0.8], pinhole camera.
  heightfield_scene 64x64 bilinear patches over [-5,5]^2, heights a sum of 8
                    seeded sines (A~U[0.05,0.3], f~U[1,8], direction and
                    phase ~U[0,2pi)), one 2x2 rectangle light colour
                    (20,20,20), grey Lambert 0.5, pinhole camera; split-diced
                    at dice_rate 1.0 (about one micropolygon per pixel at
                    1024x1024).
  bicubic_sphere_scene
                    4,096 bicubic patches tiling a unit sphere (latitude /
                    longitude grid, control points jittered by N(0, 0.02)),
                    instanced 16 times through a two-level hierarchy of
                    random rotation / scale / translation placements with
                    two transform samples each (the motion between samples
                    is a translation and scale change under a fixed
                    rotation), two sphere lights, grey Lambert, thin-lens
                    camera with aperture radius 0.05.
  many_light_scene  sphere_scene plus N sphere lights (radii log-uniform in
                    [0.01, 0.2], 0.5-3 above the ground, grey emission scale
                    U[0.1, 10]).
  subdivision_scene heightfield_scene's 65x65 vertices as one Catmull-Clark
                    control mesh (an open sheet of 64x64 quads) with 25
                    rounded boxes on it (one control mesh, a cube of 4x4
                    quads a side, instanced at random rotations about z),
                    the heightfield's light and camera.
"""
import math

import numpy as np

from . import abi
from .scene import (Assembly, Bicubic, Bilinear, Camera, RectangleLight, Scene, Sphere, SphereLight,
                    SubdivisionSurface, look_at, translate)


class JKiss:
    """The reference RNG (utils/rng.hpp), 32-bit seeding variant."""

    M = 0xFFFFFFFF

    def __init__(self, seed):
        s = seed & self.M
        self.x = ((s + 5) * 3885701021) & self.M
        self.y = ((s + 43) * 653005939) & self.M
        self.z = ((s + 13) * 1264700623) & self.M
        self.c = ((s + 67) * 37452703) & self.M
        self.next_uint()
        self.next_uint()

    def next_uint(self):
        M = self.M
        self.x = (314527869 * self.x + 1234567) & M
        y = self.y
        y ^= (y << 5) & M
        y ^= y >> 7
        y ^= (y << 22) & M
        self.y = y
        t = 4294584393 * self.z + self.c
        self.c = (t >> 32) & M
        self.z = t & M
        return (self.x + self.y + self.z) & M

    def next_float(self):
        a = (self.next_uint() >> 9) | 0x3F800000
        return float(np.array([a], np.uint32).view(np.float32)[0]) - 1.0

    def uniform(self, lo, hi):
        return lo + (hi - lo) * self.next_float()


def sphere_scene(seed=1, n_spheres=64, n_lights=2, light_radius=0.5, light_color=(50.0, 50.0, 50.0),
                 fov_deg=60.0, camera_eye=(0.0, -24.0, 14.0), camera_target=(0.0, 0.0, 1.5)):
    rng = JKiss(seed)
    root = Assembly()
    sc = Scene(Camera([look_at(camera_eye, camera_target)], [math.radians(fov_deg)]), root,
               background=(0.0, 0.0, 0.0))
    for i in range(n_spheres):
        r = rng.uniform(0.2, 1.0)
        c = (rng.uniform(-10, 10), rng.uniform(-10, 10), rng.uniform(0, 5))
        alb = rng.uniform(0.2, 0.8)
        sc.add_surface_shader(f"s{i}", abi.PSY_SHADER_LAMBERT, (alb, alb, alb))
        root.add_object(f"sphere{i}", Sphere(c, r))
        root.create_object_instance(f"sphere{i}", [], f"s{i}")
    alb = rng.uniform(0.2, 0.8)
    sc.add_surface_shader("ground", abi.PSY_SHADER_LAMBERT, (alb, alb, alb))
    root.add_object("ground", Bilinear([(-20, -20, 0), (20, -20, 0), (-20, 20, 0), (20, 20, 0)]))
    root.create_object_instance("ground", [], "ground")
    for i in range(n_lights):
        c = (rng.uniform(-8, 8), rng.uniform(-8, 8), rng.uniform(6, 9))
        root.add_object(f"light{i}", SphereLight(c, light_radius, light_color))
        root.create_object_instance(f"light{i}", [], None)
    return sc


def heightfield(seed, grid=64, extent=10.0, n_sines=8):
    """Vertex heights of the sine-sum heightfield, (grid+1)^2 float32."""
    rng = JKiss(seed)
    waves = []
    for _ in range(n_sines):
        a = rng.uniform(0.05, 0.3)
        f = rng.uniform(1.0, 8.0)
        th = rng.uniform(0.0, 2 * math.pi)
        ph = rng.uniform(0.0, 2 * math.pi)
        waves.append((a, f, th, ph))
    xs = np.linspace(-extent / 2, extent / 2, grid + 1)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    Z = np.zeros_like(X)
    for a, f, th, ph in waves:
        Z += a * np.sin(f * (X * math.cos(th) + Y * math.sin(th)) + ph)
    return X.astype(np.float32), Y.astype(np.float32), Z.astype(np.float32)


def heightfield_scene(seed=7, grid=64, extent=10.0, dice_rate=1.0, light_pos=(0.0, -3.0, 13.5),
                      light_size=2.0, light_color=(20.0, 20.0, 20.0), fov_deg=36.0,
                      camera_eye=(0.0, -1.5, 11.5), camera_target=(0.0, 0.3, 0.0)):
    """The camera frames the field so that every 1024x1024 primary ray lands on it."""
    X, Y, Z = heightfield(seed, grid, extent)
    root = Assembly()
    sc = Scene(Camera([look_at(camera_eye, camera_target)], [math.radians(fov_deg)]), root,
               background=(0.0, 0.0, 0.0), dice_rate=dice_rate)
    sc.add_surface_shader("ground", abi.PSY_SHADER_LAMBERT, (0.5, 0.5, 0.5))
    for j in range(grid):
        for i in range(grid):
            v1 = (X[j, i], Y[j, i], Z[j, i])
            v2 = (X[j, i + 1], Y[j, i + 1], Z[j, i + 1])
            v3 = (X[j + 1, i], Y[j + 1, i], Z[j + 1, i])
            v4 = (X[j + 1, i + 1], Y[j + 1, i + 1], Z[j + 1, i + 1])
            name = f"p{j}_{i}"
            root.add_object(name, Bilinear([v1, v2, v3, v4]))
            root.create_object_instance(name, [], "ground")
    root.add_object("light", RectangleLight(light_size, light_size, light_color))
    root.create_object_instance("light", [translate(*light_pos)], None)
    return sc


def _gauss(rng):
    """Standard normal sample from two JKISS uniforms (Box-Muller)."""
    u1 = max(rng.next_float(), 1e-12)
    u2 = rng.next_float()
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def _rotation(rng):
    """A random rotation (row-vector 4x4) about a random axis."""
    ax = np.array([_gauss(rng), _gauss(rng), _gauss(rng)])
    ax /= max(np.linalg.norm(ax), 1e-9)
    ang = rng.uniform(0.0, 2.0 * math.pi)
    c, s = math.cos(ang), math.sin(ang)
    x, y, z = ax
    r = np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                  [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                  [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])
    m = np.eye(4)
    m[:3, :3] = r.T  # row-vector convention
    return m


def placement(rot, scl, pos):
    """World -> local transform of an instance scaled by `scl`, rotated by
    `rot` and placed at `pos` (local -> world is S * R * T, row vectors)."""
    S = np.diag([scl, scl, scl, 1.0])
    T = np.eye(4)
    T[3, :3] = pos
    return np.linalg.inv(S @ rot @ T).astype(np.float32)


def sphere_patches(rng, grid=64, jitter=0.02):
    """Bicubic patches (16 control points each) over a latitude/longitude
    grid of the unit sphere, control points jittered by N(0, jitter)."""
    patches = []
    for j in range(grid):
        for i in range(grid):
            pts = []
            for l in range(4):
                v = math.pi * (j + l / 3.0) / grid
                for k in range(4):
                    u = 2.0 * math.pi * (i + k / 3.0) / grid
                    p = [math.sin(v) * math.cos(u), math.sin(v) * math.sin(u), math.cos(v)]
                    pts.append([p[0] + jitter * _gauss(rng), p[1] + jitter * _gauss(rng), p[2] + jitter * _gauss(rng)])
            patches.append(np.array(pts, np.float32))
    return patches


def bicubic_sphere_scene(seed=3, grid=64, jitter=0.02, n_outer=4, n_inner=4, motion=True, aperture=0.05,
                         fov_deg=45.0, camera_eye=(0.0, -26.0, 9.0), camera_target=(0.0, 0.0, 0.0),
                         dice_rate=1.0, light_color=(60.0, 60.0, 60.0)):
    rng = JKiss(seed)
    sphere = Assembly()
    for k, pts in enumerate(sphere_patches(rng, grid, jitter)):
        sphere.add_object(f"p{k}", Bicubic(pts))
        sphere.create_object_instance(f"p{k}", [], None)
    mid = Assembly()
    mid.add_assembly("sphere", sphere)
    for k in range(n_inner):
        rot = _rotation(rng)
        s = rng.uniform(0.5, 1.5)
        pos = np.array([rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(-3, 3)])
        xs = [placement(rot, s, pos)]
        if motion:
            xs.append(placement(rot, s * rng.uniform(0.95, 1.05), pos + np.array([rng.uniform(-0.3, 0.3) for _ in range(3)])))
        mid.create_assembly_instance("sphere", xs, None)
    root = Assembly()
    root.add_assembly("mid", mid)
    for k in range(n_outer):
        rot = _rotation(rng)
        s = rng.uniform(0.8, 1.2)
        pos = np.array([rng.uniform(-8, 8), rng.uniform(-8, 8), rng.uniform(-3, 3)])
        xs = [placement(rot, s, pos)]
        if motion:
            xs.append(placement(rot, s, pos + np.array([rng.uniform(-0.5, 0.5) for _ in range(3)])))
        root.create_assembly_instance("mid", xs, "grey")
    root.add_object("lamp0", SphereLight((6.0, -6.0, 12.0), 1.5, light_color))
    root.create_object_instance("lamp0", [], None)
    root.add_object("lamp1", SphereLight((-8.0, 4.0, 10.0), 1.0, light_color))
    root.create_object_instance("lamp1", [], None)
    eye = np.asarray(camera_eye, np.float64)
    focus = float(np.linalg.norm(eye - np.asarray(camera_target, np.float64)))
    cam = Camera([look_at(camera_eye, camera_target)], [math.radians(fov_deg)],
                 aperture_radii=[aperture], focus_distances=[focus])
    sc = Scene(cam, root, background=(0.05, 0.05, 0.05), dice_rate=dice_rate)
    sc.add_surface_shader("grey", abi.PSY_SHADER_LAMBERT, (0.6, 0.6, 0.6))
    return sc


def adjoint_scene(seed=5, n_spheres=256, n_lights=1000):
    """The differentiable-render scene: n_spheres spheres and the ground as
    in sphere_scene (one Lambert shader each, albedo U[0.2,0.8]), lit only
    by n_lights small sphere lights as in many_light_scene."""
    return many_light_scene(seed=seed, n_lights=n_lights, n_spheres=n_spheres, big_lights=0)


def many_light_scene(seed=1, n_lights=10000, big_lights=2, **kw):
    """sphere_scene plus n_lights small sphere lights above the ground."""
    sc = sphere_scene(seed=seed, n_lights=big_lights, **kw)
    rng = JKiss(seed + 1000)
    root = sc.root
    for i in range(n_lights):
        r = math.exp(rng.uniform(math.log(0.01), math.log(0.2)))
        c = (rng.uniform(-10, 10), rng.uniform(-10, 10), rng.uniform(0.5, 3.0))
        e = rng.uniform(0.1, 10.0)
        root.add_object(f"ml{i}", SphereLight(c, r, (e, e, e)))
        root.create_object_instance(f"ml{i}", [], None)
    return sc


def cube_mesh(k=4, half=1.0):
    """Closed quad control mesh of a cube with k x k quads a side, faces
    oriented outwards: (verts (n,3) float32, face vertex counts, indices)."""
    ids = {}
    verts = []

    def vid(p):
        if p not in ids:
            ids[p] = len(verts)
            verts.append([(c / k * 2.0 - 1.0) * half for c in p])
        return ids[p]
    faces = []
    for axis in range(3):
        for side in (0, k):
            u, v = (axis + 1) % 3, (axis + 2) % 3
            for a in range(k):
                for b in range(k):
                    quad = []
                    for (da, db) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = [0, 0, 0]
                        p[axis] = side
                        p[u] = a + da
                        p[v] = b + db
                        quad.append(vid(tuple(p)))
                    faces.append(quad if side == k else quad[::-1])
    return (np.asarray(verts, np.float32), np.full(len(faces), 4, np.int32),
            np.asarray([i for f in faces for i in f], np.int32))


def subdivision_scene(seed=7, grid=64, extent=10.0, dice_rate=1.0, n_boxes=5, light_pos=(0.0, -3.0, 13.5),
                      light_size=2.0, light_color=(20.0, 20.0, 20.0), fov_deg=36.0,
                      camera_eye=(0.0, -1.5, 11.5), camera_target=(0.0, 0.3, 0.0)):
    """Catmull-Clark surfaces: the heightfield as one control mesh and
    rounded boxes standing on it (object/subdivision_surface.hpp)."""
    X, Y, Z = heightfield(seed, grid, extent)
    root = Assembly()
    sc = Scene(Camera([look_at(camera_eye, camera_target)], [math.radians(fov_deg)]), root,
               background=(0.0, 0.0, 0.0), dice_rate=dice_rate)
    sc.add_surface_shader("ground", abi.PSY_SHADER_LAMBERT, (0.5, 0.5, 0.5))
    sc.add_surface_shader("box", abi.PSY_SHADER_LAMBERT, (0.7, 0.7, 0.7))
    V = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    n = grid + 1
    F = [[j * n + i, j * n + i + 1, (j + 1) * n + i + 1, (j + 1) * n + i] for j in range(grid) for i in range(grid)]
    root.add_object("field", SubdivisionSurface(V, [4] * len(F), np.asarray(F, np.int32).ravel()))
    root.create_object_instance("field", [], "ground")
    root.add_object("box", SubdivisionSurface(*cube_mesh(4, 0.45)))
    rng = JKiss(seed + 1)
    for b in range(n_boxes * n_boxes):
        j, i = divmod(b, n_boxes)
        x = (i + 0.5) / n_boxes * extent * 0.8 - extent * 0.4
        y = (j + 0.5) / n_boxes * extent * 0.8 - extent * 0.4
        gi = int(round((x + extent / 2) / extent * grid))
        gj = int(round((y + extent / 2) / extent * grid))
        z = float(Z[gj, gi]) + 0.35
        a = rng.uniform(0.0, 2 * math.pi)
        c, s_ = math.cos(a), math.sin(a)
        rot = np.array([[c, s_, 0, 0], [-s_, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32)
        root.create_object_instance("box", [translate(x, y, z) @ rot], "box")
    root.add_object("light", RectangleLight(light_size, light_size, light_color))
    root.create_object_instance("light", [translate(*light_pos)], None)
    return sc
