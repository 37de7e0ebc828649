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
"""
import math

import numpy as np

from . import abi
from .scene import Assembly, Bilinear, Camera, RectangleLight, Scene, Sphere, SphereLight, look_at, translate


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
