"""Control meshes for the subdivision-surface tests: (verts (n,3) float32,
face vertex counts, concatenated face vertex indices), faces oriented
counter-clockwise seen from outside (or from +z for open sheets)."""
import numpy as np


def _pack(verts, faces):
    return (np.asarray(verts, np.float32), np.array([len(f) for f in faces], np.int32),
            np.array([v for f in faces for v in f], np.int32))


def cube(half=1.0):
    V = [[x, y, z] for x in (-half, half) for y in (-half, half) for z in (-half, half)]

    def vid(x, y, z):
        return (x > 0) * 4 + (y > 0) * 2 + (z > 0)
    F = [[vid(-1, -1, -1), vid(-1, 1, -1), vid(1, 1, -1), vid(1, -1, -1)],
         [vid(-1, -1, 1), vid(1, -1, 1), vid(1, 1, 1), vid(-1, 1, 1)],
         [vid(-1, -1, -1), vid(1, -1, -1), vid(1, -1, 1), vid(-1, -1, 1)],
         [vid(-1, 1, -1), vid(-1, 1, 1), vid(1, 1, 1), vid(1, 1, -1)],
         [vid(-1, -1, -1), vid(-1, -1, 1), vid(-1, 1, 1), vid(-1, 1, -1)],
         [vid(1, -1, -1), vid(1, 1, -1), vid(1, 1, 1), vid(1, -1, 1)]]
    return _pack(V, F)


def grid(n, m, ripple=0.1, seed=None):
    """Open n x m quad sheet in the xy plane, z rippled."""
    rng = np.random.default_rng(seed) if seed is not None else None
    V = []
    for j in range(m + 1):
        for i in range(n + 1):
            z = ripple * np.sin(i + 2 * j) if rng is None else ripple * rng.normal()
            V.append([i, j, z])
    F = [[j * (n + 1) + i, j * (n + 1) + i + 1, (j + 1) * (n + 1) + i + 1, (j + 1) * (n + 1) + i]
         for j in range(m) for i in range(n)]
    return _pack(V, F)


def torus(n=8, m=6, R=3.0, r=1.0, jitter=0.0, seed=3):
    """Closed n x m quad torus: every vertex regular; face (i, j) is
    (P[j][i], P[j][i+1], P[j+1][i+1], P[j+1][i]) with indices mod n, m."""
    rng = np.random.default_rng(seed)
    V = []
    for j in range(m):
        for i in range(n):
            a, b = 2 * np.pi * i / n, 2 * np.pi * j / m
            p = np.array([(R + r * np.cos(b)) * np.cos(a), (R + r * np.cos(b)) * np.sin(a), r * np.sin(b)])
            V.append(p + jitter * rng.normal(size=3))
    F = [[j * n + i, j * n + (i + 1) % n, ((j + 1) % m) * n + (i + 1) % n, ((j + 1) % m) * n + i]
         for j in range(m) for i in range(n)]
    return _pack(V, F)


def pentaprism(h=1.0):
    """Closed: two pentagons and five quads; every vertex has valence 3."""
    ang = 2 * np.pi * np.arange(5) / 5
    V = [[np.cos(a), np.sin(a), -h] for a in ang] + [[np.cos(a), np.sin(a), h] for a in ang]
    F = [[4, 3, 2, 1, 0], [5, 6, 7, 8, 9]] + [[i, (i + 1) % 5, 5 + (i + 1) % 5, 5 + i] for i in range(5)]
    return _pack(V, F)


def tetra():
    V = [[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]
    F = [[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]]
    return _pack(V, F)


def fan(n=5, ripple=0.15):
    """Open disc: n triangles around a centre, then an annulus of n quads,
    quad k spanning inner vertices k, k+1 and outer vertices k-1, k."""
    V = [[0.0, 0.0, 0.3]]
    for k in range(n):
        a = 2 * np.pi * k / n
        V.append([np.cos(a), np.sin(a), ripple * np.cos(3 * a)])
    for k in range(n):
        a = 2 * np.pi * (k + 0.5) / n
        V.append([2.2 * np.cos(a), 2.2 * np.sin(a), -ripple * np.sin(2 * a)])
    F = [[0, 1 + k, 1 + (k + 1) % n] for k in range(n)]
    F += [[1 + k, 1 + n + (k - 1) % n, 1 + n + k, 1 + (k + 1) % n] for k in range(n)]
    return _pack(V, _orient(V, F))


def lshape(n=6, ripple=0.1):
    """Open L-shaped sheet: its concave corner is a boundary vertex with
    three faces, its five convex corners have one face each."""
    keep = set()
    for j in range(n):
        for i in range(n):
            if not (i >= n // 2 and j >= n // 2):
                keep.add((i, j))
    ids = {}
    V = []
    F = []
    for (i, j) in sorted(keep, key=lambda t: (t[1], t[0])):
        quad = []
        for (x, y) in ((i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)):
            if (x, y) not in ids:
                ids[(x, y)] = len(V)
                V.append([x, y, ripple * np.cos(1.3 * x + 0.7 * y)])
            quad.append(ids[(x, y)])
        F.append(quad)
    return _pack(V, F)


def _orient(V, F):
    """Orient every face counter-clockwise seen from +z (planar-ish sheets)."""
    V = np.asarray(V, float)
    out = []
    for f in F:
        p = V[f][:, :2]
        area = 0.5 * np.sum(p[:, 0] * np.roll(p[:, 1], -1) - np.roll(p[:, 0], -1) * p[:, 1])
        out.append(list(f) if area > 0 else list(f)[::-1])
    return out


def all_meshes():
    return {"cube": cube(), "grid": grid(6, 5), "strip": grid(5, 1), "torus": torus(jitter=0.1),
            "pentaprism": pentaprism(), "tetra": tetra(), "fan5": fan(5), "fan7": fan(7), "lshape": lshape()}


def subdiv_scene(motion=False):
    """Subdivision surfaces in a scene: a rounded box instanced three times
    (one instance moving when `motion`), a prism inside a sub-assembly
    instanced twice, a valence-7 fan, an L-shaped ground sheet, a sphere
    light and a rectangle light."""
    from psychopath_cpp_amd import abi
    from psychopath_cpp_amd.scene import (Assembly, Camera, RectangleLight, Scene, SphereLight,
                                          SubdivisionSurface, look_at, translate)
    root = Assembly()
    root.add_object("box", SubdivisionSurface(*cube(1.0)))
    root.create_object_instance("box", [translate(-2.5, 0.0, 0.2)], "grey")
    c, s = np.cos(0.6), np.sin(0.6)
    rot = np.array([[c, s, 0, 0], [-s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32)
    ts = [rot @ translate(0.5, 2.5, 0.4)]
    if motion:
        ts.append(rot @ translate(0.9, 2.3, 0.6))
    root.create_object_instance("box", ts, "red")
    root.create_object_instance("box", [translate(3.0, -1.0, 1.5)], "blue")
    sub = Assembly()
    sub.add_object("prism", SubdivisionSurface(*pentaprism(0.8)))
    sub.create_object_instance("prism", [], "blue")
    root.add_assembly("sub", sub)
    root.create_assembly_instance("sub", [translate(0.5, -1.5, 0.0)], None)
    root.create_assembly_instance("sub", [translate(-1.0, 4.5, 0.5)], "grey")
    V, cnt, idx = fan(7, ripple=0.3)
    root.add_object("fan", SubdivisionSurface(V * np.float32(1.2) + np.float32([2.5, 2.5, -0.2]), cnt, idx))
    root.create_object_instance("fan", [], "grey")
    V, cnt, idx = lshape(6, ripple=0.2)
    root.add_object("ground", SubdivisionSurface(V * np.float32(2.2) - np.float32([6.0, 6.0, 1.2]), cnt, idx))
    root.create_object_instance("ground", [], "grey")
    root.add_object("lamp", SphereLight([0, 0, 9], 1.0, (5, 5, 5)))
    root.create_object_instance("lamp", [], None)
    root.add_object("panel", RectangleLight(3.0, 2.0, (4, 4, 4)))
    root.create_object_instance("panel", [translate(2.0, 1.0, 7.0)], None)
    sc = Scene(Camera([look_at((0, -12, 7), (0, 0, 0))], [np.radians(55)]), root, dice_rate=0.5)
    sc.add_surface_shader("grey", abi.PSY_SHADER_LAMBERT, (0.6, 0.6, 0.6))
    sc.add_surface_shader("red", abi.PSY_SHADER_LAMBERT, (0.8, 0.8, 0.8))
    sc.add_surface_shader("blue", abi.PSY_SHADER_LAMBERT, (0.3, 0.3, 0.3))
    return sc
