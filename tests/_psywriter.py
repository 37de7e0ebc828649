"""Test helper: writes a Python-built Scene (psychopath_cpp_amd.scene) as
.psy text in the reference's scene format, for parse round trips.  Shaders
go in the root assembly (sub-assemblies find them up the parent chain);
objects and sub-assemblies get generated names."""
import numpy as np

from psychopath_cpp_amd import abi

DEG = np.float32(3.1415926536) / np.float32(180.0)


def num(x):
    """A float32 in positional notation that reads back exactly (the .psy
    number pattern has no exponent)."""
    return np.format_float_positional(np.float32(x), unique=True, trim="-")


def nums(a):
    return " ".join(num(x) for x in np.asarray(a, np.float32).reshape(-1))


def _assembly(sc, a, names, indent, out):
    pad = "    " * indent
    obj_names = []
    for i, o in enumerate(a.objects):
        nm = f"$o{len(names)}_{i}"
        obj_names.append(nm)
        d = o.data
        if o.kind == abi.PSY_OBJ_SPHERE:
            assert o.time_samples == 1, "the .psy Sphere has one time sample"
            out.append(f"{pad}Sphere {nm} {{ Location [{nums(d[0, :3])}] Radius [{num(d[0, 3])}] }}")
        elif o.kind in (abi.PSY_OBJ_BILINEAR, abi.PSY_OBJ_BICUBIC):
            t = "BilinearPatch" if o.kind == abi.PSY_OBJ_BILINEAR else "BicubicPatch"
            vs = " ".join(f"Vertices [{nums(row)}]" for row in d)
            out.append(f"{pad}{t} {nm} {{ {vs} }}")
        elif o.kind == abi.PSY_OBJ_SPHERE_LIGHT:
            props = " ".join(f"Location [{nums(r[:3])}] Radius [{num(r[3])}] Color [{nums(r[4:7])}]" for r in d)
            out.append(f"{pad}SphereLight {nm} {{ {props} }}")
        elif o.kind == abi.PSY_OBJ_SUBDIVISION:
            out.append(f"{pad}SubdivisionSurface {nm} {{ Vertices [{nums(d)}] "
                       f"FaceVertCounts [{' '.join(str(int(c)) for c in o.face_vert_counts)}] "
                       f"FaceVertIndices [{' '.join(str(int(i)) for i in o.face_vert_indices)}] }}")
        elif o.kind == abi.PSY_OBJ_RECTANGLE_LIGHT:
            props = " ".join(f"Dimensions [{nums(r[:2])}] Color [{nums(r[2:5])}]" for r in d)
            out.append(f"{pad}RectangleLight {nm} {{ {props} }}")
    names.append(obj_names)
    sub_names = []
    for j, s in enumerate(a.assemblies):
        nm = f"$a{len(names)}_{j}"
        sub_names.append(nm)
        out.append(f"{pad}Assembly {nm} {{")
        _assembly(sc, s, names, indent + 1, out)
        out.append(f"{pad}}}")
    for kind, idx, ti, tc, sh in a.instances:
        data = obj_names[idx] if kind == abi.PSY_INSTANCE_OBJECT else sub_names[idx]
        parts = [f"Data [{data}]"]
        for k in range(ti, ti + tc):
            parts.append(f"Transform [{nums(a.xforms[k])}]")
        if sh is not None:
            parts.append(f"SurfaceShaderBind [$s{sc.shader_map[sh]}]")
        out.append(f"{pad}Instance {{ {' '.join(parts)} }}")


def write_psy(sc, width=64, height=48, spp=4, seed=1, name="frame"):
    out = [f"Scene ${name} {{",
           f"    RenderSettings {{ Resolution [{width} {height}] SamplesPerPixel [{spp}] "
           f"DicingRate [{num(sc.dice_rate)}] Seed [{seed}] }}"]
    cam = ["    Camera {"]
    for f in sc.camera.fovs:
        cam.append(f"        Fov [{num(np.float32(f) / DEG)}]")
    for v in sc.camera.focus_distances:
        cam.append(f"        FocalDistance [{num(v)}]")
    for v in sc.camera.aperture_radii:
        cam.append(f"        ApertureRadius [{num(v)}]")
    for t in sc.camera.transforms:
        cam.append(f"        Transform [{nums(t)}]")
    out += cam + ["    }",
                  f"    World {{ BackgroundShader {{ Type [Color] Color [{nums(sc.background)}] }} }}",
                  "    Assembly {"]
    for i, (kind, col, rough, tail, fres) in enumerate(sc.shaders):
        if kind == abi.PSY_SHADER_GTR:
            out.append(f"        SurfaceShader $s{i} {{ Type [GTR] Color [{nums(col)}] Roughness [{num(rough)}] "
                       f"TailShape [{num(tail)}] Fresnel [{num(fres)}] }}")
            continue
        t = "Lambert" if kind == abi.PSY_SHADER_LAMBERT else "Emit"
        out.append(f"        SurfaceShader $s{i} {{ Type [{t}] Color [{nums(col)}] }}")
    _assembly(sc, sc.root, [], 2, out)
    out += ["    }", "}"]
    return "\n".join(out) + "\n"
