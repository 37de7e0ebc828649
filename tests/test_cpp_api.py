"""The C++ host interface (include/psychopath_gpu.hpp: the reference's
Assembly / Scene / Camera / shader / object classes, Tracer::trace and
PathTraceIntegrator over the C ABI), driven by tests/cpp/test_cpp_api.cpp
built with g++.  On the CPU: the C++ classes flatten the scene exactly as
the Python scene description does, and a missing device is reported with
the library's error text."""
import os
import subprocess

import numpy as np
import pytest

from psychopath_cpp_amd import abi
from psychopath_cpp_amd.scene import (Assembly, Bicubic, Bilinear, Camera, RectangleLight, Scene, Sphere,
                                      SphereLight, SubdivisionSurface, translate)

HERE = os.path.dirname(os.path.abspath(__file__))
CAM = np.array([1.0, -0.0, 0.0, 0.0, 0.0, 0.4472135901451111, 0.8944271802902222, 0.0, 0.0, 0.8944271802902222,
                -0.4472135901451111, 0.0, 0.0, -9.0, 5.0, 1.0], np.float32).reshape(4, 4)


def blob_mesh():
    """The Catmull-Clark cube of half-size 0.6 test_cpp_api.cpp builds."""
    v = np.array([[0.6 if x else -0.6, 0.6 if y else -0.6, 0.6 if z else -0.6]
                  for x in range(2) for y in range(2) for z in range(2)], np.float32)
    idx = [0, 2, 6, 4, 1, 5, 7, 3, 0, 4, 5, 1, 2, 3, 7, 6, 0, 1, 3, 2, 4, 6, 7, 5]
    return v, [4] * 6, idx


def cpp_equivalent_scene():
    """The scene test_cpp_api.cpp builds, through the Python description."""
    root = Assembly()
    sc = Scene(Camera([CAM], [0.8726646], [0.0], [10.0]), root, background=(0.05, 0.05, 0.05))
    sc.add_surface_shader("grey", abi.PSY_SHADER_LAMBERT, (0.6, 0.6, 0.6))
    sc.add_surface_shader("dark", abi.PSY_SHADER_LAMBERT, (0.3, 0.3, 0.3))
    root.add_object("ground", Bilinear([(-8, -8, 0), (8, -8, 0), (-8, 8, 0), (8, 8, 0)]))
    root.create_object_instance("ground", [], "grey")
    root.add_object("ball", Sphere((0.0, 0.0, 1.0), 1.0))
    root.create_object_instance("ball", [], "dark")
    cp = [(-3.0 + i, 2.0 + j, 1.5 if (i in (1, 2) and j in (1, 2)) else 0.2) for j in range(4) for i in range(4)]
    root.add_object("hill", Bicubic(cp))
    root.create_object_instance("hill", [translate(0.5, 0.0, 0.0)], "grey")
    root.add_object("blob", SubdivisionSurface(*blob_mesh()))
    root.create_object_instance("blob", [translate(2.5, 2.0, 0.8)], "grey")
    lamp = Assembly()
    lamp.add_object("bulb", SphereLight((0.0, 0.0, 0.0), 0.4, (30.0, 30.0, 30.0)))
    lamp.create_object_instance("bulb", [], None)
    lamp.add_object("panel", RectangleLight(1.0, 1.0, (10.0, 10.0, 10.0)))
    lamp.create_object_instance("panel", [translate(0.0, 0.0, 1.0)], None)
    root.add_assembly("lamp", lamp)
    root.create_assembly_instance("lamp", [translate(-2.5, 1.0, 4.0)], None)
    root.create_assembly_instance("lamp", [translate(2.0, -1.0, 3.0)], None)
    return sc


def cpp_binary():
    from tests.cpp import build as cb
    if not os.path.exists(cb.BIN) or os.path.getmtime(cb.BIN) < max(
            os.path.getmtime(cb.SRC), os.path.getmtime(os.path.join(cb.ROOT, "include", "psychopath_gpu.hpp"))):
        cb.build()
    return cb.BIN


def _describe_python(sc):
    handle = sc.to_desc()  # keeps the buffers the description points into
    d = handle.desc
    lines = [f"assemblies {d.n_assemblies}", f"shaders {d.n_shaders}"]
    for i in range(d.n_shaders):
        c = d.shaders[i].color
        lines.append(f"shader {i} {d.shaders[i].kind} {c[0]:.7g} {c[1]:.7g} {c[2]:.7g}")
    for a in range(d.n_assemblies):
        A = d.assemblies[a]
        xs = sum(float(A.transforms[k]) * (k % 16 + 1) for k in range(16 * A.n_transforms))
        lines.append(f"assembly {a} objects {A.n_objects} instances {A.n_instances} transforms {A.n_transforms} "
                     f"children {A.n_assemblies} xfsum {xs:.6f}")
        for i in range(A.n_objects):
            o = A.objects[i]
            per = 3 * o.n_vertices if o.kind == abi.PSY_OBJ_SUBDIVISION else abi.FLOATS_PER_KIND[o.kind]
            n = per * o.time_samples
            s = sum(float(o.data[k]) for k in range(n))
            lines.append(f"object {a} {i} kind {o.kind} ts {o.time_samples} sum {s:.6f}")
            if o.kind == abi.PSY_OBJ_SUBDIVISION:
                sc_ = sum(int(o.face_vert_counts[f]) for f in range(o.n_faces))
                si = sum(int(o.face_vert_indices[k]) * (k + 1) for k in range(sc_))
                lines.append(f"mesh {a} {i} verts {o.n_vertices} faces {o.n_faces} counts {sc_} indices {si}")
        for i in range(A.n_instances):
            x = A.instances[i]
            lines.append(f"instance {a} {i} type {x.type} data {x.data_index} xf {x.transform_index}+"
                         f"{x.transform_count} shader {x.shader_index}")
    return lines


def test_cpp_interface_flattens_like_python():
    out = subprocess.run([cpp_binary(), "describe"], check=True, stdout=subprocess.PIPE, text=True).stdout
    assert out.strip().splitlines() == _describe_python(cpp_equivalent_scene())


def test_cpp_parser_reads_psy_like_python(tmp_path):
    """psy::Parser (the reference's Parser over psy_parse_scene) reads a .psy
    file of the same scene into the same description the scene API builds,
    for each of its frames."""
    from tests import _psywriter
    sc = cpp_equivalent_scene()
    text = _psywriter.write_psy(sc, 32, 24, 4, 1, "one") + _psywriter.write_psy(sc, 8, 6, 2, 9, "two")
    path = tmp_path / "frames.psy"
    path.write_text(text)
    out = subprocess.run([cpp_binary(), "psyparse", str(path)], check=True, stdout=subprocess.PIPE,
                         text=True).stdout.strip().splitlines()
    want = _describe_python(sc)
    frames = [i for i, l in enumerate(out) if l.startswith("frame ")]
    assert len(frames) == 2
    assert out[frames[0]] == "frame one 32x24 spp 4 seed 1 dice 0.25 output "
    assert out[frames[1]] == "frame two 8x6 spp 2 seed 9 dice 0.25 output "
    assert out[frames[0] + 1:frames[1]] == want
    assert out[frames[1] + 1:] == want


def test_cpp_interface_reports_missing_device():
    import torch
    out = subprocess.run([cpp_binary(), "nodevice"], check=True, stdout=subprocess.PIPE, text=True).stdout
    if torch.cuda.is_available():
        assert out.strip() == "finalized"
    else:
        assert out.startswith("error psy_scene_create") or out.startswith("error psy_ctx_create")
        assert "no HIP device" in out


def test_cpp_subdivision_patches_match_python():
    """SubdivisionSurface::patches (the C++ class over
    psy_subdivision_patches) gives the patches the Python wrapper gives."""
    from psychopath_cpp_amd import gpu
    out = subprocess.run([cpp_binary(), "subdiv"], check=True, stdout=subprocess.PIPE, text=True).stdout.split()
    P = gpu.subdivision_patches(*blob_mesh()).astype(np.float64)
    w = (np.arange(len(P)) % 7 + 1)[:, None]
    s = float(((P[..., 0] + 2.0 * P[..., 1] + 3.0 * P[..., 2]) * w).sum())
    assert int(out[1]) == len(P)
    assert float(out[3]) == pytest.approx(s, rel=1e-10)
