"""The .psy scene-file parser (psy_parse_scene, psy_parse_dump_tree) through
the C ABI, on the CPU.  The expectations follow the reference's DataTree
lexer (parser/data_tree.cpp) and Parser (parser/parser.cpp): tree dumps in
print_tree's layout, numbers by its -?[0-9]+[.]?[0-9]* pattern, Fov in
degrees, name scopes of scene/assembly.hpp; and round trips of Python-built
scenes through .psy text give back their descriptions bit for bit (the
field of view to an ulp, through degrees)."""
import os

import numpy as np
import pytest

from psychopath_cpp_amd import abi, configs
from psychopath_cpp_amd import gpu, psyfile
from tests import _psywriter as W
from tests import test_gpu_parity as G
from tests import test_gpu_render as R

DEG = np.float32(3.1415926536) / np.float32(180.0)


def test_datatree_lexing():
    text = (
        "# a comment line\n"
        "Top $first\\ name {   # trailing comment\n"
        "  Leaf [ a   b\t\n c ]  Esc [x\\]y\\#z \\\\ w]\n"
        "  Inner $n2 { Deep [ 1 # inside a leaf\n 2 ] }\n"
        "  Empty {}\n"
        "  Unicode$nämé [ünï ✓]\n"
        "}\n"
        "Second [last]\n")
    want = ("Top $first name {\n"
            "    Leaf [a b c]\n"
            "    Esc [x]y#z \\ w]\n"
            "    Inner $n2 {\n"
            "        Deep [1 2]\n"
            "    }\n"
            "    Empty []\n"
            "    Unicode [ünï ✓]\n"
            "}\n"
            "Second [last]\n")
    assert psyfile.dump_tree(text) == want
    # parsing stops at the first token that is not a node type
    assert psyfile.dump_tree("A [1] ] B [2]") == "A [1]\n"
    assert psyfile.dump_tree("") == ""


@pytest.mark.parametrize("text,where", [("A { B [1] ", "line 1"), ("A {\n\n B [1", "line 3"),
                                        ("A {\n ] }", "line 2"), ("A $n B", "line 1")])
def test_datatree_errors(text, where):
    with pytest.raises(gpu.PsyError, match=where):
        psyfile.dump_tree(text)


SCENE = """
Scene $shot\\ 1 {
  Output { Path ["/tmp/out dir/shot.png"] Format [png] }
  RenderSettings { Resolution [320 240] SamplesPerPixel [8] DicingRate [0.5] Seed [7] PixelAspect [1] }
  Camera {
    Fov [45] Fov [50]
    FocalDistance [12.5] ApertureRadius [0.05]
    Transform [1 0 0 0  0 1 0 0  0 0 1 0  0 -5 2 1]
    Transform [1 0 0 0  0 1 0 0  0 0 1 0  0.5 -5 2 1]
  }
  World { BackgroundShader { Type [Color] Color [0.1 0.2 0.3] } DistantDiskLight { Radius [1] } }
  Assembly {
    SurfaceShader $red { Type [Lambert] Color [0.8 0.1 0.1] }
    SurfaceShader $glow { Type [Emit] Color [4 3 2] }
    SurfaceShader $plain { Type [Lambert] }
    SurfaceShader $shiny { Type [GTR] Roughness [0.2] }
    SurfaceShader $weird { Type [OSL] }
    Sphere $ball { Location [1e-5 .5 -.5] Radius [3.] }
    Sphere $ball { Location [9 9 9] }
    BilinearPatch $quad { Vertices [0 0 0  1 0 0  0 1 0  1 1 0.25] }
    BicubicPatch $moving {
      Vertices [0 0 0 1 0 0 2 0 0 3 0 0 0 1 0 1 1 1 2 1 1 3 1 0 0 2 0 1 2 1 2 2 1 3 2 0 0 3 0 1 3 0 2 3 0 3 3 0]
      Vertices [0 0 1 1 0 1 2 0 1 3 0 1 0 1 1 1 1 2 2 1 2 3 1 1 0 2 1 1 2 2 2 2 2 3 2 1 0 3 1 1 3 1 2 3 1 3 3 1]
    }
    SphereLight $lamp { Location [0 0 5] Location [0 0 6] Radius [0.5] Color [10 9 8] }
    RectangleLight $panel { Dimensions [2] Color [1 1 1] }
    SubdivisionSurface $subd { Vertices [0 0 0] }
    Assembly $child {
      Sphere $inner { Radius [2] }
      Instance { Data [$inner] SurfaceShaderBind [$red] }
      Instance { Data [$ball] }
    }
    Instance { Data [$ball] SurfaceShaderBind [$plain] }
    Instance { Data [$quad] SurfaceShaderBind [$weird] Transform [2 0 0 0 0 2 0 0 0 0 2 0 0 0 0 1] }
    Instance { Data [$moving] SurfaceShaderBind [$glow] }
    Instance { Data [$lamp] }
    Instance { Data [$panel] Transform [1 0 0 0 0 1 0 0 0 0 1 0 0 0 -3 1] Transform [1 0 0 0 0 1 0 0 0 0 1 0 0 0 -4 1] }
    Instance { Data [$child] Transform [1 0 0 0 0 1 0 0 0 0 1 0 3 0 0 1] }
    Instance { Data [$nothing] }
    Instance { Data [$subd] }
  }
}
Scene $second { RenderSettings { Resolution [2 3] } }
"""


def _f(a):
    return np.asarray(a, np.float32)


def test_scene_settings_camera_and_world():
    assert psyfile.scene_count(SCENE) == 2
    p = psyfile.ParsedScene(SCENE)
    s = p.settings
    assert (s["width"], s["height"], s["spp"], s["seed"]) == (320, 240, 8, 7)
    assert s["dice_rate"] == np.float32(0.5)
    assert s["name"] == "shot 1"
    assert s["output_path"] == "/tmp/out dir/shot.png"
    d = p.desc()
    assert d.dice_rate == np.float32(0.5) and d.min_upoly_size == np.float32(1e-5)
    assert np.array_equal(_f(d.background[:]), _f([0.1, 0.2, 0.3]))
    c = d.camera
    assert c.n_fovs == 2 and c.n_transforms == 2 and c.n_focus_distances == 1 and c.n_aperture_radii == 1
    # Parser::parse_camera reads degrees: (3.1415926536f / 180.0f) * value
    assert c.fovs[0] == DEG * np.float32(45) and c.fovs[1] == DEG * np.float32(50)
    assert c.focus_distances[0] == np.float32(12.5) and c.aperture_radii[0] == np.float32(0.05)
    xf = np.array(c.transforms[:32], np.float32).reshape(2, 4, 4)
    assert xf[1, 3, 0] == np.float32(0.5) and xf[0, 3, 1] == -5.0
    s2 = psyfile.ParsedScene(SCENE, 1).settings
    assert (s2["width"], s2["height"], s2["spp"], s2["seed"], s2["name"]) == (2, 3, 1, 1, "second")
    with pytest.raises(gpu.PsyError):
        psyfile.ParsedScene(SCENE, 2)


def test_assembly_objects_instances_and_warnings():
    p = psyfile.ParsedScene(SCENE)
    d = p.desc()
    # shaders: the known types in order; Lambert without Color is 0.9 grey;
    # GTR takes the reference parser's defaults for the fields it lacks
    # (parser/parser.cpp:512-515: colour 0.9, roughness 0.1, tail 2, Fresnel 0.25)
    assert d.n_shaders == 4
    sh = [(d.shaders[i].kind, tuple(d.shaders[i].color)) for i in range(4)]
    assert sh[0] == (abi.PSY_SHADER_LAMBERT, tuple(_f([0.8, 0.1, 0.1])))
    assert sh[1] == (abi.PSY_SHADER_EMIT, (4.0, 3.0, 2.0))
    assert sh[2] == (abi.PSY_SHADER_LAMBERT, tuple(_f([0.9, 0.9, 0.9])))
    assert sh[3] == (abi.PSY_SHADER_GTR, tuple(_f([0.9, 0.9, 0.9])))
    g = d.shaders[3]
    assert (g.roughness, g.tail_shape, g.fresnel) == tuple(_f([0.2, 2.0, 0.25]))
    assert d.n_assemblies == 2
    A = d.assemblies[0]
    kinds = [A.objects[i].kind for i in range(A.n_objects)]
    assert kinds == [abi.PSY_OBJ_SPHERE, abi.PSY_OBJ_SPHERE, abi.PSY_OBJ_BILINEAR, abi.PSY_OBJ_BICUBIC,
                     abi.PSY_OBJ_SPHERE_LIGHT, abi.PSY_OBJ_RECTANGLE_LIGHT, abi.PSY_OBJ_SUBDIVISION]
    # a control mesh of one point and no faces: an empty surface
    assert (A.objects[6].n_vertices, A.objects[6].n_faces, A.objects[6].time_samples) == (1, 0, 1)

    def data(o, nf):
        return np.ctypeslib.as_array(o.data, (o.time_samples * nf,)).copy()

    # -?[0-9]+[.]?[0-9]*: "1e-5 .5 -.5" reads 1, -5, 5 and 5 (a '-' before
    # '.' is no part of a number; the fourth value is dropped), "3." reads 3
    assert np.array_equal(data(A.objects[0], 4), _f([1, -5, 5, 3]))
    # the first definition of a name wins; defaults (0,0,0) r 0.5 fill in
    assert np.array_equal(data(A.objects[1], 4), _f([9, 9, 9, 0.5]))
    assert A.objects[3].time_samples == 2 and data(A.objects[3], 48)[48 + 2] == 1.0
    # a single Radius and Color hold for both Location samples
    assert np.array_equal(data(A.objects[4], 7), _f([0, 0, 5, 0.5, 10, 9, 8, 0, 0, 6, 0.5, 10, 9, 8]))
    assert np.array_equal(data(A.objects[5], 5), _f([2, 1, 1, 1, 1]))
    inst = [(A.instances[i].type, A.instances[i].data_index, A.instances[i].transform_index,
             A.instances[i].transform_count, A.instances[i].shader_index) for i in range(A.n_instances)]
    assert inst == [(0, 0, 0, 0, 2), (0, 2, 0, 1, -1), (0, 3, 1, 0, 1), (0, 4, 1, 0, -1), (0, 5, 1, 2, -1),
                    (1, 0, 3, 1, -1), (0, 6, 4, 0, -1)]
    assert A.n_transforms == 4 and A.n_assemblies == 1 and A.assemblies[0] == 1
    B = d.assemblies[1]
    # $red found up the parent chain; $ball is not in $child's scope
    assert B.n_objects == 1 and B.n_instances == 1 and B.instances[0].shader_index == 0
    assert np.array_equal(data(B.objects[0], 4), _f([0, 0, 0, 2]))
    w = p.warnings
    for msg in ("unknown surface shader type 'OSL'", "$ball, which does not exist", "$nothing, which does not exist"):
        assert msg in w, msg
    assert "$subd" not in w


@pytest.mark.parametrize("bad,msg", [
    ("SphereLight $l2 { Location [0 0 1] Location [0 0 2] Radius [1] Radius [2] Radius [3] Color [1 1 1] }",
     "motion sample counts"),
    ("SphereLight $l3 { Radius [1] Color [1 1 1] }", "needs Location"),
    ("BilinearPatch $p { Vertices [0 0 0 1 0 0] }", "needs 12"),
])
def test_unsupported_scenes_fail(bad, msg):
    text = SCENE.replace("    Instance { Data [$subd] }", "    " + bad)
    with pytest.raises(gpu.PsyError, match=msg):
        psyfile.ParsedScene(text)


def _compare_assembly(pa, pb, da, db):
    """Recursive structural equality of two psy_assembly_desc trees."""
    A, B = da.assemblies[pa], db.assemblies[pb]
    assert (A.n_objects, A.n_instances, A.n_transforms, A.n_assemblies) == \
        (B.n_objects, B.n_instances, B.n_transforms, B.n_assemblies)
    for i in range(A.n_objects):
        oa, ob = A.objects[i], B.objects[i]
        assert (oa.kind, oa.time_samples) == (ob.kind, ob.time_samples)
        if oa.kind == abi.PSY_OBJ_SUBDIVISION:
            assert (oa.n_vertices, oa.n_faces) == (ob.n_vertices, ob.n_faces)
            nf = 3 * oa.n_vertices * oa.time_samples
            cnt = np.ctypeslib.as_array(oa.face_vert_counts, (oa.n_faces,))
            assert np.array_equal(cnt, np.ctypeslib.as_array(ob.face_vert_counts, (ob.n_faces,)))
            ni = int(cnt.sum())
            assert np.array_equal(np.ctypeslib.as_array(oa.face_vert_indices, (ni,)),
                                  np.ctypeslib.as_array(ob.face_vert_indices, (ni,)))
        else:
            nf = abi.FLOATS_PER_KIND[oa.kind] * oa.time_samples
        assert np.array_equal(np.ctypeslib.as_array(oa.data, (nf,)), np.ctypeslib.as_array(ob.data, (nf,)))
    for i in range(A.n_instances):
        ia, ib = A.instances[i], B.instances[i]
        assert (ia.type, ia.data_index, ia.transform_index, ia.transform_count, ia.shader_index) == \
            (ib.type, ib.data_index, ib.transform_index, ib.transform_count, ib.shader_index)
    if A.n_transforms:
        n = 16 * A.n_transforms
        assert np.array_equal(np.ctypeslib.as_array(A.transforms, (n,)), np.ctypeslib.as_array(B.transforms, (n,)))
    for j in range(A.n_assemblies):
        _compare_assembly(A.assemblies[j], B.assemblies[j], da, db)


def round_trip(sc):
    text = W.write_psy(sc)
    p = psyfile.ParsedScene(text)
    assert p.warnings == ""
    return p


@pytest.mark.parametrize("name", ["spheres", "heightfield", "bicubic_motion", "instanced", "instanced_lights",
                                  "chromatic", "many_lights", "subdivision"])
def test_round_trip_gives_the_same_description(name):
    from tests import _meshes
    sc = {"spheres": configs.sphere_scene, "heightfield": lambda: configs.heightfield_scene(grid=8),
          "bicubic_motion": lambda: configs.bicubic_sphere_scene(grid=4, n_outer=2, n_inner=2),
          "instanced": lambda: G._instanced_scene(True), "instanced_lights": R._instanced_lights_scene,
          "chromatic": R._chromatic_scene, "many_lights": lambda: configs.many_light_scene(n_lights=50),
          "subdivision": lambda: _meshes.subdiv_scene(True)}[name]()
    p = round_trip(sc)
    handle = sc.to_desc()  # keeps the buffers the description points into
    want = handle.desc
    got = p.desc()
    assert (got.n_shaders, got.dice_rate) == (want.n_shaders, want.dice_rate)
    for i in range(want.n_shaders):
        assert got.shaders[i].kind == want.shaders[i].kind
        assert tuple(got.shaders[i].color) == tuple(want.shaders[i].color)
    assert tuple(got.background) == tuple(want.background)
    gc, wc = got.camera, want.camera
    assert (gc.n_transforms, gc.n_fovs, gc.n_aperture_radii, gc.n_focus_distances) == \
        (wc.n_transforms, wc.n_fovs, wc.n_aperture_radii, wc.n_focus_distances)
    n = 16 * wc.n_transforms
    assert np.array_equal(np.ctypeslib.as_array(gc.transforms, (n,)), np.ctypeslib.as_array(wc.transforms, (n,)))
    for k in range(wc.n_fovs):  # through degrees and back: within an ulp or two
        assert gc.fovs[k] == pytest.approx(wc.fovs[k], rel=3e-7)
    for k in range(wc.n_aperture_radii):
        assert gc.aperture_radii[k] == wc.aperture_radii[k]
    for k in range(wc.n_focus_distances):
        assert gc.focus_distances[k] == wc.focus_distances[k]
    _compare_assembly(0, 0, got, want)


REF_EXAMPLE = "/root/reference/docs/scene_format_example.psy"


@pytest.mark.skipif(not os.path.exists(REF_EXAMPLE), reason="reference tree absent")
def test_reference_example_file():
    """The reference's own example scene, read as its Parser reads it: one
    sphere light instanced, an (empty) sub-assembly instanced, the GTR
    shader unused, the CatmullClarkSubdiv nodes not a type it knows."""
    p = psyfile.load(REF_EXAMPLE)
    s = p.settings
    assert (s["width"], s["height"], s["spp"], s["seed"], s["name"]) == (1280, 720, 16, 1, "yar_0001")
    assert s["output_path"].endswith("yar_0001.png")
    d = p.desc()
    assert tuple(d.background) == tuple(_f([0.8, 0.8, 0.8]))
    assert d.camera.n_transforms == 2 and d.camera.fovs[0] == DEG * np.float32(0.785398163)
    # $grey_diffuse, then $mirror: a GTR of roughness 0 (a perfect mirror)
    assert d.n_shaders == 2 and d.shaders[0].kind == abi.PSY_SHADER_LAMBERT
    m = d.shaders[1]
    assert m.kind == abi.PSY_SHADER_GTR and (m.roughness, m.tail_shape, m.fresnel) == tuple(_f([0.0, 2.0, 0.25]))
    A = d.assemblies[0]
    assert A.n_objects == 1 and A.objects[0].kind == abi.PSY_OBJ_SPHERE_LIGHT
    assert [A.instances[i].type for i in range(A.n_instances)] == [abi.PSY_INSTANCE_OBJECT, abi.PSY_INSTANCE_ASSEMBLY]
    assert d.n_assemblies == 2 and d.assemblies[1].n_instances == 0
    assert p.warnings.count("\n") == 4


def test_subdivision_surface_fields():
    """Parser::parse_subdivision_surface (parser/parser.cpp:416-468): whole
    xyz triples of every Vertices leaf, the first leaf fixing the points per
    time sample; the last FaceVertCounts / FaceVertIndices leaves win."""
    text = """Scene $s { Output { Path [] } RenderSettings { Resolution [4 4] SamplesPerPixel [1] Seed [1] }
      Camera { Fov [40] Transform [1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1] }
      World { BackgroundShader { Type [Color] Color [0 0 0] } }
      Assembly {
        SubdivisionSurface $m {
          Vertices [0 0 0  1 0 0  1 1 0  0 1 0  9]
          FaceVertCounts [3 3]
          FaceVertCounts [4]
          FaceVertIndices [0 1 2 3]
          Vertices [0 0 1  1 0 1  1 1 1  0 1 1  2 2 2]
        }
        Instance { Data [$m] }
      }
    }"""
    p = psyfile.ParsedScene(text)
    A = p.desc().assemblies[0]
    o = A.objects[0]
    assert (o.kind, o.n_vertices, o.time_samples, o.n_faces) == (abi.PSY_OBJ_SUBDIVISION, 4, 2, 1)
    v = np.ctypeslib.as_array(o.data, (24,))
    assert v[12 + 2] == 1.0 and v[-1] == 1.0   # the fifth point of the second leaf is cut
    assert list(np.ctypeslib.as_array(o.face_vert_counts, (1,))) == [4]
    assert list(np.ctypeslib.as_array(o.face_vert_indices, (4,))) == [0, 1, 2, 3]
    bad = text.replace("FaceVertIndices [0 1 2 3]", "FaceVertIndices [0 1 2]")
    with pytest.raises(gpu.PsyError, match="FaceVertIndices"):
        psyfile.ParsedScene(bad)
