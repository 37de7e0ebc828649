"""CPU runs of the kernels' own per-ray and per-path code (psy_trace.hpp and
psy_render_dev.hpp compiled for the host, tests/cpu_build) against the
oracle.  They check the traversal and integrator logic without a GPU, with
the bars of the GPU tests; the GPU tests then check the device execution
of the same code."""
import numpy as np
import pytest

from oracle import ora
from psychopath_cpp_amd import configs
from tests.cpu_build import hostbuild
from tests import test_gpu_parity as G
from tests import test_gpu_render as R


def compare(hb, inter):
    hit_h = hb["inst"] != 0xFFFFFFFF
    hit_o = inter["hit"] != 0
    both = hit_h & hit_o
    t_ok = np.abs(hb["t"][both] - inter["t"][both]) <= 1e-5 * np.abs(inter["t"][both])
    return float((hit_h == hit_o).mean()), float(t_ok.mean()) if both.any() else 1.0, int(hit_h.sum())


@pytest.mark.parametrize("motion", [False, True])
@pytest.mark.parametrize("dice_depth", [4, 12])
def test_cpu_build_instanced_hierarchy(motion, dice_depth):
    """Transformed bicubic / bilinear / sphere instances two levels deep; a
    shallow dice cache sends most rays through the second pass."""
    sc = G._instanced_scene(motion)
    o = ora.OracleScene(sc)
    w, _ = o.camera_rays(64, 48, 2, 1)
    hb = hostbuild.trace(sc, ora.planes_from_wrays(w), world=True, dice_depth=dice_depth)
    inter = o.trace(w, threads=8, chunk=8192)
    agree, t_ok, hits = compare(hb, inter)
    assert hits > 0.1 * len(w)
    assert agree >= 0.9999 and t_ok >= 0.9999
    if dice_depth == 4:
        assert hb["retraced"] > 0


def _spinning_scene():
    """Instances that turn fast through three transform samples (up to 90
    degrees a segment, with scale changes), nested two levels deep: the
    interpolated node bounds of a moving scene must still hold every leaf
    at every ray time."""
    from psychopath_cpp_amd.scene import Assembly, Bicubic, Camera, Scene, Sphere, SphereLight, look_at

    def spin(ang_deg, scl, pos):
        a = np.radians(ang_deg)
        c, s = np.cos(a), np.sin(a)
        m = np.array([[c * scl, s * scl, 0, 0], [-s * scl, c * scl, 0, 0], [0, 0, scl, 0], [pos[0], pos[1], pos[2], 1]])
        return np.linalg.inv(m).astype(np.float32)

    rng = np.random.default_rng(21)
    inner = Assembly()
    for k in range(4):
        base = rng.normal(size=3) * 0.6 + np.array([1.0, 0.0, 0.0])
        pts = [base + np.array([i * 0.4, j * 0.4, rng.normal() * 0.2]) for j in range(4) for i in range(4)]
        inner.add_object(f"bc{k}", Bicubic(np.array(pts)))
        inner.create_object_instance(f"bc{k}", [], None)
    inner.add_object("ball", Sphere([1.5, 0.5, 0.3], 0.4))
    inner.create_object_instance("ball", [], None)
    mid = Assembly()
    mid.add_assembly("inner", inner)
    for k in range(3):
        pos = rng.normal(size=3) * 1.2
        mid.create_assembly_instance("inner", [spin(0 + 40 * k, 1.0, pos), spin(70 + 40 * k, 1.2, pos + 0.3),
                                               spin(150 + 40 * k, 0.9, pos)], None)
    root = Assembly()
    root.add_assembly("mid", mid)
    for k in range(3):
        pos = np.array([rng.uniform(-4, 4), rng.uniform(-4, 4), rng.uniform(-1, 1)])
        root.create_assembly_instance("mid", [spin(-30 * k, 1.0, pos), spin(-30 * k + 90, 1.0, pos + 0.5),
                                              spin(-30 * k + 120, 1.1, pos - 0.5)], "grey")
    root.add_object("lamp", SphereLight([0, 0, 9], 1.0, (5, 5, 5)))
    root.create_object_instance("lamp", [], None)
    from psychopath_cpp_amd import abi
    sc = Scene(Camera([look_at((0, -14, 7), (0, 0, 0))], [np.radians(50)]), root, dice_rate=0.5)
    sc.add_surface_shader("grey", abi.PSY_SHADER_LAMBERT, (0.6, 0.6, 0.6))
    return sc


def test_cpu_build_fast_motion_node_bounds():
    """Checked against the oracle's every-instance trace: the reference's own
    instance culling interpolates per-time-sample boxes (bbox_lerp in its
    BVH), which do not hold turning geometry between samples, so its BVH
    trace misses surfaces here (the ray then misses or finds one farther
    away) that the every-instance trace and the device's padded bounds
    find."""
    sc = _spinning_scene()
    o = ora.OracleScene(sc)
    w, _ = o.camera_rays(96, 72, 2, 5)
    hb = hostbuild.trace(sc, ora.planes_from_wrays(w), world=True, dice_depth=12)
    exact = o.trace(w, threads=8, chunk=8192, brute=True)
    agree, t_ok, hits = compare(hb, exact)
    assert hits > 0.1 * len(w)
    assert agree >= 0.9999 and t_ok >= 0.9999
    culled = o.trace(w, threads=8, chunk=8192)
    ch = culled["hit"] != 0
    assert not (ch & (exact["hit"] == 0)).any()
    assert (culled["t"][ch] >= exact["t"][ch]).all()


def test_cpu_build_heightfield_primary_and_object_rays():
    sc = configs.heightfield_scene(grid=16)
    o = ora.OracleScene(sc)
    w, _ = o.camera_rays(64, 64, 1, 3)
    inter = o.trace(w, threads=8, chunk=8192)
    planes = ora.planes_from_wrays(w)
    hb = hostbuild.trace(sc, planes, world=True, dice_depth=8)
    agree, t_ok, hits = compare(hb, inter)
    assert hits == len(w) and agree == 1.0 and t_ok >= 0.9999


def test_cpu_build_stack_overflow_retrace():
    sc = G._nested_spheres(n=1024)
    o = ora.OracleScene(sc)
    w, _ = o.camera_rays(24, 24, 1, 4)
    hb = hostbuild.trace(sc, ora.planes_from_wrays(w), world=True)
    inter = o.trace(w, threads=8, chunk=8192)
    agree, t_ok, hits = compare(hb, inter)
    assert hb["retraced"] > 0
    assert hits > 0 and agree >= 0.9999 and t_ok >= 0.9999


@pytest.mark.parametrize("name", ["spheres", "heightfield", "instanced_lights"])
def test_cpu_build_render_matches_oracle(name):
    sc = {"spheres": configs.sphere_scene, "heightfield": lambda: configs.heightfield_scene(grid=16),
          "instanced_lights": R._instanced_lights_scene}[name]()
    W, H, spp = 24, 24, 4
    hs, hrays = hostbuild.render(sc, W, H, spp, seed=1)
    os_, orays = ora.OracleScene(sc).render(W, H, spp, 1, threads=8)
    worst, close = R.tile_check(hs, os_)
    print(name, "rays", hrays, orays, "close", close, "worst tile z", worst)
    assert np.isfinite(hs).all() and hs.mean() > 0.0
    assert close >= 0.95 and worst <= 3.0
    assert abs(hrays - orays) <= 0.01 * orays


def test_cpu_build_chromatic_render_matches_oracle():
    """Coloured albedos, lights, emitters, background and the no-shader
    magenta emitter through the reference's spectrum grid: the integrator's
    per-wavelength colour code against the oracle's."""
    G = ora.ref_spectrum_grid()
    if G is None:
        pytest.skip("oracle/_ref not built: no spectrum grid")
    sc = R._chromatic_scene()
    W, H, spp = 24, 16, 4
    with pytest.raises(RuntimeError):
        ora.OracleScene(sc).render(W, H, spp, 1, threads=8)  # no grid: chromatic refused
    ora.spectrum_set_grid(G)
    hostbuild.set_spectrum_grid(G)
    try:
        hs, hrays = hostbuild.render(sc, W, H, spp, seed=1)
        os_, orays = ora.OracleScene(sc).render(W, H, spp, 1, threads=8)
    finally:
        ora.spectrum_set_grid(None)
        hostbuild.set_spectrum_grid(None)
    worst, close = R.tile_check(hs, os_)
    print("chromatic rays", hrays, orays, "close", close, "worst tile z", worst)
    assert np.isfinite(hs).all() and hs.mean() > 0.0
    assert close >= 0.95 and worst <= 3.0
    assert abs(hrays - orays) <= 0.01 * orays
    m = hs.reshape(-1, 3).mean(axis=0)
    assert m.max() > 1.2 * m.min()


def test_cpu_build_adjoint_matches_finite_differences():
    """The adjoint pass (psy_render_dev.hpp's path replay, compiled for the
    host) against central finite differences of the oracle's renderer and of
    the same code's forward render, every parameter of a small scene with
    12 sphere lights."""
    from tests import _adjoint as J
    sc = configs.adjoint_scene(n_spheres=16, n_lights=12)
    n_sh, n_l = hostbuild.param_counts(sc.to_desc())
    W, H, S = 32, 24, 4
    alb0 = np.array([c[1][0] for c in sc.shaders], np.float32)
    alb, esc = J.perturbed(alb0, n_l, seed=0)
    o = ora.OracleScene(sc)

    def film_oracle(a, e):
        o.set_params(a, e)
        return o.render(W, H, S, 1, threads=8)[0].astype(np.float64).mean(axis=2)

    def film_hb(a, e):
        return hostbuild.render_params(sc, W, H, S, albedo=a, escale=e)[0].astype(np.float64).mean(axis=2)

    target = film_oracle(alb0, np.ones(n_l, np.float32))
    grad, _ = hostbuild.render_params(sc, W, H, S, albedo=alb, escale=esc, d_film=J.d_loss(film_hb(alb, esc), target))
    params = list(range(n_sh + n_l))
    J.check(grad, J.central_fd(film_oracle, alb, esc, target, params), "adjoint vs oracle FD")
    # against its own forward code the agreement is far tighter than the bar
    fd_hb = J.central_fd(film_hb, alb, esc, target, params)
    sig = np.abs(fd_hb) >= 1e-3 * np.abs(fd_hb).max()
    assert (np.abs(grad - fd_hb)[sig] / np.abs(fd_hb)[sig]).max() <= 1e-3


def test_cpu_build_zero_x_direction_rays_on_split_plane():
    """Rays whose direction has x = +0 or -0, from an origin at x = 0: the
    ground patch's first split puts a half's low x plane exactly at the
    origin, where the box test's far distance is NaN.  The kernels skip such
    subtrees (no box below them can take a hit) and the oracle walks them in
    full; hits must agree exactly."""
    sc = configs.sphere_scene()
    o = ora.OracleScene(sc)
    w, _ = o.camera_rays(64, 64, 1, 3)
    sel = w[np.arange(0, len(w), len(w) // 32)][:32].copy()
    sel["d"][:, 0] = np.where(np.arange(len(sel)) % 2 == 0, 0.0, -0.0).astype(np.float32)
    assert (sel["o"][:, 0] == 0.0).all()
    inter = o.trace(sel, threads=8, chunk=64)
    hb = hostbuild.trace(sc, ora.planes_from_wrays(sel), world=True, dice_depth=8)
    agree, t_ok, hits = compare(hb, inter)
    assert hits > 0 and agree == 1.0 and t_ok == 1.0


def test_cpu_build_rejects_oversized_dice_cache():
    """Scene set-up refuses a dice depth above 24, and a cache of more than
    2^32 nodes (the main pass's node indices are 32-bit): 256 patches at
    24 levels hold 2^33."""
    sc = configs.heightfield_scene(grid=16)
    w, _ = ora.OracleScene(sc).camera_rays(4, 4, 1, 1)
    planes = ora.planes_from_wrays(w)
    for depth in (25, 24):
        with pytest.raises(RuntimeError):
            hostbuild.trace(sc, planes, world=True, dice_depth=depth)
    hb = hostbuild.trace(sc, planes, world=True, dice_depth=8)
    assert (hb["inst"] != 0xFFFFFFFF).all()


def _geometric_row_scene(n=1000):
    """Spheres along x at geometrically growing spacing: a binned SAH build
    peels them off one at a time, so an unbounded build would nest as deep
    as there are spheres."""
    from psychopath_cpp_amd.scene import Assembly, Camera, Scene, Sphere, look_at, translate
    root = Assembly()
    root.add_object("s", Sphere((0.0, 0.0, 0.0), 0.4))
    x = 0.0
    for k in range(n):
        root.create_object_instance("s", [translate(x, 0.0, 0.0)], "grey")
        x += 1.0 * (1.07 ** k)
    cam = Camera([look_at((-5.0, 3.0, 1.0), (40.0, 0.0, 0.0))], [0.9])
    sc = Scene(cam, root)
    sc.add_surface_shader("grey", 1, (0.5, 0.5, 0.5))
    return sc


def test_cpu_build_skewed_layout_bvh_depth():
    """A skewed instance layout still builds a BVH whose traversal stack
    fits the trace kernels (at most 64 entries), and traces like the
    oracle's exhaustive instance loop."""
    sc = _geometric_row_scene()
    rc, need = hostbuild.scene_stack_need(sc)
    assert rc == 0 and 1 <= need <= 64, (rc, need)
    o = ora.OracleScene(sc)
    w, _ = o.camera_rays(96, 32, 1, 1)
    hb = hostbuild.trace(sc, ora.planes_from_wrays(w), world=True, dice_depth=0)
    inter = o.trace(w, threads=8, chunk=4096, brute=True)
    agree, t_ok, hits = compare(hb, inter)
    assert hits > 50
    assert agree >= 0.9999 and t_ok >= 0.9999


@pytest.mark.parametrize("motion", [False, True])
def test_cpu_build_subdivision_surfaces(motion):
    """Catmull-Clark surfaces (a box instanced three times, one moving, a
    prism in a sub-assembly, a fan of triangles, an L-shaped sheet) against
    the oracle's SubdivisionSurface::intersect_rays over its patch BVH."""
    from tests import _meshes as Msh
    sc = Msh.subdiv_scene(motion)
    o = ora.OracleScene(sc)
    w, _ = o.camera_rays(96, 64, 2, 1)
    hb = hostbuild.trace(sc, ora.planes_from_wrays(w), world=True, dice_depth=10)
    inter = o.trace(w, threads=8, chunk=8192)
    agree, t_ok, hits = compare(hb, inter)
    assert hits > 0.3 * len(w)
    assert agree >= 0.9999 and t_ok >= 0.9999


def test_cpu_build_subdivision_render_matches_oracle():
    from tests import _meshes as Msh
    sc = Msh.subdiv_scene(False)
    W, H, spp = 32, 24, 4
    hs, hrays = hostbuild.render(sc, W, H, spp, seed=1)
    os_, orays = ora.OracleScene(sc).render(W, H, spp, 1, threads=8)
    worst, close = R.tile_check(hs, os_)
    assert np.isfinite(hs).all() and hs.mean() > 0.0
    assert close >= 0.95 and worst <= 3.0
    assert abs(hrays - orays) <= 0.01 * orays
