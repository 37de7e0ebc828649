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


def test_cpu_build_adjoint_matches_finite_differences():
    """The adjoint pass (psy_render_dev.hpp's path replay, compiled for the
    host) against central finite differences of the oracle's renderer and of
    the same code's forward render, every parameter of a small scene with
    12 sphere lights."""
    from tests import _adjoint as J
    sc = configs.adjoint_scene(n_spheres=16, n_lights=12)
    n_sh, n_l = hostbuild.param_counts(sc.to_desc())
    W, H, S = 32, 24, 4
    alb0 = np.array([c[0] for _, c in sc.shaders], np.float32)
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
