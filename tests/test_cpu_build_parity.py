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
