"""Catmull-Clark subdivision surfaces on the GPU (SubdivisionSurface,
object/subdivision_surface.cpp): the scene's surfaces become the bicubic
patches of psy_subdivision_patches (bit-equal to the oracle's extraction,
tests/test_subdiv.py) under their instance's id, traced by the device's
patch path.  Bars as for the other scenes: same hit/miss, same InstanceID,
t within 1e-5 relative on at least 99.99% of rays; renders agree per tile."""
import numpy as np
import pytest
import torch

from psychopath_cpp_amd import gpu
from tests import _meshes as M
from tests._compare import compare_hits
from tests.test_gpu_render import render_both, tile_check

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("motion", [False, True])
@pytest.mark.parametrize("dice_depth", [-1, 4])
def test_gpu_subdivision_trace_matches_oracle(gpu_ctx, oracle, motion, dice_depth):
    sc = M.subdiv_scene(motion)
    g = gpu.GpuScene(gpu_ctx, sc, dice_depth=dice_depth)
    o = oracle.OracleScene(sc)
    W, H, spp, seed = 192, 128, 2, 1
    say("subdivision scene, motion", motion, "dice_depth", dice_depth)
    w, _ = gpu.camera_rays(gpu_ctx, g, W, H, spp, seed)
    hits = gpu.trace_world(gpu_ctx, g, w)
    inter = o.trace(oracle.wrays_from_planes(w.cpu().numpy()), threads=8, chunk=8192)
    st = compare_hits(hits, g.instance_ids(), inter)
    say("primary", st)
    assert st["hits"] > 0.3 * st["n"]
    assert st["match"] >= 0.9999 and st["t_rel_max"] <= 1e-5
    b, valid = gpu.bounce_rays(gpu_ctx, g, w, hits, W, spp, seed, 9, 10)
    vb = b[:, valid.bool()].contiguous()
    hb = gpu.trace_world(gpu_ctx, g, vb)
    ib = oracle.OracleScene(sc).trace(oracle.wrays_from_planes(vb.cpu().numpy()), threads=8, chunk=8192)
    st = compare_hits(hb, g.instance_ids(), ib)
    say("bounce", st)
    assert st["match"] >= 0.9999 and st["t_rel_max"] <= 1e-5


def test_gpu_subdivision_leaves_share_their_instance_id(gpu_ctx):
    sc = M.subdiv_scene(False)
    g = gpu.GpuScene(gpu_ctx, sc)
    ids = g.instance_ids()
    counts = {k: len(gpu.subdivision_patches(*m)) for k, m in
              (("box", M.cube(1.0)), ("prism", M.pentaprism(0.8)), ("fan", M.fan(7, ripple=0.3)),
               ("ground", M.lshape(6, ripple=0.2)))}
    # 3 boxes, 2 prisms, one fan, one ground, and the two lights
    want = 3 * counts["box"] + 2 * counts["prism"] + counts["fan"] + counts["ground"] + 2
    assert len(ids) == want
    uniq, n = np.unique(ids, return_counts=True)
    assert len(uniq) == 9
    assert sorted(n.tolist()) == sorted([counts["box"]] * 3 + [counts["prism"]] * 2 +
                                        [counts["fan"], counts["ground"], 1, 1])


def test_gpu_subdivision_render_matches_oracle(gpu_ctx, oracle):
    sc = M.subdiv_scene(False)
    W, H, spp = 64, 48, 4
    say("subdivision render", W, H, spp)
    g, film, gs, rays, os_, orays = render_both(gpu_ctx, oracle, sc, W, H, spp)
    worst, close = tile_check(gs, os_)
    say("rays", rays, orays, "close", close, "worst tile z", worst)
    assert np.isfinite(gs).all() and gs.mean() > 0.0
    assert close >= 0.95 and worst <= 3.0
    assert abs(rays - orays) <= 0.01 * orays
    torch.cuda.synchronize()


def test_gpu_parsed_subdivision_scene_matches_oracle(gpu_ctx, oracle):
    """The scene written as .psy text (SubdivisionSurface nodes with
    Vertices, FaceVertCounts, FaceVertIndices) and read back by the
    parser renders as the oracle renders the parsed description."""
    from psychopath_cpp_amd import psyfile
    from tests import _psywriter
    sc = M.subdiv_scene(False)
    p = psyfile.ParsedScene(_psywriter.write_psy(sc))
    assert p.warnings == ""
    W, H, spp = 48, 32, 4
    g = gpu.GpuScene(gpu_ctx, p)
    o = oracle.OracleScene(p)
    _, samples, rays = gpu.render(gpu_ctx, g, W, H, spp, 1, samples=True)
    torch.cuda.synchronize()
    gs = samples.cpu().numpy()
    os_, orays = o.render(W, H, spp, 1, threads=8)
    worst, close = tile_check(gs, os_)
    say("parsed subdivision render: rays", rays, orays, "close", close, "worst tile z", worst)
    assert gs.mean() > 0.0 and close >= 0.95 and worst <= 3.0
    assert abs(rays - orays) <= 0.01 * orays
