"""The benchmark configurations at their own scene sizes on the GPU, against
the oracle on samples of the same rays and on reduced-resolution renders of
the same scenes:

  C1  64 x 64 bilinear heightfield patches, one rectangle light;
  C2  4,096 bicubic patches on a sphere, instanced 16 times through two
      levels with 2-sample motion, camera depth of field, path length 6;
  C3  the sphere scene plus 10,000 sphere lights (light tree + MIS).

Trace bars as in test_gpu_parity (same hit or miss, same InstanceID, t
within 1e-5 relative on >= 99.99 % of a random sample of >= 4,096 rays);
render bars as in test_gpu_render (per 8 x 8 tile within 3 sigma of the
Monte Carlo error, samples mostly within 1e-3).  Only the image size is
reduced: the scenes are the configurations' own."""
import numpy as np
import pytest
import torch

from psychopath_cpp_amd import configs, gpu
from tests.test_gpu_render import check_image

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c2(oracle):
    say("building the C2 scene (4096 bicubic patches x 16 moving instances) on the GPU and the oracle")
    sc = configs.bicubic_sphere_scene()
    return sc, oracle.OracleScene(sc)


@pytest.fixture(scope="module")
def c3(oracle):
    say("building the C3 scene (10000 sphere lights) on the GPU and the oracle")
    sc = configs.many_light_scene(n_lights=10000)
    return sc, oracle.OracleScene(sc)


def _sample_parity(oracle, g, o, wplanes, hits, n_pick, seed, what):
    """Hits of a random sample of the stream against the oracle."""
    n = wplanes.shape[1]
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(n, min(n_pick, n), replace=False))
    w = oracle.wrays_from_planes(wplanes[:, torch.from_numpy(pick).cuda()].cpu().numpy())
    inter = o.trace(w, threads=16, chunk=4096)
    inst = hits.inst.cpu().numpy().view(np.uint32)[pick]
    t_g = hits.t.cpu().numpy()[pick]
    hit_o = inter["hit"] != 0
    hit_g = inst != 0xFFFFFFFF
    ids = g.instance_ids()
    ok = hit_o == hit_g
    both = hit_o & hit_g
    ok[both] &= ids[inst[both]] == inter["id"][both]
    ok[both] &= np.abs(t_g[both] - inter["t"][both]) <= 1e-5 * np.abs(inter["t"][both])
    say(f"{what}: {len(pick)} of {n} rays sampled, oracle hits {hit_o.mean():.3f}, agree {ok.mean():.6f}")
    return ok.mean(), hit_o.mean()


def _trace_config(oracle, sc, o, W, H, spp, seed, what, min_hit):
    ctx = gpu.Context(0)
    g = None
    try:
        g = gpu.GpuScene(ctx, sc)
        say(f"{what}: {g.info.n_leaf_instances} leaves, dice depth {g.info.dice_depth}, "
            f"{g.info.dice_cache_bytes / 2**30:.2f} GiB dice cache")
        w, _ = gpu.camera_rays(ctx, g, W, H, spp, seed)
        hits = gpu.trace_world(ctx, g, w)
        second, third = ctx.pass_counts()
        say(f"{what}: camera stream of {w.shape[1]} rays, second pass {second}, third {third}")
        agree, hit = _sample_parity(oracle, g, o, w, hits, 4096, seed, what + " camera")
        assert hit >= min_hit and agree >= 0.9999
        # first-bounce rays from those hits (Lambert sample dimensions 9, 10)
        b, valid = gpu.bounce_rays(ctx, g, w, hits, W, spp, seed, 9, 10)
        idx = torch.nonzero(valid, as_tuple=False).flatten()
        bw = b[:, idx].contiguous()
        bh = gpu.trace_world(ctx, g, bw)
        agree, _ = _sample_parity(oracle, g, o, bw, bh, 4096, seed + 1, what + " bounce")
        assert agree >= 0.9999
    finally:
        if g is not None:
            g.close()
        ctx.close()


def _render_config(oracle, sc, o, W, H, spp, path_length, what):
    ctx = gpu.Context(0)
    g = None
    try:
        g = gpu.GpuScene(ctx, sc)
        film, samples, rays = gpu.render(ctx, g, W, H, spp, 1, path_length, samples=True)
        torch.cuda.synchronize()
        os_, orays = o.render(W, H, spp, 1, threads=16, path_length=path_length)
        check_image(what, film.cpu().numpy(), samples.cpu().numpy(), rays, os_, orays)
    finally:
        if g is not None:
            g.close()
        ctx.close()


def test_gpu_c2_full_scene_trace(oracle, c2):
    sc, o = c2
    _trace_config(oracle, sc, o, 1920, 1080, 1, 1, "C2 1920x1080", 0.1)


def test_gpu_c2_full_scene_render(oracle, c2):
    sc, o = c2
    _render_config(oracle, sc, o, 480, 270, 4, 6, "C2 full scene at 480x270x4, path length 6")


def test_gpu_c3_full_scene_trace(oracle, c3):
    sc, o = c3
    _trace_config(oracle, sc, o, 1024, 1024, 1, 1, "C3 1024x1024", 0.5)


def test_gpu_c3_full_scene_render(oracle, c3):
    sc, o = c3
    _render_config(oracle, sc, o, 256, 256, 4, 4, "C3 10000 lights at 256x256x4")


def test_gpu_c1_full_field_render(oracle):
    sc = configs.heightfield_scene()
    _render_config(oracle, sc, oracle.OracleScene(sc), 256, 256, 4, 4, "C1 64x64 field at 256x256x4")
