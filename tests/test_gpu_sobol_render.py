"""Renders that sample every path dimension from Sobol::sample
(PSY_RENDER_SOBOL) against the oracle's renderer in the same mode, per
tile as in test_gpu_render, with the reference's own Sobol table where
oracle/_ref was built (else matrices built from the definition); the mode
is refused without matrices and changes the samples."""
import numpy as np
import pytest
import torch

from psychopath_cpp_amd import configs, gpu
from tests.test_gpu_render import _instanced_lights_scene, check_image

pytestmark = pytest.mark.gpu


def _matrices(oracle):
    M = oracle.ref_sobol_matrices()
    return M if M is not None else oracle.make_sobol_matrices(256, seed=5)


@pytest.mark.parametrize("name", ["spheres", "instanced_lights"])
def test_gpu_sobol_render_matches_oracle(oracle, name):
    sc = {"spheres": configs.sphere_scene, "instanced_lights": _instanced_lights_scene}[name]()
    ctx = gpu.Context(0)
    g = None
    try:
        g = gpu.GpuScene(ctx, sc)
        with pytest.raises(gpu.PsyError, match="psy_sobol_set_matrices"):
            gpu.render(ctx, g, 8, 8, 1, sobol=True)
        M = _matrices(oracle)
        gpu.sobol_set_matrices(ctx, M)
        W, H, spp = 64, 48, 8
        film, samples, rays = gpu.render(ctx, g, W, H, spp, 1, samples=True, sobol=True)
        _, s_halton, _ = gpu.render(ctx, g, W, H, spp, 1, film=False, samples=True)
        torch.cuda.synchronize()
        os_, orays = oracle.OracleScene(sc).render(W, H, spp, 1, threads=8, sobol_matrices=M)
        check_image(f"{name} Sobol {W}x{H}x{spp}", film.cpu().numpy(), samples.cpu().numpy(), rays, os_, orays)
        assert not torch.equal(samples, s_halton)
    finally:
        if g is not None:
            g.close()
        ctx.close()
