"""Differentiable rendering: psy_render / psy_adjoint as a torch autograd
function of the scene's grey albedos (one per surface shader) and emission
scales (one per light object).

    film = render_image(ctx, scene, albedo, emission, width, height, spp)
    loss = ((film - target) ** 2).sum()
    loss.backward()          # albedo.grad, emission.grad from psy_adjoint

The backward pass is the library's path-replay adjoint: it re-runs every
path from its sample index with the same seed, so the forward call stores
nothing beyond the parameters.  The parameters are written into the scene
(psy_scene_set_params) before both passes."""
import torch

from . import gpu


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(fctx, albedo, emission, ctx, scene, width, height, spp, seed, path_length):
        scene.set_params(albedo, emission)
        film, _, _ = gpu.render(ctx, scene, width, height, spp, seed, path_length)
        fctx.save_for_backward(albedo, emission)
        fctx.psy = (ctx, scene, spp, seed, path_length)
        return film

    @staticmethod
    def backward(fctx, d_film):
        albedo, emission = fctx.saved_tensors
        ctx, scene, spp, seed, path_length = fctx.psy
        scene.set_params(albedo, emission)
        ga, ge, _ = gpu.adjoint(ctx, scene, d_film.float().contiguous(), spp, seed, path_length)
        return (ga.to(albedo.device, albedo.dtype), ge.to(emission.device, emission.dtype),
                None, None, None, None, None, None, None)


def render_image(ctx, scene, albedo, emission, width, height, spp, seed=1, path_length=4):
    """(height, width, 3) CIE XYZ film as a differentiable function of the
    albedo (n_shaders) and emission-scale (n_light_objects) tensors."""
    return RenderFunction.apply(albedo, emission, ctx, scene, width, height, spp, seed, path_length)
