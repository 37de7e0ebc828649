"""psy_trace_rays_static, the trace entry point that allocates nothing, against
psy_trace_rays through the C ABI: bit-identical hits on two calls in a row,
the second with a smaller input (the fixed scratch keeps the first call's
queue contents, which the second must not read), and a host-side refusal of
a stream above PSY_STATIC_MAX_RAYS.

The rays are camera rays over the sphere scene, whose ground patch dices
below its cached levels for a large share of them: both calls take the
second pass, which is where the scratch (its queues and counters) is used.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from psychopath_cpp_amd import abi, configs, gpu

pytestmark = pytest.mark.gpu


def _equal(a, b):
    return (torch.equal(a.inst, b.inst) and torch.equal(a.t, b.t) and torch.equal(a.u, b.u)
            and torch.equal(a.v, b.v) and torch.equal(a.aux, b.aux))


def test_gpu_trace_rays_static_matches_trace_rays():
    ctx = gpu.Context(0)
    try:
        g = gpu.GpuScene(ctx, configs.sphere_scene())
        for k, (w, h, spp, seed) in enumerate(((1024, 1024, 2, 3), (640, 360, 1, 8))):
            wr, _ = gpu.camera_rays(ctx, g, w, h, spp, seed)
            rays = gpu.world_to_rays(ctx, wr)
            n = rays.shape[1]
            assert n <= abi.PSY_STATIC_MAX_RAYS
            want = gpu.trace_rays(ctx, g, rays)
            second, _ = ctx.pass_counts()
            got = gpu.trace_rays_static(ctx, g, rays)
            torch.cuda.synchronize()
            hit = float((want.inst != -1).float().mean().item())
            print(f"call {k}: {n} rays, hit fraction {hit:.3f}, second-pass rays {second}", flush=True)
            assert second > 1000
            assert _equal(got, want), k
        # above the fixed scratch's size: refused on the host, nothing launched
        big = torch.empty((abi.PSY_RAY_PLANES, 1), dtype=torch.float32, device="cuda")
        hs = gpu.HitBuffers(1, big.device).struct()
        rc = gpu.lib().psy_trace_rays_static(ctx.handle, g.handle, C.c_void_p(big.data_ptr()),
                                             abi.PSY_STATIC_MAX_RAYS + 1, C.byref(hs), 0, None)
        assert rc == abi.PSY_ERR_INVALID
        assert b"PSY_STATIC_MAX_RAYS" in gpu.lib().psy_last_error()
        g.close()
    finally:
        ctx.close()
