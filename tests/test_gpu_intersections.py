"""psy_intersections, the rest of the reference's Intersection for each
closest hit (differential geometry in the object's space, offset, space
transform, backfacing, light pdf, and the closure the shader binds),
against the oracle's Intersection records of the same rays
(basics/intersection.hpp:22-49, tracer/tracer.cpp:241-250).

Bars: on rays where device and oracle hit the same instance at the same t
bit for bit, the closure kind and colour and the backfacing flag agree
exactly and every vector field within 1e-5 of its scale (the same
arithmetic from the same t, u and v); on every ray with the same instance
and t within 1e-5, within 1e-3."""
import numpy as np
import pytest
import torch

from psychopath_cpp_amd import abi, configs, gpu
from tests.test_gpu_parity import _instanced_scene
from tests.test_gpu_render import _instanced_lights_scene

pytestmark = pytest.mark.gpu

VEC_FIELDS = [("p", 0), ("n", 3), ("dpdu", 6), ("dpdv", 9), ("dndu", 12), ("dndv", 15), ("offset", 18)]


def say(*a):
    print(*a, flush=True)


def _compare(oracle, ctx, g, o, w, what):
    hits = gpu.trace_world(ctx, g, w)
    rec = gpu.intersections(ctx, g, w, hits).cpu().numpy()
    inter = o.trace(oracle.wrays_from_planes(w.cpu().numpy()), threads=8, chunk=8192)
    inst = hits.inst.cpu().numpy().view(np.uint32)
    t = hits.t.cpu().numpy()
    ids = g.instance_ids()
    hit = (inst != 0xFFFFFFFF) & (inter["hit"] != 0)
    same = hit.copy()
    same[hit] = ids[inst[hit]] == inter["id"][hit]
    close = same.copy()
    close[same] = np.abs(t[same] - inter["t"][same]) <= 1e-5 * np.abs(inter["t"][same])
    exact = close & (t.view(np.uint32) == inter["t"].view(np.uint32))
    say(f"{what}: {int(hit.sum())} hits, same instance {int(same.sum())}, t close {int(close.sum())}, "
        f"t exact {int(exact.sum())}")
    assert close.sum() >= 0.999 * hit.sum() and exact.sum() >= 0.9 * close.sum()
    # misses carry closure kind 0
    miss = inst == 0xFFFFFFFF
    assert (rec[39][miss].view(np.uint32) == 0).all()
    assert (rec[37][miss] == 9999.0).all() and (rec[43][miss].view(np.int32) == -1).all()
    assert (rec[21:37].T[miss] == np.eye(4, dtype=np.float32).ravel()).all()
    # Intersection's light_pdf default where no light set it
    assert np.array_equal(rec[37][close & (inter["light_pdf"] == 9999.0)],
                          inter["light_pdf"][close & (inter["light_pdf"] == 9999.0)])
    # the closure and the flags
    kind = rec[39].view(np.uint32)
    assert np.array_equal(kind[close], inter["closure"][close].astype(np.uint32))
    assert np.array_equal(rec[40:43].T[close], inter["closure_col"][close])
    assert np.array_equal(rec[38].view(np.uint32)[exact], inter["backfacing"][exact].astype(np.uint32))
    lights = close & (rec[43].view(np.int32) == -1) & (kind == abi.PSY_SHADER_EMIT) & (inter["light_pdf"] != 9999.0)
    if lights.any():
        rel = np.abs(rec[37][lights] - inter["light_pdf"][lights]) / inter["light_pdf"][lights]
        assert rel.max() <= 1e-5, rel.max()
    # the space transform (the instance's merged transform at the ray time)
    sp = rec[21:37].T
    assert np.abs(sp[close] - inter["space"][close]).max() <= 1e-6 * max(1.0, np.abs(inter["space"][close]).max())
    worst = {}
    # a light hit sets only p, n and the offset of the geometry (the
    # reference leaves the derivatives as an earlier candidate hit left them)
    light_hit = (rec[43].view(np.int32) == -1) & (kind == abi.PSY_SHADER_EMIT) & (rec[37] != 9999.0)
    for name, p0 in VEC_FIELDS:
        g_v = rec[p0:p0 + 3].T
        o_v = inter[name]
        scale = np.maximum(np.linalg.norm(o_v, axis=1), 1e-20)
        err = np.linalg.norm(g_v - o_v, axis=1) / scale
        nz = np.linalg.norm(o_v, axis=1) > 0
        if name.startswith("d"):
            nz &= ~light_hit
            assert (np.abs(g_v[light_hit]) == 0).all()
        e_exact = float(err[exact & nz].max()) if (exact & nz).any() else 0.0
        e_close = float(err[close & nz].max()) if (close & nz).any() else 0.0
        worst[name] = (e_exact, e_close)
        assert e_exact <= 1e-5, (name, e_exact)
        assert e_close <= 1e-3, (name, e_close)
    say(f"{what}: worst relative error (t exact, t close) {worst}")
    return hits


@pytest.mark.parametrize("name", ["instanced", "spheres", "instanced_lights"])
def test_gpu_intersection_records_match_oracle(oracle, name):
    sc = {"instanced": lambda: _instanced_scene(False), "spheres": configs.sphere_scene,
          "instanced_lights": _instanced_lights_scene}[name]()
    ctx = gpu.Context(0)
    g = None
    try:
        g = gpu.GpuScene(ctx, sc)
        o = oracle.OracleScene(sc)
        w, _ = gpu.camera_rays(ctx, g, 96, 64, 2, 3)
        hits = _compare(oracle, ctx, g, o, w, name + " camera")
        b, valid = gpu.bounce_rays(ctx, g, w, hits, 96, 2, 3, 9, 10)
        vb = b[:, valid.bool()].contiguous()
        _compare(oracle, ctx, g, o, vb, name + " bounce")
    finally:
        if g is not None:
            g.close()
        ctx.close()
