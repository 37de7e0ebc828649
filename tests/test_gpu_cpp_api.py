"""The C++ host interface on the GPU: Tracer::trace and
PathTraceIntegrator::integrate of include/psychopath_gpu.hpp (driven by
tests/cpp/test_cpp_api, built with g++) give the same hits and film, bit
for bit, as the C ABI called directly on the same scene."""
import os
import subprocess
import tempfile

import numpy as np
import pytest
import torch

from psychopath_cpp_amd import gpu
from tests.test_cpp_api import cpp_binary, cpp_equivalent_scene

pytestmark = pytest.mark.gpu


def test_gpu_cpp_parser_renderer(gpu_ctx, tmp_path):
    """psy::Parser::parse_next_frame + Renderer::render on a .psy file give
    the same film, bit for bit, as the scene built through the scene API and
    rendered through the C ABI."""
    from tests import _psywriter
    sc = cpp_equivalent_scene()
    path = tmp_path / "frame.psy"
    path.write_text(_psywriter.write_psy(sc, 24, 16, 3, 5, "frame"))
    out = tmp_path / "film.bin"
    r = subprocess.run([cpp_binary(), "psyrender", str(path), str(out)], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = out.read_bytes()
    W, H, spp = np.frombuffer(raw[:12], np.int32)
    film_cpp = np.frombuffer(raw[12:12 + 4 * W * H * 3], np.float32).reshape(H, W, 3)
    g = gpu.GpuScene(gpu_ctx, sc)
    film, _, _ = gpu.render(gpu_ctx, g, int(W), int(H), int(spp), 5)
    assert np.array_equal(film.cpu().numpy().view(np.uint32), film_cpp.view(np.uint32))


def test_gpu_cpp_tracer_and_integrator(gpu_ctx):
    binary = cpp_binary()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "out.bin")
        r = subprocess.run([binary, "gpu", path], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                           timeout=120)
        print(r.stdout, r.stderr, flush=True)
        assert r.returncode == 0
        raw = open(path, "rb").read()
    n, W, H = np.frombuffer(raw[:12], np.int32)
    off = 12
    rays = np.frombuffer(raw[off:off + 80 * n], np.float32).reshape(n, 20)
    off += 80 * n
    rec = np.frombuffer(raw[off:off + 24 * n], np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("id", "<u8"),
                                                         ("inst", "<u4")]))
    off += 24 * n
    film_cpp = np.frombuffer(raw[off:off + 4 * W * H * 3], np.float32).reshape(H, W, 3)
    off += 4 * W * H * 3
    rays_shot = int(np.frombuffer(raw[off:off + 8], np.int64)[0])
    off += 8
    isect_cpp = np.frombuffer(raw[off:off + 4 * 44 * n], np.float32).reshape(n, 44)
    sc = cpp_equivalent_scene()
    g = gpu.GpuScene(gpu_ctx, sc)
    planes = np.ascontiguousarray(rays.T)
    planes[19] = np.array([1], np.uint32).view(np.float32)[0]  # WorldRay::CAMERA
    hits = gpu.trace_world(gpu_ctx, g, torch.from_numpy(planes).cuda())
    torch.cuda.synchronize()
    inst = hits.inst.cpu().numpy().view(np.uint32)
    assert (inst != 0xFFFFFFFF).sum() > n // 2
    assert np.array_equal(rec["inst"], inst)
    assert np.array_equal(rec["t"].view(np.uint32), hits.t.cpu().numpy().view(np.uint32))
    assert np.array_equal(rec["u"].view(np.uint32), hits.u.cpu().numpy().view(np.uint32))
    assert np.array_equal(rec["v"].view(np.uint32), hits.v.cpu().numpy().view(np.uint32))
    ids = g.instance_ids()
    hit = inst != 0xFFFFFFFF
    assert np.array_equal(rec["id"][hit], ids[inst[hit]])
    # the Intersection's geometry, flags and re-materialised closure are the
    # C ABI's records of the same hits
    isect = gpu.intersections(gpu_ctx, g, torch.from_numpy(planes).cuda(), hits).cpu().numpy().T
    assert np.array_equal(isect_cpp[hit].view(np.uint32), isect[hit].view(np.uint32))
    kinds = isect_cpp[hit][:, 39].view(np.uint32)
    assert set(np.unique(kinds)) <= {1, 2} and (kinds == 1).any()
    film, _, rays_py = gpu.render(gpu_ctx, g, int(W), int(H), 4, 1)
    assert rays_shot == rays_py
    assert np.array_equal(film.cpu().numpy().view(np.uint32), film_cpp.view(np.uint32))
