"""Pins of the oracle's adaptive-sampling film (film/film.hpp:135-165 in the
loop of integrator/path_trace_integrator.cpp:296-376) and of the bucket
size (:28-34), from hand computation and invariants.  CPU only."""
import math

import numpy as np
import pytest

f32 = np.float32


def _hcol(x):
    return f32(max(float(x), 0.0)) ** f32(1.0 / 2.2)


def test_bucket_size_formula():
    """sqrt(65536 / spp) truncated, at most sqrt(W H / (4 threads)), at least 1."""
    from psychopath_cpp_amd import gpu
    for W, H, spp, thr, want in ((1920, 1080, 16, 16, 64), (1920, 1080, 64, 16, 32), (64, 64, 1, 64, 4),
                                 (10, 10, 100000, 1, 1), (1024, 1024, 4, 1, 128), (7, 5, 1, 8, 1)):
        assert gpu.bucket_size(W, H, spp, thr) == want, (W, H, spp, thr)
        assert want == max(1, min(int(math.sqrt((W * H) / (thr * 4.0))), int(math.sqrt(65536.0 / spp))))


def test_constant_samples_stop_after_one_pass(oracle):
    """Equal samples never move the running gamma mean, so the variance
    estimate is 0 after the first pass of >= 2 samples and every bucket
    stops there; the film is the sample value."""
    s = np.full((12, 20, 4, 3), 0.37, np.float32)
    film, var, passes = oracle.film_adaptive(s, 64, 1e-9, 8)
    assert (passes == 1).all() and np.all(var == 0.0) and np.allclose(film, 0.37, rtol=1e-6)


def test_two_sample_variance_by_hand(oracle):
    """Samples 0 then 1 in one pixel: after the second, the running sum of
    mean changes is |hcol(0) - hcol(1/2)| * 1, so the estimate is that
    over (2 - 1), over sqrt(2)."""
    s = np.zeros((1, 1, 2, 3), np.float32)
    s[0, 0, 1, :] = 1.0
    film, var, passes = oracle.film_adaptive(s, 2, -1.0, 1)
    want = (abs(0.0 - float(_hcol(0.5))) / 1.0) / math.sqrt(2.0)
    assert var[0, 0, 0] == pytest.approx(want, rel=1e-6) and film[0, 0, 0] == pytest.approx(0.5)
    assert passes[0] == 1


def test_passes_run_to_spp_max_without_a_variance_bound(oracle):
    """variance_max -1 (the reference's default) never stops a bucket, so it
    takes ceil(spp_max / spp) passes of the same samples; the film mean
    stays the one-pass mean and the estimate falls as passes add samples."""
    rng = np.random.default_rng(3)
    s = rng.uniform(0, 1, size=(16, 16, 4, 3)).astype(np.float32)
    f1, v1, p1 = oracle.film_adaptive(s, 4, -1.0, 8)
    f3, v3, p3 = oracle.film_adaptive(s, 10, -1.0, 8)
    assert (p1 == 1).all() and (p3 == 3).all()
    assert np.allclose(f3, f1, rtol=1e-5) and np.allclose(f1, s.mean(axis=2), rtol=1e-5)
    assert (v3[..., 0] < v1[..., 0]).mean() > 0.9


def test_variance_bound_stops_buckets_independently(oracle):
    """A bucket of equal samples stops after one pass while a noisy bucket
    goes on to spp_max; non-finite samples are skipped (a pixel with none
    finite keeps count 0: film 0, estimate +inf, and its bucket runs every
    pass)."""
    rng = np.random.default_rng(5)
    s = np.full((8, 16, 4, 3), 0.2, np.float32)
    s[:, 8:] = rng.uniform(0, 4, size=(8, 8, 4, 3)).astype(np.float32)
    s[0, 0, 1, 2] = np.nan
    film, var, passes = oracle.film_adaptive(s, 16, 1e-4, 8)
    assert passes[0] == 1 and passes[1] == 4
    assert film[0, 0, 0] == pytest.approx(0.2) and var[0, 0, 0] == 0.0
    s[3, 3, :, 0] = np.inf
    film, var, passes = oracle.film_adaptive(s, 16, 1e-4, 8)
    assert film[3, 3, 1] == 0.0 and np.isinf(var[3, 3, 0]) and passes[0] == 4
