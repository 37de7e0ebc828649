"""Shared pieces of the adjoint tests: perturbed parameters, the L2 image
loss and its finite differences.  Bar (per parameter): central finite
differences (h = 1e-3, common random numbers: same seed and sample streams
for both sides) and the adjoint agree to 1e-2 relative; parameters whose
finite difference is below 1e-3 of the largest one are held to an absolute
bar of 1e-4 of that largest one instead.  The loss difference is taken as
sum((f1 - f2) * (f1 + f2 - 2 t)) in float64, so that pixels a parameter
does not reach (bit-identical in both renders) add nothing."""
import numpy as np

H = 1e-3
REL = 1e-2


def perturbed(albedo0, n_light, seed=0):
    rng = np.random.default_rng(seed)
    alb = (np.asarray(albedo0, np.float64) * rng.uniform(0.5, 1.5, len(albedo0))).astype(np.float32)
    esc = rng.uniform(0.5, 1.5, n_light).astype(np.float32)
    return alb, esc


def loss(film, target):
    return float(((np.asarray(film, np.float64) - np.asarray(target, np.float64)) ** 2).sum())


def d_loss(film, target):
    return (2.0 * (np.asarray(film, np.float64) - np.asarray(target, np.float64))).astype(np.float32)


def central_fd(film_fn, alb, esc, target, params):
    """Central differences of loss(film_fn(albedo, emission)) for parameter
    indices `params` (albedos first, then emission scales)."""
    out = []
    na = len(alb)
    for k in params:
        a1, a2, e1, e2 = alb.copy(), alb.copy(), esc.copy(), esc.copy()
        if k < na:
            a1[k] += H
            a2[k] -= H
        else:
            e1[k - na] += H
            e2[k - na] -= H
        f1 = np.asarray(film_fn(a1, e1), np.float64)
        f2 = np.asarray(film_fn(a2, e2), np.float64)
        t = np.asarray(target, np.float64)
        out.append(float(((f1 - f2) * (f1 + f2 - 2.0 * t)).sum()) / (2.0 * H))
    return np.array(out)


def check(grad, fd, label, say=print):
    grad = np.asarray(grad, np.float64)
    fd = np.asarray(fd, np.float64)
    scale = np.abs(fd).max()
    assert scale > 0.0, "every finite difference is zero"
    sig = np.abs(fd) >= 1e-3 * scale
    rel = np.abs(grad - fd) / np.maximum(np.abs(fd), 1e-30)
    worst = float(rel[sig].max()) if sig.any() else 0.0
    small = float(np.abs(grad - fd)[~sig].max() / scale) if (~sig).any() else 0.0
    say(f"{label}: {int(sig.sum())} significant parameters, worst relative error {worst:.2e}; "
        f"{int((~sig).sum())} small, worst error {small:.2e} of the largest")
    assert worst <= REL, f"{label}: relative error {worst:.3e}"
    assert small <= 1e-4, f"{label}: small-gradient error {small:.3e}"
    return worst
