"""Film output through libpsy_gpu.so: the (height, width, 3) XYZ film of
gpu.render as dithered 8-bit sRGB (Film::scanline_image_8bbc), a PNG file
(write_png_from_film) or a float PFM dump.

    film, _, _ = gpu.render(ctx, scene, W, H, spp)
    film_out.write_png("frame.png", film)
"""
import ctypes as C

import numpy as np

from .gpu import check, lib


def _host_film(film):
    if hasattr(film, "detach"):
        film = film.detach().cpu().numpy()
    f = np.ascontiguousarray(film, np.float32)
    if f.ndim != 3 or f.shape[2] != 3:
        raise ValueError("film must be (height, width, 3)")
    return f


def to_srgb8(film, sample_counts=None, dither_seed=1):
    """(height, width, 3) uint8 sRGB; pixels with a zero sample count show
    a grey checkerboard; dither_seed 0 turns the dither off."""
    f = _host_film(film)
    h, w, _ = f.shape
    cnt = None
    if sample_counts is not None:
        cnt = np.ascontiguousarray(sample_counts, np.int32).reshape(h, w)
    out = np.empty((h, w, 3), np.uint8)
    check(lib().psy_film_to_srgb8(f.ctypes.data_as(C.c_void_p), None if cnt is None else cnt.ctypes.data_as(C.c_void_p),
                                  w, h, int(dither_seed) & 0xffffffff, out.ctypes.data_as(C.c_void_p)),
          "psy_film_to_srgb8")
    return out


def write_png(path, film, sample_counts=None, dither_seed=1):
    img = to_srgb8(film, sample_counts, dither_seed)
    h, w, _ = img.shape
    check(lib().psy_write_png(str(path).encode(), img.ctypes.data_as(C.c_void_p), w, h), "psy_write_png")


def write_pfm(path, film):
    f = _host_film(film)
    h, w, _ = f.shape
    check(lib().psy_write_pfm(str(path).encode(), f.ctypes.data_as(C.c_void_p), w, h), "psy_write_pfm")
