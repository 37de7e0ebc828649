"""TEST INFRASTRUCTURE ONLY: ctypes bindings of the oracle (liboracle.so)
and of oracle/_ref/libpsyref.so (the reference's own self-contained
sources).  Imported only by tests/, the smoke check and bench.py's
CPU-baseline leg, as the checker -- never by the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libpsyref.so")
REFERENCE_TREE = "/root/reference"

WRAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("odx", "<f4", 3), ("ody", "<f4", 3),
                       ("ddx", "<f4", 3), ("ddy", "<f4", 3), ("time", "<f4"), ("type", "<u4")])
INTER_DTYPE = np.dtype([("hit", "<i4"), ("backfacing", "<i4"), ("id", "<u8"), ("t", "<f4"),
                        ("light_pdf", "<f4"), ("space", "<f4", 16), ("u", "<f4"), ("v", "<f4"),
                        ("p", "<f4", 3), ("dpdu", "<f4", 3), ("dpdv", "<f4", 3), ("n", "<f4", 3),
                        ("dndu", "<f4", 3), ("dndv", "<f4", 3), ("offset", "<f4", 3),
                        ("closure", "<i4"), ("closure_col", "<f4", 3), ("closure_prob", "<f4"),
                        ("patch_dim", "<f4"), ("closure_param", "<f4", 3)], align=True)
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("max_t", "<f4"), ("time", "<f4"),
                      ("owx", "<f4"), ("owy", "<f4"), ("dwx", "<f4"), ("dwy", "<f4"), ("fwx", "<f4"),
                      ("fwy", "<f4"), ("flags", "<u4")])

_lib = None
_ref = None


def build(with_ref=None):
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if with_ref is None:
        with_ref = os.path.isdir(REFERENCE_TREE)
    if with_ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={REFERENCE_TREE}"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(with_ref=False)
        L = C.CDLL(LIB)
        vp, i64, u32 = C.c_void_p, C.c_int64, C.c_uint32
        L.ora_halton.restype = C.c_float
        L.ora_halton.argtypes = [u32, u32]
        L.ora_halton_batch.argtypes = [u32, vp, i64, vp]
        L.ora_scene_set_params.restype = C.c_int
        L.ora_scene_set_params.argtypes = [vp, vp, C.c_int, vp, C.c_int]
        L.ora_sobol_batch.argtypes = [vp, u32, u32, vp, i64, C.c_uint64, vp]
        L.ora_hash_u32.restype = u32
        L.ora_hash_u32.argtypes = [u32, u32]
        L.ora_sampler_offset.restype = u32
        L.ora_sampler_offset.argtypes = [u32, u32, u32, u32]
        L.ora_sampler_get.restype = C.c_float
        L.ora_sampler_get.argtypes = [u32, u32]
        L.ora_matrix_inverse.argtypes = [vp, vp]
        L.ora_bbox_ray.restype = C.c_int
        L.ora_bbox_ray.argtypes = [vp, vp, vp, vp, C.c_float, vp, vp]
        L.ora_bbox4_ray.restype = C.c_uint
        L.ora_bbox4_ray.argtypes = [vp, vp, vp, C.c_float, vp]
        L.ora_world_to_rays.argtypes = [vp, i64, vp]
        L.ora_vec3_length.restype = C.c_float
        L.ora_vec3_length.argtypes = [vp]
        L.ora_vec3_normalize.argtypes = [vp, vp]
        L.ora_vec3_dot.restype = C.c_float
        L.ora_vec3_dot.argtypes = [vp, vp]
        L.ora_vec3_cross.argtypes = [vp, vp, vp]
        L.ora_partition_ints.restype = C.c_int
        L.ora_partition_ints.argtypes = [vp, C.c_int, C.c_int, vp]
        L.ora_scene_create.restype = vp
        L.ora_scene_create.argtypes = [vp]
        L.ora_scene_destroy.argtypes = [vp]
        L.ora_scene_bvh4_nodes.restype = i64
        L.ora_scene_bvh4_nodes.argtypes = [vp, C.c_int]
        L.ora_trace.restype = i64
        L.ora_trace.argtypes = [vp, vp, i64, vp, C.c_int, i64]
        L.ora_trace_brute.restype = i64
        L.ora_trace_brute.argtypes = [vp, vp, i64, vp, C.c_int, i64]
        L.ora_camera_rays.argtypes = [vp, C.c_int, C.c_int, C.c_int, u32, i64, i64, vp, vp]
        L.ora_bounce_rays.argtypes = [vp, vp, vp, i64, i64, C.c_int, C.c_int, u32, u32, u32, vp, vp]
        L.ora_render.restype = i64
        L.ora_render.argtypes = [vp, C.c_int, C.c_int, C.c_int, u32, C.c_int, C.c_int, vp]
        L.ora_render_sobol.restype = i64
        L.ora_render_sobol.argtypes = [vp, C.c_int, C.c_int, C.c_int, u32, C.c_int, C.c_int, vp, vp, u32]
        L.ora_color_to_spectrum.restype = C.c_float
        L.ora_color_to_spectrum.argtypes = [vp]
        L.ora_spectrum_set_grid.restype = C.c_int
        L.ora_spectrum_set_grid.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int, C.c_int, C.c_float, C.c_float, vp]
        L.ora_spectrum_xyz_to_p.restype = C.c_float
        L.ora_spectrum_xyz_to_p.argtypes = [C.c_float, vp]
        L.ora_color_to_spectral_sample.argtypes = [vp, C.c_float, vp]
        f = C.c_float
        L.ora_kat_sphere_t.restype = C.c_int
        L.ora_kat_sphere_t.argtypes = [vp, vp, vp, f, f, vp]
        L.ora_kat_closest_ray_t2.argtypes = [vp, vp, vp, vp, vp, vp]
        L.ora_kat_min_width.restype = f
        L.ora_kat_min_width.argtypes = [vp, f, f]
        L.ora_kat_patch_split.argtypes = [C.c_int, vp, C.c_int, vp, vp]
        L.ora_kat_patch_dg.argtypes = [C.c_int, vp, f, f, vp, vp, vp, vp, vp]
        L.ora_kat_instance_id.restype = C.c_uint64
        L.ora_kat_instance_id.argtypes = [vp, vp, C.c_int]
        L.ora_kat_sphere_light_pdf.restype = f
        L.ora_kat_sphere_light_pdf.argtypes = [vp, f, vp]
        L.ora_kat_rect_light_pdf.restype = f
        L.ora_kat_rect_light_pdf.argtypes = [f, f, vp]
        L.ora_kat_sphere_light_sample.argtypes = [vp, f, vp, f, f, vp, vp]
        L.ora_kat_rect_light_sample.argtypes = [f, f, vp, f, f, vp, vp]
        L.ora_kat_faster_ln.restype = f
        L.ora_kat_faster_ln.argtypes = [f]
        L.ora_kat_fast_logit.restype = f
        L.ora_kat_fast_logit.argtypes = [f, f]
        L.ora_kat_square_to_circle.argtypes = [vp, vp]
        L.ora_kat_cie_1931.argtypes = [f, vp]
        L.ora_kat_sphere_lambert.restype = f
        L.ora_kat_sphere_lambert.argtypes = [f, f]
        L.ora_kat_lambert.argtypes = [f, vp, vp, vp, vp, vp]
        L.ora_kat_gtr_params.argtypes = [vp, vp]
        L.ora_kat_gtr_dist.restype = f
        L.ora_kat_gtr_dist.argtypes = [vp, f]
        L.ora_kat_gtr_half_theta.restype = f
        L.ora_kat_gtr_half_theta.argtypes = [vp, f]
        L.ora_kat_gtr.argtypes = [f, vp, vp, vp, vp, vp, vp]
        L.ora_kat_gtr_sample.argtypes = [vp, vp, vp, f, f, vp]
        L.ora_kat_gtr_estimate.restype = f
        L.ora_kat_gtr_estimate.argtypes = [vp, vp, vp, f, vp]
        L.ora_kat_axis_angle.argtypes = [vp, f, vp]
        L.ora_film_adaptive.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, f, C.c_int, vp, vp, vp]
        L.ora_subdiv_patches_ex.restype = i64
        L.ora_subdiv_patches_ex.argtypes = [vp, C.c_int, vp, C.c_int, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(C.c_void_p)]
        L.ora_free.argtypes = [vp]
        L.ora_kat_light_select.restype = C.c_int
        L.ora_kat_light_select.argtypes = [vp, vp, vp, vp, f, f, f, f, vp, vp, vp, vp]
        L.ora_kat_total_lights.restype = i64
        L.ora_kat_total_lights.argtypes = [vp]
        L.ora_bounce_dir.argtypes = [vp, vp, f, f, vp, vp, vp, vp]
        L.ora_camera_ray_from_dims.argtypes = [vp, f, f, f, f, f, f, f, vp]
        _lib = L
    return _lib


def ref():
    """oracle/_ref/libpsyref.so or None when it was not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB):
            return None
        R = C.CDLL(REF_LIB)
        R.psyref_halton.restype = C.c_float
        R.psyref_halton.argtypes = [C.c_uint32, C.c_uint32]
        R.psyref_sobol.restype = C.c_double
        R.psyref_sobol.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64]
        R.psyref_sobol_matrix.restype = C.c_uint64
        R.psyref_sobol_matrix.argtypes = [C.c_uint32]
        R.psyref_sobol_dims.restype = C.c_uint32
        R.psyref_hash_u32.restype = C.c_uint32
        R.psyref_hash_u32.argtypes = [C.c_uint32, C.c_uint32]
        R.psyref_invert44.restype = C.c_float
        R.psyref_invert44.argtypes = [C.c_void_p]
        R.psyref_spectrum_xyz_to_p.restype = C.c_float
        R.psyref_spectrum_xyz_to_p.argtypes = [C.c_float, C.c_void_p]
        if hasattr(R, "psyref_spectrum_tables"):
            pp = C.POINTER
            R.psyref_spectrum_tables.argtypes = [pp(C.c_void_p), pp(C.c_int), pp(C.c_int), pp(C.c_void_p),
                                                 pp(C.c_int), pp(C.c_int), pp(C.c_float), pp(C.c_float),
                                                 pp(C.c_void_p)]
            R.psyref_color_to_spectrum.restype = C.c_float
            R.psyref_color_to_spectrum.argtypes = [C.c_void_p, C.c_float]
        _ref = R
    return _ref


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def halton(dim, index):
    index = np.ascontiguousarray(index, np.uint32)
    out = np.empty(index.shape, np.float32)
    lib().ora_halton_batch(dim, _p(index), index.size, _p(out))
    return out


def sobol(matrices, dim, index, scramble=0):
    """Sobol::sample over an index array with the given generator matrices
    (uint64, 52 columns per dimension)."""
    m = np.ascontiguousarray(matrices, np.uint64).reshape(-1)
    index = np.ascontiguousarray(index, np.uint64)
    out = np.empty(index.shape, np.float64)
    lib().ora_sobol_batch(_p(m), m.size, dim, _p(index), index.size, int(scramble), _p(out))
    return out


def ref_sobol_matrices():
    """The reference's own generator table (Sobol::Matrices::matrices) read
    from oracle/_ref, or None when that library was not built."""
    R = ref()
    if R is None:
        return None
    n = R.psyref_sobol_dims() * 52
    return np.array([R.psyref_sobol_matrix(i) for i in range(n)], np.uint64)


class SpectrumGrid:
    """The spectrum grid tables (color/spectrum_grid.h) as arrays: cells
    (w*h, 8) int32 = inside, num_points, idx[6]; points (n, 4 + n_samples)
    float32 = xystar, uv, spectrum; the xy -> uv matrix (6,)."""

    def __init__(self, cells, w, h, points, n_samples, smin, smax, mat):
        self.cells = np.ascontiguousarray(cells, np.int32).reshape(w * h, 8)
        self.points = np.ascontiguousarray(points, np.float32).reshape(-1, 4 + n_samples)
        self.w, self.h, self.n_samples = int(w), int(h), int(n_samples)
        self.smin, self.smax = float(smin), float(smax)
        self.mat = np.ascontiguousarray(mat, np.float32).reshape(6)

    @property
    def n_points(self):
        return self.points.shape[0]


def ref_spectrum_grid():
    """The reference's own spectrum grid tables read from oracle/_ref, or
    None when that library was not built."""
    R = ref()
    if R is None or not hasattr(R, "psyref_spectrum_tables"):
        return None
    cells, pts, mat = C.c_void_p(), C.c_void_p(), C.c_void_p()
    w, h, n, ns = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    smin, smax = C.c_float(), C.c_float()
    R.psyref_spectrum_tables(C.byref(cells), C.byref(w), C.byref(h), C.byref(pts), C.byref(n), C.byref(ns),
                             C.byref(smin), C.byref(smax), C.byref(mat))
    cv = np.ctypeslib.as_array(C.cast(cells, C.POINTER(C.c_int32)), (w.value * h.value * 8,)).copy()
    pv = np.ctypeslib.as_array(C.cast(pts, C.POINTER(C.c_float)), (n.value * (4 + ns.value),)).copy()
    mv = np.ctypeslib.as_array(C.cast(mat, C.POINTER(C.c_float)), (6,)).copy()
    return SpectrumGrid(cv, w.value, h.value, pv, ns.value, smin.value, smax.value, mv)


def spectrum_set_grid(grid):
    """Hand a SpectrumGrid to the oracle (None clears it: grey colours only)."""
    if grid is None:
        rc = lib().ora_spectrum_set_grid(None, 0, 0, None, 0, 0, 0.0, 0.0, None)
    else:
        rc = lib().ora_spectrum_set_grid(_p(grid.cells), grid.w, grid.h, _p(grid.points), grid.n_points,
                                         grid.n_samples, grid.smin, grid.smax, _p(grid.mat))
    if rc != 0:
        raise ValueError("ora_spectrum_set_grid: bad table")


def spectrum_xyz_to_p(lam, xyz):
    xyz = np.ascontiguousarray(xyz, np.float32)
    return np.float32(lib().ora_spectrum_xyz_to_p(float(lam), _p(xyz)))


def color_to_spectral_sample(rgb, hero):
    rgb = np.ascontiguousarray(rgb, np.float32)
    out = np.zeros(4, np.float32)
    lib().ora_color_to_spectral_sample(_p(rgb), float(hero), _p(out))
    return out


def make_sobol_matrices(n_dims, seed=1):
    """Generator matrices of a Sobol' sequence built from its definition
    (Bratley & Fox 1988): dimension 0 is the van der Corput sequence, each
    further one uses a primitive polynomial over GF(2) with random odd
    initial direction numbers.  A valid Sobol' construction, not the
    Joe-Kuo table; for tests that need matrices without the reference."""
    rng = np.random.default_rng(seed)
    polys = []  # (degree, interior coefficient bits)
    deg = 1
    while len(polys) < n_dims:
        for a in range(1 << (deg - 1)):
            p = (1 << deg) | (a << 1) | 1
            if _primitive_gf2(p, deg):
                polys.append((deg, a))
                if len(polys) >= n_dims:
                    break
        deg += 1
    out = np.zeros((n_dims, 52), np.uint64)
    for k in range(52):
        out[0, k] = 1 << (51 - k)
    for d in range(1, n_dims):
        s, a = polys[d - 1]
        m = [int(rng.integers(0, 1 << i)) * 2 + 1 if i else 1 for i in range(s)]
        for i in range(s, 52):
            v = m[i - s] ^ (m[i - s] << s)
            for k in range(1, s):
                if (a >> (s - 1 - k)) & 1:
                    v ^= m[i - k] << k
            m.append(v)
        for i in range(52):
            out[d, i] = m[i] << (51 - i)
    return out.reshape(-1)


def _primitive_gf2(p, deg):
    order = (1 << deg) - 1
    if deg == 1:
        return True
    x, e = 2, 1
    while True:  # smallest e with x^e == 1 mod p
        x <<= 1
        if x >> deg:
            x ^= p
        e += 1
        if x == 1:
            return e == order
        if e > order:
            return False


def matrix_inverse(m):
    m = np.ascontiguousarray(m, np.float32).reshape(16)
    out = np.empty(16, np.float32)
    lib().ora_matrix_inverse(_p(m), _p(out))
    return out.reshape(4, 4)


def world_to_rays(wrays):
    out = np.empty(wrays.shape[0], RAY_DTYPE)
    lib().ora_world_to_rays(_p(wrays), wrays.shape[0], _p(out))
    return out


class OracleScene:
    def __init__(self, scene):
        self._desc = scene.to_desc() if hasattr(scene, "to_desc") else scene
        self.handle = lib().ora_scene_create(self._desc.ptr())
        if not self.handle:
            raise RuntimeError("ora_scene_create failed")

    def close(self):
        if self.handle:
            lib().ora_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, albedo=None, emission_scale=None):
        """Grey albedo per shader and emission scale per light object (the
        product's psy_scene_set_params)."""
        a = None if albedo is None else np.ascontiguousarray(albedo, np.float32)
        e = None if emission_scale is None else np.ascontiguousarray(emission_scale, np.float32)
        rc = lib().ora_scene_set_params(self.handle, None if a is None else _p(a), 0 if a is None else a.size,
                                        None if e is None else _p(e), 0 if e is None else e.size)
        if rc:
            raise ValueError("ora_scene_set_params: parameter counts differ from the scene")

    def trace(self, wrays, threads=1, chunk=65536, brute=False):
        wrays = np.ascontiguousarray(wrays, WRAY_DTYPE)
        out = np.zeros(wrays.shape[0], INTER_DTYPE)
        fn = lib().ora_trace_brute if brute else lib().ora_trace
        fn(self.handle, _p(wrays), wrays.shape[0], _p(out), int(threads), int(chunk))
        return out

    def camera_rays(self, width, height, spp, seed, first=0, count=None):
        n = width * height * spp if count is None else count
        out = np.zeros(n, WRAY_DTYPE)
        wl = np.zeros(n, np.float32)
        lib().ora_camera_rays(self.handle, width, height, spp, seed, first, n, _p(out), _p(wl))
        return out, wl

    def render(self, width, height, spp, seed, threads=8, path_length=4, sobol_matrices=None):
        """Per-sample film contributions (height, width, spp, 3) and rays traced; with sobol_matrices
        (uint64, 52 columns a dimension) every path dimension comes from Sobol::sample."""
        out = np.zeros((height, width, spp, 3), np.float32)
        if sobol_matrices is None:
            rays = lib().ora_render(self.handle, width, height, spp, seed, int(path_length), int(threads), _p(out))
        else:
            m = np.ascontiguousarray(sobol_matrices, np.uint64)
            rays = lib().ora_render_sobol(self.handle, width, height, spp, seed, int(path_length), int(threads),
                                          _p(out), _p(m), m.size)
        if rays < 0:
            raise RuntimeError("oracle render: chromatic colours are not restated")
        return out, int(rays)

    def bounce_rays(self, wrays, inters, width, spp, seed, first=0, dim_u=9, dim_v=10):
        n = wrays.shape[0]
        out = np.zeros(n, WRAY_DTYPE)
        valid = np.zeros(n, np.uint32)
        lib().ora_bounce_rays(self.handle, _p(np.ascontiguousarray(wrays)), _p(np.ascontiguousarray(inters)),
                              first, n, width, spp, seed, dim_u, dim_v, _p(out), _p(valid))
        return out, valid


def wrays_from_planes(planes):
    """(20, n) float32 plane layout -> WRAY_DTYPE records."""
    planes = np.asarray(planes, np.float32)
    n = planes.shape[1]
    w = np.zeros(n, WRAY_DTYPE)
    w["o"] = planes[0:3].T
    w["d"] = planes[3:6].T
    w["odx"] = planes[6:9].T
    w["ody"] = planes[9:12].T
    w["ddx"] = planes[12:15].T
    w["ddy"] = planes[15:18].T
    w["time"] = planes[18]
    w["type"] = planes[19].view(np.uint32)
    return w


def planes_from_wrays(w):
    """WRAY_DTYPE records -> (20, n) float32 plane layout."""
    n = w.shape[0]
    p = np.zeros((20, n), np.float32)
    p[0:3] = w["o"].T
    p[3:6] = w["d"].T
    p[6:9] = w["odx"].T
    p[9:12] = w["ody"].T
    p[12:15] = w["ddx"].T
    p[15:18] = w["ddy"].T
    p[18] = w["time"]
    p[19] = w["type"].astype(np.uint32).view(np.float32)
    return p


def film_adaptive(samples, spp_max, variance_max, bucket_size):
    """ora_film_adaptive over samples of shape (H, W, spp, 3): (film (H, W, 3),
    variance (H, W, 3), passes per bucket in row-major bucket order)."""
    s = np.ascontiguousarray(samples, np.float32)
    H, W, spp, _ = s.shape
    film = np.zeros((H, W, 3), np.float32)
    var = np.zeros((H, W, 3), np.float32)
    nb = ((W + bucket_size - 1) // bucket_size) * ((H + bucket_size - 1) // bucket_size)
    passes = np.zeros(nb, np.int32)
    lib().ora_film_adaptive(_p(s), W, H, spp, int(spp_max), np.float32(variance_max), int(bucket_size), _p(film),
                            _p(var), _p(passes))
    return film, var, passes


def subdiv_patches(verts, face_vert_counts, face_vert_indices, max_level=5, min_level=0, prune=True,
                   cap_all=False):
    """ora_subdiv_patches_ex: Bezier patches (n, 16, 3) of a Catmull-Clark
    control mesh; None when the oracle refuses the mesh."""
    v = np.ascontiguousarray(np.asarray(verts, np.float32).reshape(-1, 3))
    cnt = np.ascontiguousarray(np.asarray(face_vert_counts, np.int32).reshape(-1))
    idx = np.ascontiguousarray(np.asarray(face_vert_indices, np.int32).reshape(-1))
    out = C.c_void_p()
    n = lib().ora_subdiv_patches_ex(_p(v), v.shape[0], _p(cnt), cnt.size, _p(idx), int(max_level), int(min_level),
                                    int(bool(prune)), int(bool(cap_all)), C.byref(out))
    if n < 0:
        return None
    res = np.zeros((n, 16, 3), np.float32)
    if n:
        C.memmove(res.ctypes.data, out.value, n * 48 * 4)
    lib().ora_free(out)
    return res
