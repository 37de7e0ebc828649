"""The oracle pinned against the reference's own known-answer tests
(tests/golden/reference_kats.json) and against vectors recorded from the
reference's self-contained sources (tests/golden/ref_vectors.npz)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def kats():
    with open(os.path.join(GOLD, "reference_kats.json")) as f:
        return json.load(f)


def refvec():
    return np.load(os.path.join(GOLD, "ref_vectors.npz"))


def _f3(v):
    return np.array(v, np.float32)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def approx(a, x, eps):
    """Catch v1 Approx: |a - x| < eps * (1 + max(|a|, |x|))."""
    return abs(a - x) < eps * (1.0 + max(abs(a), abs(x)))


def test_bbox_ray_kats(oracle):
    L = oracle.lib()
    for c in kats()["bbox_ray"]["cases"]:
        h0 = C.c_float()
        h1 = C.c_float()
        hit = L.ora_bbox_ray(_p(_f3(c["bmin"])), _p(_f3(c["bmax"])), _p(_f3(c["o"])), _p(_f3(c["d"])),
                             float("inf"), C.byref(h0), C.byref(h1))
        assert bool(hit) == c["hit"], c["name"]
        if c["hit"]:
            assert h0.value == c["h0"], c["name"]
            assert c["h1_min"] <= h1.value <= c["h1_max"], c["name"]


def test_bbox2_ray_kats_as_bbox4(oracle):
    L = oracle.lib()
    for c in kats()["bbox2_ray"]["cases"]:
        b = np.zeros((6, 4), np.float32)
        b[0::2, :] = np.inf
        b[1::2, :] = -np.inf
        for lane, (lo, hi) in enumerate(c["boxes"]):
            for ax in range(3):
                b[2 * ax, lane] = lo[ax]
                b[2 * ax + 1, lane] = hi[ax]
        near = np.zeros(4, np.float32)
        mask = L.ora_bbox4_ray(_p(np.ascontiguousarray(b)), _p(_f3(c["o"])), _p(_f3(c["d"])), float("inf"), _p(near))
        assert mask == c["mask"], c["name"]
        for lane, v in c["near"].items():
            assert near[int(lane)] == v, c["name"]


def test_matrix_inverse_kat(oracle):
    k = kats()["matrix_inverse"]
    inv = oracle.matrix_inverse(np.array(k["m"], np.float32)).reshape(16)
    for a, x in zip(inv, k["inverse"]):
        assert approx(float(a), x, k["epsilon"])
    sing = oracle.matrix_inverse(np.array(k["singular"], np.float32))
    assert np.isnan(sing).all()


def test_matrix_inverse_matches_reference_bits(oracle):
    rv = refvec()
    for m, want in zip(rv["inv_in"], rv["inv_out"]):
        got = oracle.matrix_inverse(m)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_vector_kats(oracle):
    L = oracle.lib()
    k = kats()["vector"]
    v1, v2 = _f3(k["v1"]), _f3(k["v2"])
    assert approx(L.ora_vec3_length(_p(v1)), k["length"], k["length_eps"])
    n = np.zeros(3, np.float32)
    L.ora_vec3_normalize(_p(v1), _p(n))
    for a, x in zip(n, k["normalized"]):
        assert approx(float(a), x, k["normalize_eps"])
    assert approx(L.ora_vec3_dot(_p(v1), _p(v2)), k["dot"], k["dot_eps"])
    cr = np.zeros(3, np.float32)
    L.ora_vec3_cross(_p(v1), _p(v2), _p(cr))
    for a, x in zip(cr, k["cross"]):
        assert approx(float(a), x, k["cross_eps"])


def test_vector_length_tiny_branch(oracle):
    # Imath takes the scaled path below 2*FLT_MIN; a plain sqrt of the
    # underflowed square would return 0 here.
    L = oracle.lib()
    v = np.array([3e-30, 4e-30, 0.0], np.float32)
    assert L.ora_vec3_length(_p(v)) == pytest.approx(5e-30, rel=1e-6)


def test_mutable_partition_kat(oracle):
    k = kats()["mutable_partition"]
    v = np.array(k["input"], np.int32)
    calls = C.c_int()
    split = oracle.lib().ora_partition_ints(_p(v), v.size, k["key"], C.byref(calls))
    assert calls.value == k["calls"]
    assert split == k["split"]
    assert v.tolist() == k["output"]


def test_halton_matches_reference_bits(oracle):
    rv = refvec()
    got = np.array([oracle.lib().ora_halton(int(d), int(i)) for d, i in zip(rv["halton_dim"], rv["halton_index"])],
                   np.float32)
    assert np.array_equal(got.view(np.uint32), rv["halton_value"].view(np.uint32))


def test_halton_radical_inverse_closed_form(oracle):
    # Faure's digit permutations: base 3 -> (0,1,2) (plain radical inverse),
    # base 5 -> (0,3,2,1,4).  Sample i of a dimension is the permuted
    # radical inverse of i scaled by (1 - 2^-24); checks digit order and the
    # permutation recurrence (sampling/halton.py:46-87).
    def radinv(i, b, perm, ndig=6):
        r, f = 0.0, 1.0 / b
        for _ in range(ndig):
            r += perm[i % b] * f
            i //= b
            f /= b
        return r
    for i in range(200):
        assert oracle.lib().ora_halton(1, i) == pytest.approx(radinv(i, 3, (0, 1, 2)) * (1 - 2**-24), abs=2e-7)
        assert oracle.lib().ora_halton(2, i) == pytest.approx(radinv(i, 5, (0, 3, 2, 1, 4)) * (1 - 2**-24), abs=2e-7)
    # base 2 is the plain bit reversal
    for i in range(64):
        r = int("{:032b}".format(i)[::-1], 2)
        assert oracle.lib().ora_halton(0, i) == np.float32(r) * np.float32(2**-32)


def test_hash_matches_reference(oracle):
    rv = refvec()
    L = oracle.lib()
    got = np.array([L.ora_hash_u32(int(a), int(b)) for a, b in zip(rv["hash_n"], rv["hash_seed"])], np.uint32)
    assert np.array_equal(got, rv["hash_value"])


def test_oracle_against_ref_build_if_present(oracle):
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(5)
    for _ in range(2000):
        d = int(rng.integers(0, 256))
        i = int(rng.integers(0, 2**32))
        a = np.float32(oracle.lib().ora_halton(d, i))
        b = np.float32(R.psyref_halton(d, i))
        assert a.view(np.uint32) == b.view(np.uint32)


def _bitrev52(i):
    return int("{:052b}".format(i)[::-1], 2)


def test_sobol_closed_form_and_net_property(oracle):
    # Generator matrices built from the definition (van der Corput first
    # dimension, primitive polynomials after it).  Dimension 0 is the 52-bit
    # bit reversal; the scramble XORs its low 52 bits; dimensions 0 and 1
    # form a (0, m, 2)-net for every m: each elementary interval of area
    # 2^-m holds exactly one of the first 2^m points (sampling/sobol.hpp:40-53).
    m = oracle.make_sobol_matrices(4)
    idx = np.arange(1 << 10, dtype=np.uint64)
    d0 = oracle.sobol(m, 0, idx)
    assert np.array_equal(d0, np.array([_bitrev52(int(i)) for i in idx], np.float64) * 2.0**-52)
    scr = 0xFFF0_1234_5678_9ABC
    d0s = oracle.sobol(m, 0, idx, scr)
    want = np.array([_bitrev52(int(i)) ^ (scr & (2**52 - 1)) for i in idx], np.float64) * 2.0**-52
    assert np.array_equal(d0s, want)
    d1 = oracle.sobol(m, 1, idx)
    for mm in range(1, 11):
        n = 1 << mm
        for a in range(mm + 1):
            cx = np.floor(d0[:n] * (1 << a)).astype(np.int64)
            cy = np.floor(d1[:n] * (1 << (mm - a))).astype(np.int64)
            cells = cx * (1 << (mm - a)) + cy
            assert np.unique(cells).size == n, (mm, a)
    # index bits 52 and up read the next dimension's columns
    big = np.array([1 << 52, (1 << 52) | 1], np.uint64)
    got = oracle.sobol(m, 0, big)
    assert got[0] == float(int(m[52])) * 2.0**-52
    assert got[1] == float(int(m[52]) ^ int(m[0])) * 2.0**-52


def test_sobol_matches_reference_vectors(oracle):
    M = oracle.ref_sobol_matrices()
    if M is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    # the reference's table: dimension 0 is van der Corput, dimension 1 the
    # direction numbers 1, 3, 5, 15, ... of x + 1
    assert all(int(M[k]) == 1 << (51 - k) for k in range(52))
    assert [int(M[52 + k]) >> (51 - k) for k in range(4)] == [1, 3, 5, 15]
    rv = refvec()
    for d in np.unique(rv["sobol_dim"]):
        sel = rv["sobol_dim"] == d
        for c in np.unique(rv["sobol_scramble"][sel]):
            s2 = sel & (rv["sobol_scramble"] == c)
            got = oracle.sobol(M, int(d), rv["sobol_index"][s2], int(c))
            assert np.array_equal(got.view(np.uint64), rv["sobol_value"][s2].view(np.uint64)), (d, c)


def test_spectrum_white_point_is_flat(oracle):
    """spectrum_xyz_to_p at the equal-energy white point, recorded from the
    reference (color/spectrum_grid.h), is flat: times 1/equal_energy_
    reflectance it is 1 = (X+Y+Z)/3 at every wavelength, to 2e-6.  That is
    the value the grey-only (no grid) path uses."""
    rv = refvec()
    lam, spec = rv["spectrum_lambda"], rv["spectrum_grey"]
    assert lam.min() == 380.0 and lam.max() == 700.0
    val = spec.astype(np.float64) * (1.0 / np.float64(np.float32(0.009358239977091027)))
    assert np.abs(val - 1.0).max() <= 2e-6
    c = np.ones(3, np.float32)
    assert abs(oracle.lib().ora_color_to_spectrum(_p(c)) - 1.0) <= 1e-6


def _random_xyz(rng, n):
    """Chromaticities over the whole grid and beyond it: XYZ of random
    linear RGB (in gamut), plus arbitrary triples (out of gamut, negative
    sums, zero)."""
    out = []
    M = np.array([[0.4339, 0.3762, 0.1899], [0.2126, 0.7152, 0.0722], [0.0177, 0.1095, 0.8728]], np.float32)
    for k in range(n):
        if k % 4 == 0:
            out.append(rng.uniform(-0.2, 2.0, 3).astype(np.float32))
        else:
            rgb = (rng.uniform(0, 1, 3) ** 3 * rng.uniform(0.01, 100)).astype(np.float32)
            out.append((M @ rgb).astype(np.float32))
    out[0] = np.zeros(3, np.float32)
    return out


def test_spectrum_grid_matches_reference_bits(oracle):
    """The oracle's spectrum_xyz_to_p and Color_to_SpectralSample over the
    reference's own grid tables against the reference's own functions
    (oracle/_ref), bit for bit: bilinear cells, fan cells, outside the
    grid, zero and negative brightness."""
    G = oracle.ref_spectrum_grid()
    if G is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    R = oracle.ref()
    assert G.w * G.h == G.cells.shape[0] and (G.cells[:, 0] == 1).any() and (G.cells[:, 0] == 0).any()
    oracle.spectrum_set_grid(G)
    try:
        rng = np.random.default_rng(3)
        cells_seen = set()
        for xyz in _random_xyz(rng, 6000):
            lam = np.float32(rng.uniform(380, 700))
            a = oracle.spectrum_xyz_to_p(lam, xyz)
            b = np.float32(R.psyref_spectrum_xyz_to_p(float(lam), _p(xyz)))
            assert a.view(np.uint32) == b.view(np.uint32), (lam, xyz, a, b)
            s = float(xyz.sum())
            if s > 0:
                u = G.mat[0] * xyz[0] / s + G.mat[1] * xyz[1] / s + G.mat[2]
                v = G.mat[3] * xyz[0] / s + G.mat[4] * xyz[1] / s + G.mat[5]
                if 0 <= u < G.w and 0 <= v < G.h:
                    cells_seen.add(int(G.cells[int(u) + G.w * int(v), 0]))
        assert cells_seen == {0, 1}  # both interpolation paths ran
        for k in range(500):
            rgb = (rng.uniform(0, 1, 3) * rng.uniform(0.1, 50)).astype(np.float32)
            hero = np.float32(rng.uniform(380, 700))
            got = oracle.color_to_spectral_sample(rgb, hero)
            for i in range(4):
                lam = np.float32(hero + np.float32(i) * np.float32(80.0))
                if lam > 700.0:
                    lam = np.float32(lam - np.float32(320.0))
                want = np.float32(R.psyref_color_to_spectrum(_p(rgb), float(lam)))
                assert got[i].view(np.uint32) == want.view(np.uint32)
        # brightness scales the spectrum (spectrum_xyz_to_p divides by 1/(X+Y+Z))
        for xyz in _random_xyz(rng, 200)[1:]:
            lam = np.float32(rng.uniform(380, 700))
            a = float(oracle.spectrum_xyz_to_p(lam, xyz))
            b = float(oracle.spectrum_xyz_to_p(lam, (xyz * np.float32(4.0)).astype(np.float32)))
            assert abs(b - 4.0 * a) <= 1e-5 * max(abs(b), 1e-20)
    finally:
        oracle.spectrum_set_grid(None)
