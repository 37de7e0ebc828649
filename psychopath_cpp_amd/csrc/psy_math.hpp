// psy_math.hpp -- scalar math shared by the HIP kernels and the host-side
// scene builder.  Float operation order follows the reference so that
// results are bit-identical when compiled with -ffp-contract=off and IEEE
// division / square root:
//   Vec3 = Imath::Vec3<float> (math/vector.hpp:186): dot left-associative,
//   length with the tiny-vector branch, normalize divides by the length;
//   Matrix44 row-vector products (math/matrix.hpp:99-114, :148-156);
//   lerp (utils/utils.hpp:51-91): (1.0 - alpha) is evaluated in double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <math.h>

#define PSY_HD __host__ __device__ __forceinline__

namespace psy {

// bit patterns of u32 fields stored in float planes
PSY_HD uint32_t f2u(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }
PSY_HD float u2f(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }

struct V3 { float x, y, z; };

PSY_HD V3 mk(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
PSY_HD V3 splat(float s) { return mk(s, s, s); }
PSY_HD V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
PSY_HD V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
PSY_HD V3 mulv(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
PSY_HD V3 divv(V3 a, V3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
PSY_HD V3 scale(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
PSY_HD V3 sdiv(V3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
PSY_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PSY_HD V3 cross(V3 a, V3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
PSY_HD float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

PSY_HD float length_tiny(V3 a) {
    float ax = (a.x >= 0.0f) ? a.x : -a.x;
    float ay = (a.y >= 0.0f) ? a.y : -a.y;
    float az = (a.z >= 0.0f) ? a.z : -a.z;
    float mx = ax;
    if (mx < ay) mx = ay;
    if (mx < az) mx = az;
    if (mx == 0.0f) return 0.0f;
    ax /= mx; ay /= mx; az /= mx;
    return mx * sqrtf(ax * ax + ay * ay + az * az);
}
PSY_HD float length2(V3 a) { return dot(a, a); }
PSY_HD float length(V3 a) {
    const float l2 = dot(a, a);
    if (l2 < 2.0f * FLT_MIN) return length_tiny(a);
    return sqrtf(l2);
}
PSY_HD V3 normalized(V3 a) {
    const float l = length(a);
    if (l == 0.0f) return splat(0.0f);
    return mk(a.x / l, a.y / l, a.z / l);
}
// vector.hpp:208-224 and std::min / std::max
PSY_HD V3 vmin(V3 a, V3 b) { return mk(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z); }
PSY_HD V3 vmax(V3 a, V3 b) { return mk(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z); }
PSY_HD float smin(float a, float b) { return (b < a) ? b : a; }
PSY_HD float smax(float a, float b) { return (a < b) ? b : a; }
// _mm_min_ps / _mm_max_ps lane semantics (the second operand when either
// is NaN), as BBox4::intersect_ray uses them (basics/bbox.hpp:566-595)
PSY_HD float sse_min(float a, float b) { return a < b ? a : b; }
PSY_HD float sse_max(float a, float b) { return a > b ? a : b; }
PSY_HD float longest_axis(V3 v) { return smax(smax(fabsf(v.x), fabsf(v.y)), fabsf(v.z)); }

// ---- lerp ------------------------------------------------------------------
PSY_HD float lerp_fac(float alpha) { return (float)(1.0 - (double)alpha); }
PSY_HD float lerpf(float alpha, float a, float b) {
    return (float)((double)a * (1.0 - (double)alpha) + (double)(b * alpha));
}
PSY_HD V3 lerpv(float alpha, V3 a, V3 b) { return add(scale(a, lerp_fac(alpha)), scale(b, alpha)); }
// index split of lerp_seq; -1 single sample, -2 last sample
PSY_HD int lerp_seq_index(float alpha, int n, float* nalpha) {
    if (n == 1) return -1;
    if (alpha >= 1.0f) return -2;
    const float temp = alpha * (float)(n - 1);
    const int index = (int)temp;
    *nalpha = temp - (float)index;
    return index;
}

// ---- Matrix44 (row-major, row-vector convention) ---------------------------
struct M44 { float m[16]; };

PSY_HD M44 m_identity() {
    M44 r;
    for (int i = 0; i < 16; ++i) r.m[i] = 0.0f;
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}
// operator*: r[i][j] = ((b[0][j]*a[i][0] + b[1][j]*a[i][1]) + b[2][j]*a[i][2]) + b[3][j]*a[i][3]
PSY_HD M44 m_mul(const M44& a, const M44& b) {
    M44 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i * 4 + j] = (((b.m[0 * 4 + j] * a.m[i * 4 + 0]) + (b.m[1 * 4 + j] * a.m[i * 4 + 1])) +
                              (b.m[2 * 4 + j] * a.m[i * 4 + 2])) + (b.m[3 * 4 + j] * a.m[i * 4 + 3]);
    return r;
}
PSY_HD M44 m_lerp(float alpha, const M44& a, const M44& b) {
    const float f = lerp_fac(alpha);
    M44 r;
    for (int i = 0; i < 16; ++i) r.m[i] = (a.m[i] * f) + (b.m[i] * alpha);
    return r;
}
PSY_HD M44 m_lerp_seq(float alpha, const M44* seq, int n) {
    float na = 0.0f;
    const int i = lerp_seq_index(alpha, n, &na);
    if (i == -1) return seq[0];
    if (i == -2) return seq[n - 1];
    return m_lerp(na, seq[i], seq[i + 1]);
}
// multVecMatrix (divides every lane by w)
PSY_HD V3 m_pos(const M44& a, V3 s) {
    float r[4];
    for (int j = 0; j < 4; ++j)
        r[j] = (((a.m[0 * 4 + j] * s.x) + (a.m[1 * 4 + j] * s.y)) + (a.m[2 * 4 + j] * s.z)) + a.m[3 * 4 + j];
    const float w = r[3];
    return mk(r[0] / w, r[1] / w, r[2] / w);
}
// multDirMatrix
PSY_HD V3 m_dir(const M44& a, V3 s) {
    float r[3];
    for (int j = 0; j < 3; ++j)
        r[j] = ((a.m[0 * 4 + j] * s.x) + (a.m[1 * 4 + j] * s.y)) + (a.m[2 * 4 + j] * s.z);
    return mk(r[0], r[1], r[2]);
}
// dir_transform_transpose (math/matrix.hpp:415-418), horizontal-add order
PSY_HD V3 m_dir_transpose(const M44& a, V3 s) {
    float r[3];
    for (int i = 0; i < 3; ++i) {
        const float p0 = a.m[i * 4 + 0] * s.x, p1 = a.m[i * 4 + 1] * s.y, p2 = a.m[i * 4 + 2] * s.z,
                    p3 = a.m[i * 4 + 3] * 0.0f;
        r[i] = (p0 + p1) + (p2 + p3);
    }
    return mk(r[0], r[1], r[2]);
}

// Matrix44::inverse: cofactor method in the lane order of Intel AP-928
// (utils/simd.hpp:110-205), finished with an exactly rounded 1/det refined
// by the same Newton step.  The reference's reciprocal estimate (rcpss) is
// CPU-specific; on the device the exact reciprocal is used, which is the
// value that step converges to.
struct L4 { float v[4]; };
PSY_HD L4 l4(float a, float b, float c, float d) { L4 r; r.v[0] = a; r.v[1] = b; r.v[2] = c; r.v[3] = d; return r; }
PSY_HD L4 l4add(L4 a, L4 b) { return l4(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]); }
PSY_HD L4 l4sub(L4 a, L4 b) { return l4(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2], a.v[3] - b.v[3]); }
PSY_HD L4 l4mul(L4 a, L4 b) { return l4(a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]); }
PSY_HD L4 l4sel(L4 a, L4 b, unsigned s) {
    return l4(a.v[s & 3], a.v[(s >> 2) & 3], b.v[(s >> 4) & 3], b.v[(s >> 6) & 3]);
}
PSY_HD L4 l4pairs(L4 a) { return l4sel(a, a, 0xB1); }
PSY_HD L4 l4halves(L4 a) { return l4sel(a, a, 0x4E); }

PSY_HD M44 m_inverse(const M44& a) {
    const float* s = a.m;
    L4 t = l4(s[0], s[1], s[4], s[5]);
    L4 c1 = l4(s[8], s[9], s[12], s[13]);
    const L4 c0 = l4sel(t, c1, 0x88);
    c1 = l4sel(c1, t, 0xDD);
    t = l4(s[2], s[3], s[6], s[7]);
    L4 c3 = l4(s[10], s[11], s[14], s[15]);
    L4 c2 = l4sel(t, c3, 0x88);
    c3 = l4sel(c3, t, 0xDD);
    L4 k0, k1, k2, k3, p;
    p = l4pairs(l4mul(c2, c3));
    k0 = l4mul(c1, p); k1 = l4mul(c0, p);
    p = l4halves(p);
    k0 = l4sub(l4mul(c1, p), k0); k1 = l4sub(l4mul(c0, p), k1);
    k1 = l4halves(k1);
    p = l4pairs(l4mul(c1, c2));
    k0 = l4add(l4mul(c3, p), k0); k3 = l4mul(c0, p);
    p = l4halves(p);
    k0 = l4sub(k0, l4mul(c3, p)); k3 = l4sub(l4mul(c0, p), k3);
    k3 = l4halves(k3);
    p = l4pairs(l4mul(l4halves(c1), c3));
    c2 = l4halves(c2);
    k0 = l4add(l4mul(c2, p), k0); k2 = l4mul(c0, p);
    p = l4halves(p);
    k0 = l4sub(k0, l4mul(c2, p)); k2 = l4sub(l4mul(c0, p), k2);
    k2 = l4halves(k2);
    p = l4pairs(l4mul(c0, c1));
    k2 = l4add(l4mul(c3, p), k2); k3 = l4sub(l4mul(c2, p), k3);
    p = l4halves(p);
    k2 = l4sub(l4mul(c3, p), k2); k3 = l4sub(k3, l4mul(c2, p));
    p = l4pairs(l4mul(c0, c3));
    k1 = l4sub(k1, l4mul(c2, p)); k2 = l4add(l4mul(c1, p), k2);
    p = l4halves(p);
    k1 = l4add(l4mul(c2, p), k1); k2 = l4sub(k2, l4mul(c1, p));
    p = l4pairs(l4mul(c0, c2));
    k1 = l4add(l4mul(c3, p), k1); k3 = l4sub(k3, l4mul(c1, p));
    p = l4halves(p);
    k1 = l4sub(k1, l4mul(c3, p)); k3 = l4add(l4mul(c1, p), k3);
    L4 d = l4mul(c0, k0);
    d = l4add(l4halves(d), d);
    const float det0 = l4pairs(d).v[0] + d.v[0];
    const float est = 1.0f / det0;
    const float inv = (est + est) - det0 * (est * est);
    M44 r;
    for (int j = 0; j < 4; ++j) {
        r.m[0 * 4 + j] = inv * k0.v[j];
        r.m[1 * 4 + j] = inv * k1.v[j];
        r.m[2 * 4 + j] = inv * k2.v[j];
        r.m[3 * 4 + j] = inv * k3.v[j];
    }
    if (inv == 0.0f)
        for (int i = 0; i < 16; ++i) r.m[i] = __builtin_nanf("");
    return r;
}

// ---- BBox (basics/bbox.hpp:138-177) ----------------------------------------
#define PSY_BBOX_MAXT_ADJUST 1.00000024f
// The dicing width std::max(ray.min_width(..) * dice_rate, min_upoly_size)
// of object/patch_utils.hpp:77 for a footprint that is a number.  A NaN
// footprint (NaN ray differentials: a transfer across a zero cosine) makes
// the reference split until its boxes stop being hit, which need not end;
// here it takes min_upoly_size, so such a ray dices to the smallest
// micropolygons and ends.
PSY_HD float dice_width(float scaled_width, float min_upoly) {
    return scaled_width > min_upoly ? scaled_width : min_upoly;
}

PSY_HD bool bbox_ray(V3 bmin, V3 bmax, V3 o, V3 dinv, float t, float* h0, float* h1) {
    const float tx1 = (bmin.x - o.x) * dinv.x;
    const float tx2 = (bmax.x - o.x) * dinv.x;
    float a0 = smin(tx1, tx2);
    float a1 = smax(tx1, tx2);
    const float ty1 = (bmin.y - o.y) * dinv.y;
    const float ty2 = (bmax.y - o.y) * dinv.y;
    a0 = smax(a0, smin(ty1, ty2));
    a1 = smin(a1, smax(ty1, ty2));
    const float tz1 = (bmin.z - o.z) * dinv.z;
    const float tz2 = (bmax.z - o.z) * dinv.z;
    a0 = smax(0.0f, smax(a0, smin(tz1, tz2)));
    a1 = smin(a1, smax(tz1, tz2)) * PSY_BBOX_MAXT_ADJUST;
    *h0 = a0;
    *h1 = a1;
    return a0 <= smin(t, a1);
}

// bbox_ray for a ray whose inverse direction is finite in every component.
// Then no slab distance is NaN, and IEEE min/max give the values of the
// std::min/std::max sequence above (up to the sign of a zero, which no
// result depends on), in fewer instructions.
PSY_HD bool bbox_ray_fast(V3 bmin, V3 bmax, V3 o, V3 dinv, float t, float* h0, float* h1) {
    const float tx1 = (bmin.x - o.x) * dinv.x;
    const float tx2 = (bmax.x - o.x) * dinv.x;
    const float ty1 = (bmin.y - o.y) * dinv.y;
    const float ty2 = (bmax.y - o.y) * dinv.y;
    const float tz1 = (bmin.z - o.z) * dinv.z;
    const float tz2 = (bmax.z - o.z) * dinv.z;
    const float a0 = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fmaxf(fminf(tz1, tz2), 0.0f));
    const float a1 = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fmaxf(tz1, tz2)) * PSY_BBOX_MAXT_ADJUST;
    *h0 = a0;
    *h1 = a1;
    return a0 <= fminf(t, a1);
}
PSY_HD bool finite3(V3 v) { return fabsf(v.x) < INFINITY && fabsf(v.y) < INFINITY && fabsf(v.z) < INFINITY; }

// ---- ray footprint (basics/ray.hpp:122-153, :207-305; utils/utils.hpp:443-456)
struct RayW { float owx, owy, dwx, dwy, fwx, fwy; };

PSY_HD void closest_ray_t2(V3 o1, V3 d1, V3 o2, V3 d2, float* t_out, float* dist_out) {
    const V3 dd = sub(d1, d2);
    const float dd2 = dot(dd, dd);
    if (dd2 < 0.00001f) { *t_out = -1.0f; *dist_out = -1.0f; return; }
    const V3 w = sub(o1, o2);
    const float t = -dot(w, dd) / dd2;
    *t_out = t;
    *dist_out = length(sub(add(o1, scale(d1, t)), add(o2, scale(d2, t))));
}
PSY_HD RayW ray_widths(V3 o, V3 d, V3 odx, V3 ody, V3 ddx, V3 ddy) {
    RayW w;
    float tdx, lx, tdy, ly;
    closest_ray_t2(o, d, add(o, odx), add(d, ddx), &tdx, &lx);
    closest_ray_t2(o, d, add(o, ody), add(d, ddy), &tdy, &ly);
    w.owx = length(odx);
    if (tdx <= 0.0f) { w.dwx = length(ddx); w.fwx = 0.0f; }
    else { w.dwx = (lx - w.owx) / tdx; w.fwx = lx; }
    w.owy = length(ody);
    if (tdy <= 0.0f) { w.dwy = length(ddy); w.fwy = 0.0f; }
    else { w.dwy = (ly - w.owy) / tdy; w.fwy = ly; }
    return w;
}
// Ray::min_width with its two flip distances precomputed (they depend on
// the footprint only): min_width(r, a, b) == min_width_flip(r, flip_x(r),
// flip_y(r), a, b) bit for bit.
PSY_HD float flip_x(const RayW& r) { return (r.owx - r.fwx) / r.dwx; }
PSY_HD float flip_y(const RayW& r) { return (r.owy - r.fwy) / r.dwy; }
PSY_HD float min_width_flip(const RayW& r, float tflipx, float tflipy, float tnear, float tfar) {
    float minx, miny;
    if (tnear < tflipx && tfar > tflipx) minx = r.fwx;
    else minx = smin(fabsf((r.owx - r.fwx) + (r.dwx * tnear)) + r.fwx,
                     fabsf((r.owx - r.fwx) + (r.dwx * tfar)) + r.fwx);
    if (tnear < tflipy && tfar > tflipy) miny = r.fwy;
    else miny = smin(fabsf((r.owy - r.fwy) + (r.dwy * tnear)) + r.fwy,
                     fabsf((r.owy - r.fwy) + (r.dwy * tfar)) + r.fwy);
    return smin(minx, miny);
}
PSY_HD float min_width(const RayW& r, float tnear, float tfar) {
    const float tflipx = (r.owx - r.fwx) / r.dwx;
    const float tflipy = (r.owy - r.fwy) / r.dwy;
    float minx, miny;
    if (tnear < tflipx && tfar > tflipx) minx = r.fwx;
    else minx = smin(fabsf((r.owx - r.fwx) + (r.dwx * tnear)) + r.fwx,
                     fabsf((r.owx - r.fwx) + (r.dwx * tfar)) + r.fwx);
    if (tnear < tflipy && tfar > tflipy) miny = r.fwy;
    else miny = smin(fabsf((r.owy - r.fwy) + (r.dwy * tnear)) + r.fwy,
                     fabsf((r.owy - r.fwy) + (r.dwy * tfar)) + r.fwy);
    return smin(minx, miny);
}

}  // namespace psy
