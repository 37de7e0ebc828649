/*
 * ora_core.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement of the reference's scalar math, in the reference's float
 * operation order.  Only tests/, the smoke check and bench.py's CPU-baseline
 * leg may load the oracle; the product never links it.
 *
 * Build with -ffp-contract=off and SSE float math (x86-64 default), as the
 * reference is built without FMA (CMakeLists.txt:29,37-47).
 */
#ifndef ORA_CORE_H
#define ORA_CORE_H

#include <math.h>
#include <float.h>
#include <stdint.h>
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Vec3 = Imath::Vec3<float> (math/vector.hpp:186) ------- */
typedef struct { float x, y, z; } v3;

static inline v3 v3m(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 v3s(float s) { v3 r = {s, s, s}; return r; }
static inline v3 v3add(v3 a, v3 b) { return v3m(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3sub(v3 a, v3 b) { return v3m(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3mul(v3 a, v3 b) { return v3m(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 v3div(v3 a, v3 b) { return v3m(a.x / b.x, a.y / b.y, a.z / b.z); }
/* Vec3 * T and T * Vec3 (Imath: component * scalar) */
static inline v3 v3scale(v3 a, float s) { return v3m(a.x * s, a.y * s, a.z * s); }
static inline v3 v3sdiv(v3 a, float s) { return v3m(a.x / s, a.y / s, a.z / s); }
static inline float v3get(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline void v3set(v3* a, int i, float f) { if (i == 0) a->x = f; else if (i == 1) a->y = f; else a->z = f; }
/* Imath dot: x*v.x + y*v.y + z*v.z, left-associative */
static inline float v3dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* Imath cross */
static inline v3 v3cross(v3 a, v3 b) {
    return v3m(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float v3len2(v3 a) { return v3dot(a, a); }
/* Imath Vec3::lengthTiny */
static inline float v3len_tiny(v3 a) {
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
/* Imath Vec3::length: tiny-vector branch below 2*FLT_MIN */
static inline float v3len(v3 a) {
    float l2 = v3dot(a, a);
    if (l2 < 2.0f * FLT_MIN) return v3len_tiny(a);
    return sqrtf(l2);
}
/* Imath Vec3::normalized: divide by length (zero vector stays zero) */
static inline v3 v3norm(v3 a) {
    float l = v3len(a);
    if (l == 0.0f) return v3s(0.0f);
    return v3m(a.x / l, a.y / l, a.z / l);
}
/* vector.hpp:208-224 */
static inline v3 v3min(v3 a, v3 b) {
    return v3m(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z);
}
static inline v3 v3max(v3 a, v3 b) {
    return v3m(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z);
}
/* std::min / std::max on floats */
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }
static inline float longest_axis(v3 v) {
    return smax(smax(fabsf(v.x), fabsf(v.y)), fabsf(v.z));
}

/* ---------------- lerp (utils/utils.hpp:51-55) ---------------------------
 * lerp<float>: a*(1.0-alpha) is a double product; the float product b*alpha
 * is promoted; the double sum is rounded to float on return.
 * lerp<Vec3/BBox/Matrix/...>: (1.0-alpha) is rounded to float first. */
static inline float lerpf(float alpha, float a, float b) {
    return (float)((double)a * (1.0 - (double)alpha) + (double)(b * alpha));
}
static inline float lerp_fac(float alpha) { return (float)(1.0 - (double)alpha); }
static inline v3 lerpv(float alpha, v3 a, v3 b) {
    return v3add(v3scale(a, lerp_fac(alpha)), v3scale(b, alpha));
}
/* lerp_seq index/alpha split (utils/utils.hpp:77-91); returns the index,
 * *nalpha the local factor; -1 means "single sample", -2 "last sample". */
static inline int lerp_seq_index(float alpha, int n, float* nalpha) {
    if (n == 1) return -1;
    if (alpha >= 1.0f) return -2;
    float temp = alpha * (float)(n - 1);
    int index = (int)temp;
    *nalpha = temp - (float)index;
    return index;
}
static inline float lerp_seq_f(float alpha, const float* seq, int n, int stride) {
    float na;
    int i = lerp_seq_index(alpha, n, &na);
    if (i == -1) return seq[0];
    if (i == -2) return seq[(n - 1) * stride];
    return lerpf(na, seq[i * stride], seq[(i + 1) * stride]);
}

/* ---------------- BBox (basics/bbox.hpp:24-177) ------------------------- */
typedef struct { v3 min, max; } bbox;
#define BBOX_MAXT_ADJUST 1.00000024f

static inline bbox bbox_empty(void) {
    bbox b; b.min = v3s(INFINITY); b.max = v3s(-INFINITY); return b;
}
static inline bbox bbox_merge(bbox a, bbox b) {  /* merge_with / operator| */
    bbox r;
    r.min = v3m(a.min.x < b.min.x ? a.min.x : b.min.x, a.min.y < b.min.y ? a.min.y : b.min.y,
                a.min.z < b.min.z ? a.min.z : b.min.z);
    r.max = v3m(a.max.x > b.max.x ? a.max.x : b.max.x, a.max.y > b.max.y ? a.max.y : b.max.y,
                a.max.z > b.max.z ? a.max.z : b.max.z);
    return r;
}
static inline bbox bbox_lerp(float alpha, bbox a, bbox b) {
    float f = lerp_fac(alpha);
    bbox r;
    r.min = v3add(v3scale(a.min, f), v3scale(b.min, alpha));
    r.max = v3add(v3scale(a.max, f), v3scale(b.max, alpha));
    return r;
}
static inline bbox bbox_lerp_seq(float alpha, const bbox* seq, int n) {
    float na;
    int i = lerp_seq_index(alpha, n, &na);
    if (i == -1) return seq[0];
    if (i == -2) return seq[n - 1];
    return bbox_lerp(na, seq[i], seq[i + 1]);
}
/* BBox::intersect_ray(ray, &hitt0, &hitt1, t): std::min/max semantics */
static inline int bbox_intersect(const bbox* b, v3 o, v3 dinv, float t, float* h0, float* h1) {
    const float tx1 = (b->min.x - o.x) * dinv.x;
    const float tx2 = (b->max.x - o.x) * dinv.x;
    float a0 = smin(tx1, tx2);
    float a1 = smax(tx1, tx2);
    const float ty1 = (b->min.y - o.y) * dinv.y;
    const float ty2 = (b->max.y - o.y) * dinv.y;
    a0 = smax(a0, smin(ty1, ty2));
    a1 = smin(a1, smax(ty1, ty2));
    const float tz1 = (b->min.z - o.z) * dinv.z;
    const float tz2 = (b->max.z - o.z) * dinv.z;
    a0 = smax(0.0f, smax(a0, smin(tz1, tz2)));
    a1 = smin(a1, smax(tz1, tz2)) * BBOX_MAXT_ADJUST;
    *h0 = a0;
    *h1 = a1;
    return a0 <= smin(t, a1);
}

/* BBox4 (basics/bbox.hpp:461-595): lanes with SSE min/max semantics,
 * _mm_min_ps(a,b) = a<b?a:b, _mm_max_ps(a,b) = a>b?a:b. */
typedef struct { float b[6][4]; } bbox4;  /* minx maxx miny maxy minz maxz */
static inline float sse_min(float a, float b) { return a < b ? a : b; }
static inline float sse_max(float a, float b) { return a > b ? a : b; }
static inline unsigned bbox4_intersect(const bbox4* n, v3 o, v3 dinv, float tmax, float near_out[4]) {
    const int ds0 = dinv.x < 0.0f, ds1 = dinv.y < 0.0f, ds2 = dinv.z < 0.0f;
    unsigned mask = 0;
    for (int i = 0; i < 4; ++i) {
        const float xlo = (n->b[0 + ds0][i] - o.x) * dinv.x;
        const float xhi = (n->b[1 - ds0][i] - o.x) * dinv.x;
        const float ylo = (n->b[2 + ds1][i] - o.y) * dinv.y;
        const float yhi = (n->b[3 - ds1][i] - o.y) * dinv.y;
        const float zlo = (n->b[4 + ds2][i] - o.z) * dinv.z;
        const float zhi = (n->b[5 - ds2][i] - o.z) * dinv.z;
        const float mins = sse_max(sse_max(xlo, ylo), sse_max(zlo, 0.0f));
        const float maxs = sse_max(sse_min(sse_min(xhi, yhi), zhi), -INFINITY) * BBOX_MAXT_ADJUST;
        if ((mins < tmax) && (mins <= maxs)) mask |= 1u << i;
        near_out[i] = mins;
    }
    return mask;
}
static inline bbox4 bbox4_from(bbox a, bbox b, bbox c, bbox d) {
    bbox4 r;
    const bbox* bs[4] = {&a, &b, &c, &d};
    for (int i = 0; i < 4; ++i) {
        r.b[0][i] = bs[i]->min.x; r.b[1][i] = bs[i]->max.x;
        r.b[2][i] = bs[i]->min.y; r.b[3][i] = bs[i]->max.y;
        r.b[4][i] = bs[i]->min.z; r.b[5][i] = bs[i]->max.z;
    }
    return r;
}

/* ---------------- Matrix44 (math/matrix.hpp:18-400) ---------------------- */
typedef struct { float m[4][4]; } mat4;

static inline mat4 mat4_identity(void) {
    mat4 r; memset(&r, 0, sizeof r);
    r.m[0][0] = r.m[1][1] = r.m[2][2] = r.m[3][3] = 1.0f;
    return r;
}
static inline mat4 mat4_nan(void) {
    mat4 r;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) r.m[i][j] = NAN;
    return r;
}
/* operator*: r.row[i] = m.row0*a[i][0] + m.row1*a[i][1] + m.row2*a[i][2] + m.row3*a[i][3] */
static inline mat4 mat4_mul(const mat4* a, const mat4* b) {
    mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = (((b->m[0][j] * a->m[i][0]) + (b->m[1][j] * a->m[i][1])) +
                         (b->m[2][j] * a->m[i][2])) + (b->m[3][j] * a->m[i][3]);
    return r;
}
static inline mat4 mat4_lerp(float alpha, const mat4* a, const mat4* b) {
    float f = lerp_fac(alpha);
    mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = (a->m[i][j] * f) + (b->m[i][j] * alpha);
    return r;
}
static inline mat4 mat4_lerp_seq(float alpha, const mat4* seq, int n) {
    float na;
    int i = lerp_seq_index(alpha, n, &na);
    if (i == -1) return seq[0];
    if (i == -2) return seq[n - 1];
    return mat4_lerp(na, &seq[i], &seq[i + 1]);
}
/* multVecMatrix: r = row0*x + row1*y + row2*z + row3; r /= r[3] */
static inline v3 mat4_pos(const mat4* a, v3 s) {
    float r[4];
    for (int j = 0; j < 4; ++j)
        r[j] = (((a->m[0][j] * s.x) + (a->m[1][j] * s.y)) + (a->m[2][j] * s.z)) + a->m[3][j];
    const float w = r[3];
    return v3m(r[0] / w, r[1] / w, r[2] / w);
}
/* multDirMatrix */
static inline v3 mat4_dir(const mat4* a, v3 s) {
    float r[3];
    for (int j = 0; j < 3; ++j)
        r[j] = ((a->m[0][j] * s.x) + (a->m[1][j] * s.y)) + (a->m[2][j] * s.z);
    return v3m(r[0], r[1], r[2]);
}
/* dir_transform_transpose (math/matrix.hpp:415-418): SSE hadd order
 * ((a0+a1)+(a2+a3)) with a3 = 0. */
static inline v3 mat4_dir_transpose(const mat4* a, v3 s) {
    float r[3];
    for (int i = 0; i < 3; ++i) {
        const float p0 = a->m[i][0] * s.x, p1 = a->m[i][1] * s.y, p2 = a->m[i][2] * s.z,
                    p3 = a->m[i][3] * 0.0f;
        r[i] = (p0 + p1) + (p2 + p3);
    }
    return v3m(r[0], r[1], r[2]);
}

/* Matrix44::inverse via the cofactor/Cramer SSE method of Intel AP-928
 * (utils/simd.hpp:110-205); defined in ora_math.c. */
mat4 ora_mat4_inverse(const mat4* a);
float ora_rcp_newton(float det);

#ifdef __cplusplus
}
#endif

#endif /* ORA_CORE_H */
