/*
 * ora_math.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Matrix44::inverse() restated (math/matrix.hpp:319-329 calling
 * SIMD::invert_44_matrix, utils/simd.hpp:110-205).  The reference uses the
 * cofactor ("Cramer") method published by Intel (AP-928) on 4-lane SSE
 * registers, finishing with an approximate reciprocal of the determinant
 * (_mm_rcp_ss) refined by one Newton step.  Bit parity needs the same
 * lane-wise operation order, so the steps below follow that method; the
 * 4-lane values are modelled by a plain struct.
 *
 * The reciprocal estimate is the host CPU's own rcpss instruction when built
 * for x86 (it is what the reference computes on that host; its low bits are
 * CPU-vendor specific), else an exactly rounded 1/det.
 */
#include "ora_core.h"

#if defined(__x86_64__) || defined(__i386__)
#include <xmmintrin.h>
#define ORA_HAVE_RCPSS 1
#endif

typedef struct { float v[4]; } lane4;

static inline lane4 l4(float a, float b, float c, float d) { lane4 r = {{a, b, c, d}}; return r; }
static inline lane4 l4add(lane4 a, lane4 b) { return l4(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]); }
static inline lane4 l4sub(lane4 a, lane4 b) { return l4(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2], a.v[3] - b.v[3]); }
static inline lane4 l4mul(lane4 a, lane4 b) { return l4(a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]); }
/* SSE shufps selector semantics: two lanes from a, two from b */
static inline lane4 l4shuf(lane4 a, lane4 b, unsigned imm) {
    return l4(a.v[imm & 3], a.v[(imm >> 2) & 3], b.v[(imm >> 4) & 3], b.v[(imm >> 6) & 3]);
}
/* pair swap (1,0,3,2) and half swap (2,3,0,1) */
static inline lane4 l4pairs(lane4 a) { return l4shuf(a, a, 0xB1); }
static inline lane4 l4halves(lane4 a) { return l4shuf(a, a, 0x4E); }

float ora_rcp_newton(float det) {
    float est;
#ifdef ORA_HAVE_RCPSS
    est = _mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(det)));
#else
    est = 1.0f / det;
#endif
    /* one Newton step: (e+e) - det*(e*e) */
    return (est + est) - det * (est * est);
}

mat4 ora_mat4_inverse(const mat4* a) {
    const float* s = &a->m[0][0];
    /* columns of the matrix, two of them with swapped halves */
    lane4 t = l4(s[0], s[1], s[4], s[5]);
    lane4 c1 = l4(s[8], s[9], s[12], s[13]);
    const lane4 c0 = l4shuf(t, c1, 0x88);
    c1 = l4shuf(c1, t, 0xDD);
    t = l4(s[2], s[3], s[6], s[7]);
    lane4 c3 = l4(s[10], s[11], s[14], s[15]);
    lane4 c2 = l4shuf(t, c3, 0x88);
    c3 = l4shuf(c3, t, 0xDD);

    lane4 k0, k1, k2, k3, p;
    /* products of column pairs (2,3) */
    p = l4pairs(l4mul(c2, c3));
    k0 = l4mul(c1, p);
    k1 = l4mul(c0, p);
    p = l4halves(p);
    k0 = l4sub(l4mul(c1, p), k0);
    k1 = l4sub(l4mul(c0, p), k1);
    k1 = l4halves(k1);
    /* (1,2) */
    p = l4pairs(l4mul(c1, c2));
    k0 = l4add(l4mul(c3, p), k0);
    k3 = l4mul(c0, p);
    p = l4halves(p);
    k0 = l4sub(k0, l4mul(c3, p));
    k3 = l4sub(l4mul(c0, p), k3);
    k3 = l4halves(k3);
    /* (1 swapped, 3) */
    p = l4pairs(l4mul(l4halves(c1), c3));
    c2 = l4halves(c2);
    k0 = l4add(l4mul(c2, p), k0);
    k2 = l4mul(c0, p);
    p = l4halves(p);
    k0 = l4sub(k0, l4mul(c2, p));
    k2 = l4sub(l4mul(c0, p), k2);
    k2 = l4halves(k2);
    /* (0,1) */
    p = l4pairs(l4mul(c0, c1));
    k2 = l4add(l4mul(c3, p), k2);
    k3 = l4sub(l4mul(c2, p), k3);
    p = l4halves(p);
    k2 = l4sub(l4mul(c3, p), k2);
    k3 = l4sub(k3, l4mul(c2, p));
    /* (0,3) */
    p = l4pairs(l4mul(c0, c3));
    k1 = l4sub(k1, l4mul(c2, p));
    k2 = l4add(l4mul(c1, p), k2);
    p = l4halves(p);
    k1 = l4add(l4mul(c2, p), k1);
    k2 = l4sub(k2, l4mul(c1, p));
    /* (0,2) */
    p = l4pairs(l4mul(c0, c2));
    k1 = l4add(l4mul(c3, p), k1);
    k3 = l4sub(k3, l4mul(c1, p));
    p = l4halves(p);
    k1 = l4sub(k1, l4mul(c3, p));
    k3 = l4add(l4mul(c1, p), k3);

    /* determinant: horizontal sum in the SSE order */
    lane4 d = l4mul(c0, k0);
    d = l4add(l4halves(d), d);
    const float det0 = l4pairs(d).v[0] + d.v[0];
    const float inv = ora_rcp_newton(det0);

    mat4 r;
    for (int j = 0; j < 4; ++j) {
        r.m[0][j] = inv * k0.v[j];
        r.m[1][j] = inv * k1.v[j];
        r.m[2][j] = inv * k2.v[j];
        r.m[3][j] = inv * k3.v[j];
    }
    if (inv == 0.0f) return mat4_nan();
    return r;
}
