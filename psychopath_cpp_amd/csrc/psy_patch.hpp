// psy_patch.hpp -- bilinear / bicubic split-dice primitives for the dice
// cache.  The split tree of a patch is fixed by its geometry alone
// (object/patch_utils.hpp:146-192: split U when ulen > vlen of time sample
// 0, the halves by midpoint subdivision), so every node of that tree can be
// produced independently by replaying the splits along its path from the
// root.
//   Bilinear ops: object/bilinear.hpp:44-141; Bicubic: object/bicubic.hpp:55-189.
#pragma once

#include "psy_math.hpp"

namespace psy {

enum { PATCH_BILINEAR = 2, PATCH_BICUBIC = 3 };

template <int NPTS>
struct Patch { V3 p[NPTS]; };

PSY_HD V3 mid(V3 a, V3 b) { return scale(add(a, b), 0.5f); }

PSY_HD float patch_ulen(const Patch<4>& p) { return longest_axis(sub(p.p[0], p.p[1])); }
PSY_HD float patch_vlen(const Patch<4>& p) { return longest_axis(sub(p.p[0], p.p[2])); }
PSY_HD float patch_ulen(const Patch<16>& p) { return longest_axis(sub(p.p[0], p.p[3])); }
PSY_HD float patch_vlen(const Patch<16>& p) { return longest_axis(sub(p.p[0], p.p[12])); }

// keep the `second` half of a U split (bilinear.hpp:63-73)
PSY_HD void split_u_keep(Patch<4>& p, bool second) {
    const V3 m01 = mid(p.p[0], p.p[1]);
    const V3 m23 = mid(p.p[2], p.p[3]);
    if (second) { p.p[0] = m01; p.p[2] = m23; }
    else { p.p[1] = m01; p.p[3] = m23; }
}
PSY_HD void split_v_keep(Patch<4>& p, bool second) {
    const V3 m02 = mid(p.p[0], p.p[2]);
    const V3 m13 = mid(p.p[1], p.p[3]);
    if (second) { p.p[0] = m02; p.p[1] = m13; }
    else { p.p[2] = m02; p.p[3] = m13; }
}
// bicubic halves by de Casteljau at 1/2 per row / column (bicubic.hpp:74-110)
PSY_HD void split_u_keep(Patch<16>& p, bool second) {
    for (int r = 0; r < 4; ++r) {
        const int rr = r * 4;
        const V3 tmp = mid(p.p[rr + 1], p.p[rr + 2]);
        const V3 b3 = p.p[rr + 3];
        const V3 b2 = mid(p.p[rr + 3], p.p[rr + 2]);
        const V3 b1 = mid(tmp, b2);
        const V3 a0 = p.p[rr + 0];
        const V3 a1 = mid(p.p[rr + 0], p.p[rr + 1]);
        const V3 a2 = mid(tmp, a1);
        const V3 a3 = mid(a2, b1);
        if (second) { p.p[rr + 0] = a3; p.p[rr + 1] = b1; p.p[rr + 2] = b2; p.p[rr + 3] = b3; }
        else { p.p[rr + 0] = a0; p.p[rr + 1] = a1; p.p[rr + 2] = a2; p.p[rr + 3] = a3; }
    }
}
PSY_HD void split_v_keep(Patch<16>& p, bool second) {
    for (int c = 0; c < 4; ++c) {
        const V3 tmp = mid(p.p[c + 4], p.p[c + 8]);
        const V3 b3 = p.p[c + 12];
        const V3 b2 = mid(p.p[c + 12], p.p[c + 8]);
        const V3 b1 = mid(tmp, b2);
        const V3 a0 = p.p[c + 0];
        const V3 a1 = mid(p.p[c + 0], p.p[c + 4]);
        const V3 a2 = mid(tmp, a1);
        const V3 a3 = mid(a2, b1);
        if (second) { p.p[c + 0] = a3; p.p[c + 4] = b1; p.p[c + 8] = b2; p.p[c + 12] = b3; }
        else { p.p[c + 0] = a0; p.p[c + 4] = a1; p.p[c + 8] = a2; p.p[c + 12] = a3; }
    }
}

template <int N>
PSY_HD void patch_bound(const Patch<N>& p, V3* bmin, V3* bmax) {
    V3 lo = p.p[0], hi = p.p[0];
    for (int i = 1; i < N; ++i) { lo = vmin(lo, p.p[i]); hi = vmax(hi, p.p[i]); }
    *bmin = lo;
    *bmax = hi;
}

// One split step along a path, tracking the uv rectangle exactly as the
// reference's uv_stack does ((a+b)*0.5 per split).
template <int N>
PSY_HD void patch_step(Patch<N>& p, bool second, float uv[4]) {
    if (patch_ulen(p) > patch_vlen(p)) {
        split_u_keep(p, second);
        const float m = (uv[0] + uv[1]) * 0.5f;
        if (second) uv[0] = m; else uv[1] = m;
    } else {
        split_v_keep(p, second);
        const float m = (uv[2] + uv[3]) * 0.5f;
        if (second) uv[2] = m; else uv[3] = m;
    }
}

// Heap numbering, 1-based: node k has children 2k (first half) and 2k+1
// (second half); its depth is floor(log2 k) and its path from the root is
// the bits of k below the leading one, MSB first (1 = second half).
PSY_HD int heap_depth(uint32_t k) { return 31 - __builtin_clz(k); }

// Replay the splits from `root` down to heap node k.
template <int N>
PSY_HD void patch_at_node(Patch<N>& p, uint32_t k, float uv[4]) {
    uv[0] = 0.0f; uv[1] = 1.0f; uv[2] = 0.0f; uv[3] = 1.0f;
    const int depth = heap_depth(k);
    for (int l = depth - 1; l >= 0; --l) patch_step(p, ((k >> l) & 1u) != 0, uv);
}

// Differential geometry at (u, v) of the un-split patch.
template <int N>
PSY_HD void patch_dg(const Patch<N>& p, float u, float v, V3* n, V3* dpdu, V3* dpdv, V3* dndu, V3* dndv);

PSY_HD V3 bl_eval_p(float u, V3 p0, V3 p1) { return add(scale(p0, 1.0f - u), scale(p1, u)); }
PSY_HD V3 bl_eval_pd(V3 p0, V3 p1) { return add(scale(p0, -1.0f), scale(p1, 1.0f)); }

PSY_HD void dg_tail(V3 dpdu_, V3 dpdv_, V3 n_, V3 d2pduu, V3 d2pduv, V3 d2pdvv, V3* dndu, V3* dndv) {
    const float E = dot(dpdu_, dpdu_), F = dot(dpdu_, dpdv_), G = dot(dpdv_, dpdv_);
    const float e = dot(n_, d2pduu), f = dot(n_, d2pduv), g = dot(n_, d2pdvv);
    const float invEGF2 = 1.0f / ((E * G) - (F * F));
    *dndu = add(scale(dpdu_, ((f * F) - (e * G)) * invEGF2), scale(dpdv_, ((e * F) - (f * E)) * invEGF2));
    *dndv = add(scale(dpdu_, ((g * F) - (f * G)) * invEGF2), scale(dpdv_, ((f * F) - (g * E)) * invEGF2));
}

template <>
PSY_HD void patch_dg<4>(const Patch<4>& p, float u, float v, V3* n, V3* dpdu, V3* dpdv, V3* dndu, V3* dndv) {
    *dpdu = bl_eval_pd(bl_eval_p(v, p.p[0], p.p[2]), bl_eval_p(v, p.p[1], p.p[3]));
    *dpdv = bl_eval_pd(bl_eval_p(u, p.p[0], p.p[1]), bl_eval_p(u, p.p[2], p.p[3]));
    *n = normalized(cross(*dpdv, *dpdu));
    const V3 d2pduv = bl_eval_pd(bl_eval_pd(p.p[0], p.p[1]), bl_eval_pd(p.p[2], p.p[3]));
    dg_tail(*dpdu, *dpdv, *n, splat(0.0f), d2pduv, splat(0.0f), dndu, dndv);
}

PSY_HD V3 bc_eval_p(float u, V3 p0, V3 p1, V3 p2, V3 p3) {
    const float iu = 1.0f - u;
    const float b0 = iu * iu * iu, b1 = 3.0f * u * iu * iu, b2 = 3.0f * u * u * iu, b3 = u * u * u;
    return add(add(add(scale(p0, b0), scale(p1, b1)), scale(p2, b2)), scale(p3, b3));
}
PSY_HD V3 bc_eval_pd(float u, V3 p0, V3 p1, V3 p2, V3 p3) {
    const float iu = 1.0f - u;
    const float d0 = -3.0f * iu * iu;
    const float d1 = (3.0f * iu * iu) - (6.0f * iu * u);
    const float d2 = (6.0f * iu * u) - (3.0f * u * u);
    const float d3 = 3.0f * u * u;
    return add(add(add(scale(p0, d0), scale(p1, d1)), scale(p2, d2)), scale(p3, d3));
}
PSY_HD V3 bc_eval_pdd(float u, V3 p0, V3 p1, V3 p2, V3 p3) {
    const float iu = 1.0f - u;
    const float dd0 = 6.0f * iu, dd1 = (6.0f * u) - (12.0f * iu), dd2 = (6.0f * iu) - (12.0f * u), dd3 = 6.0f * u;
    return add(add(add(scale(p0, dd0), scale(p1, dd1)), scale(p2, dd2)), scale(p3, dd3));
}

template <>
PSY_HD void patch_dg<16>(const Patch<16>& p, float u, float v, V3* n, V3* dpdu, V3* dpdv, V3* dndu, V3* dndv) {
    V3 pu[4], pv[4], pdv[4];
    for (int i = 0; i < 4; ++i) {
        pu[i] = bc_eval_p(v, p.p[i], p.p[i + 4], p.p[i + 8], p.p[i + 12]);
        pv[i] = bc_eval_p(u, p.p[i * 4], p.p[i * 4 + 1], p.p[i * 4 + 2], p.p[i * 4 + 3]);
        pdv[i] = bc_eval_pd(u, p.p[i * 4], p.p[i * 4 + 1], p.p[i * 4 + 2], p.p[i * 4 + 3]);
    }
    *dpdu = bc_eval_pd(u, pu[0], pu[1], pu[2], pu[3]);
    *dpdv = bc_eval_pd(v, pv[0], pv[1], pv[2], pv[3]);
    *n = normalized(cross(*dpdv, *dpdu));
    dg_tail(*dpdu, *dpdv, *n, bc_eval_pdd(u, pu[0], pu[1], pu[2], pu[3]),
            bc_eval_pd(v, pdv[0], pdv[1], pdv[2], pdv[3]), bc_eval_pdd(v, pv[0], pv[1], pv[2], pv[3]),
            dndu, dndv);
}

}  // namespace psy
