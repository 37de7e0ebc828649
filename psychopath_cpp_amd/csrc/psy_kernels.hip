// psy_kernels.hip -- HIP kernels for gfx950 (MI355X).
//
// The reference traces breadth-first: it partitions a ray stream at every
// BVH4 node and every patch split (tracer/tracer.cpp:38-313,
// accel/bvh4.cpp:163-236, object/patch_utils.hpp:19-200).  On CDNA4 the
// same closest hit is found with one lane per ray: the lane walks the BVH4
// nearest-first with a short stack in LDS, and walks each patch's split
// tree through the precomputed dice cache without any per-ray stack (heap
// order: the next node is found from the current index alone).  A patch
// hit's t is the midpoint of a micro-bbox slab interval, which does not
// depend on the ray's max_t, so the result equals the reference's for any
// visit order except exact ties.
//
// Arithmetic follows the reference's operation order; this file must be
// compiled with -ffp-contract=off and IEEE division / square root.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "psy_device.hpp"
#include "psy_patch.hpp"
#include "../../include/psy_gpu.h"

namespace psy {

// ---------------------------------------------------------------------------
// Halton + Sampler (sampling/halton.cpp, sampling/image_sampler.hpp:16-38,
// utils/hash.hpp:6-16)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float halton_sample(const HaltonDim* dims, const uint16_t* perm, uint32_t dim,
                                               uint32_t index) {
    if (dim == 0) {
        const uint32_t r = __builtin_bitreverse32(index);
        return (float)r * (1.0f / 4294967296.0f);
    }
    if (dim >= 256) return 0.0f;
    const HaltonDim& h = dims[dim];
    uint32_t sum = 0;
    for (uint32_t t = 0; t < h.n_terms; ++t)
        sum += (uint32_t)perm[h.perm_offset + (index / h.div[t]) % h.pow_base] * h.power[t];
    return (float)((double)sum * h.scale);
}

__device__ __forceinline__ uint32_t hash_u32(uint32_t n, uint32_t seed) {
    uint32_t hash = n;
    for (int i = 0; i < 3; ++i) {
        hash *= 1936502639u;
        hash ^= hash >> 16;
        hash += seed;
    }
    return hash;
}

__device__ __forceinline__ uint32_t sampler_offset(uint32_t x, uint32_t y, uint32_t n, uint32_t seed) {
    return hash_u32(x ^ ((y >> 16) | (y << 16)), seed) + n;
}

__device__ __forceinline__ float sampler_get(const HaltonDim* dims, const uint16_t* perm, uint32_t off,
                                             uint32_t dim) {
    // dimension reorder (sampling/image_sampler.hpp:26)
    uint32_t d = dim;
    if (dim < 11) {
        const uint8_t tbl[11] = {10, 7, 6, 5, 4, 2, 9, 8, 3, 1, 0};
        d = tbl[dim];
    }
    return halton_sample(dims, perm, d, off);
}

__global__ void k_halton(const HaltonDim* dims, const uint16_t* perm, uint32_t dim, const uint32_t* idx,
                         int64_t n, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = halton_sample(dims, perm, dim, idx[i]);
}

__global__ void k_pixel_samples(const HaltonDim* dims, const uint16_t* perm, int x0, int y0, int w, int h,
                                int spp, uint32_t seed, uint32_t dim0, int n_dims, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)w * h * spp;
    if (i >= total) return;
    const int s = (int)(i % spp);
    const int64_t pix = i / spp;
    const int px = x0 + (int)(pix % w);
    const int py = y0 + (int)(pix / w);
    const uint32_t off = sampler_offset((uint32_t)px, (uint32_t)py, (uint32_t)s, seed);
    for (int k = 0; k < n_dims; ++k) out[i * n_dims + k] = sampler_get(dims, perm, off, dim0 + (uint32_t)k);
}

// ---------------------------------------------------------------------------
// Dice cache builder: one lane per split-tree node, replaying the splits
// from the patch root (object/patch_utils.hpp:38-46, :146-192).  Node k of a
// patch (1-based heap) lives at entries 2k, 2k+1 of the patch's slot, so the
// two children of a node form one aligned 64-byte pair.
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void load_patch(const float* odata, int32_t off, Patch<N>& p) {
    for (int i = 0; i < N; ++i) p.p[i] = mk(odata[off + 3 * i], odata[off + 3 * i + 1], odata[off + 3 * i + 2]);
}

template <int N>
__device__ void build_node(const float* odata, const DevObject& ob, uint32_t k, float4* out) {
    Patch<N> p;
    load_patch<N>(odata, ob.data, p);
    float uv[4];
    patch_at_node<N>(p, k, uv);
    V3 lo, hi;
    patch_bound<N>(p, &lo, &hi);
    out[0] = make_float4(lo.x, lo.y, lo.z, (uv[0] + uv[1]) * 0.5f);
    out[1] = make_float4(hi.x, hi.y, hi.z, (uv[2] + uv[3]) * 0.5f);
}

__global__ void k_build_dice(const float* odata, const DevObject* objs, const int32_t* patch_ids, int n_patches,
                             int levels, float4* dice) {
    const int64_t per = (int64_t)1 << (levels + 1);
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= per * n_patches) return;
    const int pi = (int)(gid / per);
    const uint32_t k = (uint32_t)(gid % per);
    const DevObject ob = objs[patch_ids[pi]];
    float4* out = dice + 2 * (ob.dice + k);
    if (k == 0) {  // unused slot of the 1-based heap
        out[0] = make_float4(0.f, 0.f, 0.f, 0.f);
        out[1] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    if (ob.kind == K_BILINEAR) build_node<4>(odata, ob, k, out);
    else build_node<16>(odata, ob, k, out);
}

// ---------------------------------------------------------------------------
// Traversal
// ---------------------------------------------------------------------------
struct Best {
    float t;
    uint32_t inst;
    float u, v, aux;
};

__device__ __forceinline__ V3 xyz(float4 a) { return mk(a.x, a.y, a.z); }

// The leaf rule of intersect_rays_with_patch (object/patch_utils.hpp:76-133):
// a hit micro-bbox no larger than the ray footprint (times dice_rate) ends
// the descent; its hit distance is the middle of the slab interval.
__device__ __forceinline__ bool accept_leaf(float h0, float h1, float u, float v, float maxd, uint32_t inst, Best& b) {
    const float tt = (h0 + h1) * 0.5f;
    if (tt > 0.0f && tt < b.t) {
        b.t = tt;
        b.inst = inst;
        b.u = u;
        b.v = v;
        b.aux = maxd;
        return true;
    }
    return false;
}

// Split-dice below the cached depth: depth-first over the subtree of heap
// node `k` (already found to need splitting), replaying the splits from the
// patch root for every visited node.  `sidx` is the reference's split-stack
// index of k (object/patch_utils.hpp:80: a node at stack index 63 is a leaf).
template <int N>
__device__ __noinline__ bool dice_deep(const DevScene& s, const DevObject& ob, uint32_t k0, uint32_t inst, V3 o,
                                       V3 dinv, RayW w, bool occl, Best& b) {
    Patch<N> base;
    load_patch<N>(s.odata, ob.data, base);
    float uvb[4];
    patch_at_node<N>(base, k0, uvb);
    const int sidx0 = __builtin_popcount(k0) - 1;
    int k = 1;
    uint64_t pb = 1;
    int sidx = sidx0 + 1;
    for (;;) {
        Patch<N> p = base;
        float uv[4] = {uvb[0], uvb[1], uvb[2], uvb[3]};
        for (int l = k - 1; l >= 0; --l) patch_step<N>(p, ((pb >> l) & 1ull) != 0, uv);
        V3 lo, hi;
        patch_bound<N>(p, &lo, &hi);
        float h0, h1;
        bool down = false;
        if (bbox_ray(lo, hi, o, dinv, b.t, &h0, &h1)) {
            const float width = smax(min_width(w, h0, h1) * s.dice_rate, s.min_upoly);
            const float max_dim = longest_axis(sub(hi, lo));
            if (max_dim <= width || sidx == 63 || k == 62) {
                if (accept_leaf(h0, h1, (uv[0] + uv[1]) * 0.5f, (uv[2] + uv[3]) * 0.5f, max_dim, inst, b) && occl)
                    return true;
            } else {
                down = true;
            }
        }
        if (down) { ++k; pb = (pb << 1) | 1ull; ++sidx; continue; }
        for (;;) {
            if (k == 0) return false;
            if (pb & 1ull) { pb ^= 1ull; --sidx; break; }
            pb >>= 1; --k;
        }
        if (k == 0) return false;
    }
}

template <bool DEEP>
__device__ __forceinline__ bool deep_or_flag(const DevScene& s, const DevObject& ob, uint32_t k, uint32_t inst, V3 o,
                                             V3 dinv, const RayW& w, bool occl, Best& b, bool& need_deep) {
    if (!DEEP) {
        need_deep = true;
        return false;
    }
    return ob.kind == K_BILINEAR ? dice_deep<4>(s, ob, k, inst, o, dinv, w, occl, b)
                                 : dice_deep<16>(s, ob, k, inst, o, dinv, w, occl, b);
}

// Walk one patch's cached split tree, nearest child first, without a
// per-ray stack: `trail` keeps one bit per depth whose other child is still
// pending, and a node's sibling and ancestors follow from its heap index.
// Any visit order finds the reference's closest hit (the leaf t does not
// depend on max_t); within a node the second half is taken on ties, as the
// reference visits it first.  Returns true when an occlusion ray is done.
template <bool DEEP>
__device__ __forceinline__ bool dice_walk(const DevScene& s, const DevObject& ob, uint32_t inst, V3 o, V3 dinv,
                                          const RayW& w, bool occl, Best& b, bool& need_deep) {
    const float4* nodes = s.dice + 2 * ob.dice;
    const int levels = ob.dice_levels;
    {
        const float4 a = nodes[2], c = nodes[3];
        float h0, h1;
        if (!bbox_ray(xyz(a), xyz(c), o, dinv, b.t, &h0, &h1)) return false;
        const float maxd = longest_axis(sub(xyz(c), xyz(a)));
        if (maxd <= smax(min_width(w, h0, h1) * s.dice_rate, s.min_upoly))
            return accept_leaf(h0, h1, a.w, c.w, maxd, inst, b) && occl;
        if (levels == 0) return deep_or_flag<DEEP>(s, ob, 1, inst, o, dinv, w, occl, b, need_deep);
    }
    uint32_t k = 1;
    int depth = 0;
    uint32_t trail = 0;
    for (;;) {
        // expand k: its children 2k, 2k+1 are at depth+1 <= levels
        const uint32_t c0 = 2 * k;
        const float4* q = nodes + 2 * c0;
        const float4 a0 = q[0], e0 = q[1], a1 = q[2], e1 = q[3];
        float h00, h01, h10, h11;
        const bool hit0 = bbox_ray(xyz(a0), xyz(e0), o, dinv, b.t, &h00, &h01);
        const bool hit1 = bbox_ray(xyz(a1), xyz(e1), o, dinv, b.t, &h10, &h11);
        bool split0 = false, split1 = false;
        if (hit1) {
            const float maxd = longest_axis(sub(xyz(e1), xyz(a1)));
            if (maxd <= smax(min_width(w, h10, h11) * s.dice_rate, s.min_upoly)) {
                if (accept_leaf(h10, h11, a1.w, e1.w, maxd, inst, b) && occl) return true;
            } else {
                split1 = true;
            }
        }
        if (hit0) {
            const float maxd = longest_axis(sub(xyz(e0), xyz(a0)));
            if (maxd <= smax(min_width(w, h00, h01) * s.dice_rate, s.min_upoly)) {
                if (accept_leaf(h00, h01, a0.w, e0.w, maxd, inst, b) && occl) return true;
            } else {
                split0 = true;
            }
        }
        split0 = split0 && h00 <= b.t;
        split1 = split1 && h10 <= b.t;
        const int cd = depth + 1;
        if (cd == levels) {
            if (split1 && deep_or_flag<DEEP>(s, ob, c0 + 1, inst, o, dinv, w, occl, b, need_deep)) return true;
            if (split0 && deep_or_flag<DEEP>(s, ob, c0, inst, o, dinv, w, occl, b, need_deep)) return true;
            split0 = split1 = false;
        }
        if (split0 && split1) {
            k = (h00 < h10) ? c0 : c0 + 1;
            trail |= 1u << cd;
            depth = cd;
            continue;
        }
        if (split0 || split1) {
            k = split0 ? c0 : c0 + 1;
            depth = cd;
            continue;
        }
        // resume at the deepest pending sibling
        for (;;) {
            if (trail == 0) return false;
            const int dd = 31 - __builtin_clz(trail);
            k = (k >> (depth - dd)) ^ 1u;
            depth = dd;
            trail &= ~(1u << dd);
            const float4 a = nodes[2 * k], c = nodes[2 * k + 1];
            float h0, h1;
            if (!bbox_ray(xyz(a), xyz(c), o, dinv, b.t, &h0, &h1)) continue;
            const float maxd = longest_axis(sub(xyz(c), xyz(a)));
            if (maxd <= smax(min_width(w, h0, h1) * s.dice_rate, s.min_upoly)) {
                if (accept_leaf(h0, h1, a.w, c.w, maxd, inst, b) && occl) return true;
                continue;
            }
            break;
        }
    }
}

// Sphere::intersect_ray / SphereLight::intersect_ray quadratic
// (object/sphere.cpp:67-140)
__device__ __forceinline__ bool sphere_t(V3 ro, V3 rd, float max_t, V3 cent, float radi, float* t_out) {
    const V3 o = sub(ro, cent);
    const float a = dot(rd, rd);
    const float bq = 2.0f * dot(rd, o);
    const float c = dot(o, o) - radi * radi;
    float disc = bq * bq - 4.0f * a * c;
    if (disc < 0.0f) return false;
    disc = sqrtf(disc);
    float q;
    if (bq < 0.0f) q = -0.5f * (bq - disc);
    else q = -0.5f * (bq + disc);
    float t0 = q / a, t1;
    if (q != 0.0f) t1 = c / q; else t1 = max_t;
    if (t0 > t1) { const float tmp = t1; t1 = t0; t0 = tmp; }
    if (t0 >= max_t || t1 < 0.0001f) return false;
    if (t0 >= 0.0001f) *t_out = t0;
    else if (t1 < max_t) *t_out = t1;
    else return false;
    return true;
}

__device__ __forceinline__ void sphere_at(const float* d, int ts, int stride, float time, V3* cent, float* radi) {
    float na = 0.0f;
    const int i = lerp_seq_index(time, ts, &na);
    if (i == -1) { *cent = mk(d[0], d[1], d[2]); *radi = d[3]; return; }
    if (i == -2) {
        const float* e = d + (ts - 1) * stride;
        *cent = mk(e[0], e[1], e[2]); *radi = e[3]; return;
    }
    const float* a = d + i * stride;
    const float* b = d + (i + 1) * stride;
    *cent = lerpv(na, mk(a[0], a[1], a[2]), mk(b[0], b[1], b[2]));
    *radi = lerpf(na, a[3], b[3]);
}

// RectangleLight::intersect_ray (lights/rectangle_light.hpp:110-147)
__device__ __forceinline__ bool rect_t(const float* d, int ts, float time, V3 o, V3 dir, V3 dinv, float max_t,
                                       float* t_out) {
    if (dir.z == 0.0f) return false;
    float wd, ht;
    float na = 0.0f;
    const int i = lerp_seq_index(time, ts, &na);
    if (i == -1 || i == -2) {
        const float* e = d + (i == -1 ? 0 : (ts - 1) * 5);
        wd = e[0]; ht = e[1];
    } else {
        const float f = lerp_fac(na);
        wd = (d[i * 5] * f) + (d[(i + 1) * 5] * na);
        ht = (d[i * 5 + 1] * f) + (d[(i + 1) * 5 + 1] * na);
    }
    const float t = -o.z * dinv.z;
    if (t <= 0.0f || t > max_t) return false;
    const float x = o.x + (dir.x * t);
    const float y = o.y + (dir.y * t);
    // the y test compares against the width, as the reference does
    if (x >= (wd * -0.5f) && x <= (wd * 0.5f) && y >= (ht * -0.5f) && y <= (wd * 0.5f)) {
        *t_out = t;
        return true;
    }
    return false;
}

struct TRay {
    V3 o, d, dinv;
    V3 odx, ody, ddx, ddy;  // world differentials (world-ray form only)
    RayW w;                 // footprint in world space
    float time;
    bool occl;
};

template <bool WORLD, bool DEEP>
__device__ __forceinline__ bool intersect_leaf(const DevScene& s, uint32_t li, const TRay& r, Best& b,
                                               bool& need_deep) {
    const DevInstance I = s.inst[li];
    const DevObject ob = s.obj[I.object];
    V3 o = r.o, d = r.d, dinv = r.dinv;
    RayW w = r.w;
    if (I.xf_count > 0) {
        const M44 x = m_lerp_seq(r.time, s.xf + I.xf_offset, I.xf_count);
        o = m_pos(x, r.o);
        d = m_dir(x, r.d);
        dinv = divv(splat(1.0f), d);
        if (WORLD && (ob.kind == K_BILINEAR || ob.kind == K_BICUBIC))
            w = ray_widths(o, d, m_dir(x, r.odx), m_dir(x, r.ody), m_dir(x, r.ddx), m_dir(x, r.ddy));
    }
    const float* od = s.odata + ob.data;
    float t;
    switch (ob.kind) {
        case K_BILINEAR:
        case K_BICUBIC:
            return dice_walk<DEEP>(s, ob, li, o, dinv, w, r.occl, b, need_deep);
        case K_SPHERE: {
            V3 c; float rad;
            sphere_at(od, ob.ts, 4, r.time, &c, &rad);
            if (sphere_t(o, d, b.t, c, rad, &t)) {
                b.t = t; b.inst = li; b.aux = 0.0f; b.u = 0.0f; b.v = 0.0f;
                return r.occl;
            }
            return false;
        }
        case K_SPHERE_LIGHT: {
            V3 c; float rad;
            sphere_at(od, ob.ts, 7, r.time, &c, &rad);
            if (sphere_t(o, d, b.t, c, rad, &t)) {
                b.t = t; b.inst = li; b.aux = 0.0f; b.u = 0.0f; b.v = 0.0f;
                return r.occl;
            }
            return false;
        }
        case K_RECT_LIGHT:
            if (rect_t(od, ob.ts, r.time, o, d, dinv, b.t, &t)) {
                b.t = t; b.inst = li; b.aux = 0.0f; b.u = 0.0f; b.v = 0.0f;
                return r.occl;
            }
            return false;
        default:
            return false;
    }
}

// Latitude / longitude parameters of a sphere hit, computed once for the
// final hit (object/sphere.cpp:131-150).
__device__ __forceinline__ void sphere_uv(const DevScene& s, Best& b, const TRay& r) {
    const DevInstance I = s.inst[b.inst];
    if (I.kind != K_SPHERE) return;
    const DevObject ob = s.obj[I.object];
    V3 o = r.o, d = r.d;
    if (I.xf_count > 0) {
        const M44 x = m_lerp_seq(r.time, s.xf + I.xf_offset, I.xf_count);
        o = m_pos(x, r.o);
        d = m_dir(x, r.d);
    }
    V3 c; float rad;
    sphere_at(s.odata + ob.data, ob.ts, 4, r.time, &c, &rad);
    const V3 p = add(o, scale(d, b.t));
    const V3 unit_p = normalized(sub(p, c));
    const float lat_sin = sqrtf((unit_p.x * unit_p.x) + (unit_p.y * unit_p.y));
    const float long_cos = unit_p.x / lat_sin;
    const float latitude = acosf(unit_p.z);
    float longitude = 0.0f;
    if (unit_p.x != 0.0f || unit_p.y != 0.0f) {
        longitude = acosf(long_cos);
        if (unit_p.y < 0.0f) longitude = (float)((2.0 * 3.1415926535897932384626433) - (double)longitude);
    }
    b.u = longitude / (float)(3.1415926535897932384626433 * 2);
    b.v = (float)((double)latitude / 3.1415926535897932384626433);
}

__device__ __forceinline__ void cmpswap(float& ka, int32_t& ra, float& kb, int32_t& rb) {
    if (kb < ka) {
        const float tk = ka; ka = kb; kb = tk;
        const int32_t tr = ra; ra = rb; rb = tr;
    }
}

constexpr int TRACE_BLOCK = 128;

// A ray of the stream: world form (PSY_WRAY_PLANES, footprint derived here,
// basics/ray.hpp:310-323) or object form (PSY_RAY_PLANES).
template <bool WORLD>
__device__ __forceinline__ void load_ray(const float* __restrict__ rays, int64_t n, int64_t i, uint32_t flags,
                                         TRay& r, float& max_t) {
    if (WORLD) {
        r.o = mk(rays[0 * n + i], rays[1 * n + i], rays[2 * n + i]);
        r.d = mk(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        r.odx = mk(rays[6 * n + i], rays[7 * n + i], rays[8 * n + i]);
        r.ody = mk(rays[9 * n + i], rays[10 * n + i], rays[11 * n + i]);
        r.ddx = mk(rays[12 * n + i], rays[13 * n + i], rays[14 * n + i]);
        r.ddy = mk(rays[15 * n + i], rays[16 * n + i], rays[17 * n + i]);
        r.time = rays[18 * n + i];
        const uint32_t type = __float_as_uint(rays[19 * n + i]);
        r.occl = type == PSY_RAY_OCCLUSION;
        max_t = r.occl ? 1.0f : INFINITY;
        r.w = ray_widths(r.o, r.d, r.odx, r.ody, r.ddx, r.ddy);
    } else {
        r.o = mk(rays[0 * n + i], rays[1 * n + i], rays[2 * n + i]);
        r.d = mk(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        max_t = rays[6 * n + i];
        r.time = rays[7 * n + i];
        r.w.owx = rays[8 * n + i];
        r.w.owy = rays[9 * n + i];
        r.w.dwx = rays[10 * n + i];
        r.w.dwy = rays[11 * n + i];
        r.w.fwx = rays[12 * n + i];
        r.w.fwy = rays[13 * n + i];
        r.occl = (__float_as_uint(rays[14 * n + i]) & (1u << 30)) != 0;
        r.odx = r.ody = r.ddx = r.ddy = splat(0.0f);
    }
    if (flags & PSY_TRACE_OCCLUSION_ONLY) r.occl = true;
    r.dinv = divv(splat(1.0f), r.d);
}


// One ray: BVH4 nearest-first with an LDS stack, leaves through
// intersect_leaf.  Returns true if the ray needs the below-cache pass.
template <bool WORLD, int STACK, bool DEEP>
__device__ __forceinline__ bool trace_one(const DevScene& s, const float* __restrict__ rays, int64_t n, int64_t i,
                                          const DevHits& h, uint32_t flags, int32_t (*stk_ref)[TRACE_BLOCK],
                                          float (*stk_near)[TRACE_BLOCK], int tid) {
    TRay r;
    float max_t;
    load_ray<WORLD>(rays, n, i, flags, r, max_t);

    Best b;
    b.t = max_t;
    b.inst = 0xffffffffu;
    b.u = b.v = b.aux = 0.0f;
    bool need_deep = false;

    int sp = 0;
    if (s.root != CHILD_EMPTY) {
        stk_ref[0][tid] = s.root;
        stk_near[0][tid] = 0.0f;
        sp = 1;
    }
    const bool ds0 = r.dinv.x < 0.0f, ds1 = r.dinv.y < 0.0f, ds2 = r.dinv.z < 0.0f;
    while (sp > 0) {
        --sp;
        const int32_t ref = stk_ref[sp][tid];
        const float nearv = stk_near[sp][tid];
        if (!(nearv < b.t)) continue;
        if (ref >= 0) {
            const float4* q = reinterpret_cast<const float4*>(s.bvh + ref);
            const float4 lx = q[0], hx = q[1], ly = q[2], hy = q[3], lz = q[4], hz = q[5];
            const int4 ch = reinterpret_cast<const int4*>(s.bvh + ref)[6];
            // BBox4::intersect_ray lanes (basics/bbox.hpp:566-595); keys are
            // the near distances of hit children, +inf otherwise
#define PSY_LANE(L, KEY)                                                                             \
    float KEY;                                                                                       \
    {                                                                                                \
        const float xlo = ((ds0 ? hx.L : lx.L) - r.o.x) * r.dinv.x;                                  \
        const float xhi = ((ds0 ? lx.L : hx.L) - r.o.x) * r.dinv.x;                                  \
        const float ylo = ((ds1 ? hy.L : ly.L) - r.o.y) * r.dinv.y;                                  \
        const float yhi = ((ds1 ? ly.L : hy.L) - r.o.y) * r.dinv.y;                                  \
        const float zlo = ((ds2 ? hz.L : lz.L) - r.o.z) * r.dinv.z;                                  \
        const float zhi = ((ds2 ? lz.L : hz.L) - r.o.z) * r.dinv.z;                                  \
        const float mins = fmaxf(fmaxf(xlo, ylo), fmaxf(zlo, 0.0f));                                 \
        const float maxs = fminf(fminf(xhi, yhi), zhi) * PSY_BBOX_MAXT_ADJUST;                       \
        KEY = ((mins < b.t) && (mins <= maxs)) ? mins : INFINITY;                                    \
    }
            PSY_LANE(x, k0)
            PSY_LANE(y, k1)
            PSY_LANE(z, k2)
            PSY_LANE(w, k3)
#undef PSY_LANE
            int32_t r0 = ch.x, r1 = ch.y, r2 = ch.z, r3 = ch.w;
            // sorting network, ascending
            cmpswap(k0, r0, k1, r1);
            cmpswap(k2, r2, k3, r3);
            cmpswap(k0, r0, k2, r2);
            cmpswap(k1, r1, k3, r3);
            cmpswap(k1, r1, k2, r2);
            // push farthest first so the nearest is popped next
            if (k3 < INFINITY) { stk_ref[sp][tid] = r3; stk_near[sp][tid] = k3; ++sp; }
            if (k2 < INFINITY) { stk_ref[sp][tid] = r2; stk_near[sp][tid] = k2; ++sp; }
            if (k1 < INFINITY) { stk_ref[sp][tid] = r1; stk_near[sp][tid] = k1; ++sp; }
            if (k0 < INFINITY) { stk_ref[sp][tid] = r0; stk_near[sp][tid] = k0; ++sp; }
        } else {
            if (intersect_leaf<WORLD, DEEP>(s, (uint32_t)(~ref), r, b, need_deep)) break;
        }
    }
    const bool hit = b.inst != 0xffffffffu;
    if (hit && !r.occl) sphere_uv(s, b, r);
    h.t[i] = hit ? b.t : INFINITY;
    h.inst[i] = b.inst;
    h.u[i] = b.u;
    h.v[i] = b.v;
    if (h.aux) h.aux[i] = b.aux;
    return need_deep && !(r.occl && hit);
}

// ---------------------------------------------------------------------------
// Main pass.  A lane's traversal state lives in registers (plus its LDS
// stack) and one lane_step does one unit of work -- one BVH4 node, or one
// step of a patch's split-tree walk (expand a node's two children, or
// resume at a pending sibling) -- so the 64 lanes of a wave advance
// independently instead of waiting at every BVH leaf for the longest
// split-tree walk in the wave.  Rays that reach the bottom of the dice cache
// with a node still to split, or whose BVH stack would outgrow the main
// pass's MAIN_STACK entries, are handed to the second pass, which retraces
// them with a stack sized for the scene's worst case.
// ---------------------------------------------------------------------------
constexpr int MAIN_STACK = 16;

struct Lane {
    TRay r;
    Best b;
    int64_t idx;
    // split-tree walk (valid while `walking`): patch-space ray, heap node k
    // at `depth`, one `trail` bit per depth whose other child is pending
    const float4* nodes;
    V3 wo, wdinv;
    RayW ww;
    uint32_t winst, k, trail;
    int levels, depth;
    bool walking, expand;
    // BVH: the nearest hit child is visited next without a stack round trip
    int32_t next;
    bool has_next;
    int sp;
    bool need_deep, retrace;
};

template <bool WORLD>
__device__ __forceinline__ void lane_init(const DevScene& s, const float* __restrict__ rays, int64_t n, int64_t i,
                                          uint32_t flags, Lane& L) {
    float max_t;
    load_ray<WORLD>(rays, n, i, flags, L.r, max_t);
    L.b.t = max_t;
    L.b.inst = 0xffffffffu;
    L.b.u = L.b.v = L.b.aux = 0.0f;
    L.idx = i;
    L.walking = false;
    L.expand = true;
    L.nodes = nullptr;
    L.levels = 0;
    L.winst = 0;
    L.k = 1;
    L.trail = 0;
    L.depth = 0;
    L.next = s.root;
    L.has_next = s.root != CHILD_EMPTY;
    L.sp = 0;
    L.need_deep = false;
    L.retrace = false;
}

// One unit of work; true when the ray is finished (or must be retraced).
template <bool WORLD>
__device__ __forceinline__ bool lane_step(const DevScene& s, Lane& L, int32_t (*stk_ref)[TRACE_BLOCK],
                                          float (*stk_near)[TRACE_BLOCK], int tid) {
    if (L.walking) {
        if (L.expand) {
            const uint32_t c0 = 2 * L.k;
            const float4* q = L.nodes + 2 * c0;
            const float4 a0 = q[0], e0 = q[1], a1 = q[2], e1 = q[3];
            float h00, h01, h10, h11;
            const bool hit0 = bbox_ray(xyz(a0), xyz(e0), L.wo, L.wdinv, L.b.t, &h00, &h01);
            const bool hit1 = bbox_ray(xyz(a1), xyz(e1), L.wo, L.wdinv, L.b.t, &h10, &h11);
            bool split0 = false, split1 = false;
            // the second half first: the reference visits it first on ties
            if (hit1) {
                const float maxd = longest_axis(sub(xyz(e1), xyz(a1)));
                if (maxd <= smax(min_width(L.ww, h10, h11) * s.dice_rate, s.min_upoly)) {
                    if (accept_leaf(h10, h11, a1.w, e1.w, maxd, L.winst, L.b) && L.r.occl) return true;
                } else {
                    split1 = true;
                }
            }
            if (hit0) {
                const float maxd = longest_axis(sub(xyz(e0), xyz(a0)));
                if (maxd <= smax(min_width(L.ww, h00, h01) * s.dice_rate, s.min_upoly)) {
                    if (accept_leaf(h00, h01, a0.w, e0.w, maxd, L.winst, L.b) && L.r.occl) return true;
                } else {
                    split0 = true;
                }
            }
            split0 = split0 && h00 <= L.b.t;
            split1 = split1 && h10 <= L.b.t;
            const int cd = L.depth + 1;
            if (cd == L.levels) {
                if (split0 || split1) L.need_deep = true;
                split0 = split1 = false;
            }
            if (split0 && split1) {
                L.k = (h00 < h10) ? c0 : c0 + 1;
                L.trail |= 1u << cd;
                L.depth = cd;
            } else if (split0 || split1) {
                L.k = split0 ? c0 : c0 + 1;
                L.depth = cd;
            } else {
                L.expand = false;
            }
            return false;
        }
        if (L.trail == 0) {
            L.walking = false;
            return false;
        }
        // resume at the deepest pending sibling
        const int dd = 31 - __builtin_clz(L.trail);
        L.k = (L.k >> (L.depth - dd)) ^ 1u;
        L.depth = dd;
        L.trail &= ~(1u << dd);
        const float4 a = L.nodes[2 * L.k], c = L.nodes[2 * L.k + 1];
        float h0, h1;
        if (bbox_ray(xyz(a), xyz(c), L.wo, L.wdinv, L.b.t, &h0, &h1)) {
            const float maxd = longest_axis(sub(xyz(c), xyz(a)));
            if (maxd <= smax(min_width(L.ww, h0, h1) * s.dice_rate, s.min_upoly)) {
                if (accept_leaf(h0, h1, a.w, c.w, maxd, L.winst, L.b) && L.r.occl) return true;
            } else {
                L.expand = true;
            }
        }
        return false;
    }

    int32_t ref;
    if (L.has_next) {
        ref = L.next;
        L.has_next = false;
    } else {
        if (L.sp == 0) return true;
        --L.sp;
        ref = stk_ref[L.sp][tid];
        if (!(stk_near[L.sp][tid] < L.b.t)) return false;
    }
    if (ref >= 0) {
        const float4* q = reinterpret_cast<const float4*>(s.bvh + ref);
        const float4 lx = q[0], hx = q[1], ly = q[2], hy = q[3], lz = q[4], hz = q[5];
        const int4 ch = reinterpret_cast<const int4*>(s.bvh + ref)[6];
        const TRay& r = L.r;
        const bool ds0 = r.dinv.x < 0.0f, ds1 = r.dinv.y < 0.0f, ds2 = r.dinv.z < 0.0f;
        // BBox4::intersect_ray lanes (basics/bbox.hpp:566-595); keys are the
        // near distances of hit children, +inf otherwise
#define PSY_LANE(LN, KEY)                                                                            \
    float KEY;                                                                                       \
    {                                                                                                \
        const float xlo = ((ds0 ? hx.LN : lx.LN) - r.o.x) * r.dinv.x;                                \
        const float xhi = ((ds0 ? lx.LN : hx.LN) - r.o.x) * r.dinv.x;                                \
        const float ylo = ((ds1 ? hy.LN : ly.LN) - r.o.y) * r.dinv.y;                                \
        const float yhi = ((ds1 ? ly.LN : hy.LN) - r.o.y) * r.dinv.y;                                \
        const float zlo = ((ds2 ? hz.LN : lz.LN) - r.o.z) * r.dinv.z;                                \
        const float zhi = ((ds2 ? lz.LN : hz.LN) - r.o.z) * r.dinv.z;                                \
        const float mins = fmaxf(fmaxf(xlo, ylo), fmaxf(zlo, 0.0f));                                 \
        const float maxs = fminf(fminf(xhi, yhi), zhi) * PSY_BBOX_MAXT_ADJUST;                       \
        KEY = ((mins < L.b.t) && (mins <= maxs)) ? mins : INFINITY;                                  \
    }
        PSY_LANE(x, k0)
        PSY_LANE(y, k1)
        PSY_LANE(z, k2)
        PSY_LANE(w, k3)
#undef PSY_LANE
        int32_t r0 = ch.x, r1 = ch.y, r2 = ch.z, r3 = ch.w;
        cmpswap(k0, r0, k1, r1);
        cmpswap(k2, r2, k3, r3);
        cmpswap(k0, r0, k2, r2);
        cmpswap(k1, r1, k3, r3);
        cmpswap(k1, r1, k2, r2);
        const int npush = (k1 < INFINITY ? 1 : 0) + (k2 < INFINITY ? 1 : 0) + (k3 < INFINITY ? 1 : 0);
        if (L.sp + npush > MAIN_STACK) {
            L.retrace = true;
            return true;
        }
        if (k3 < INFINITY) { stk_ref[L.sp][tid] = r3; stk_near[L.sp][tid] = k3; ++L.sp; }
        if (k2 < INFINITY) { stk_ref[L.sp][tid] = r2; stk_near[L.sp][tid] = k2; ++L.sp; }
        if (k1 < INFINITY) { stk_ref[L.sp][tid] = r1; stk_near[L.sp][tid] = k1; ++L.sp; }
        if (k0 < INFINITY) { L.next = r0; L.has_next = true; }
        return false;
    }
    // leaf instance
    const uint32_t li = (uint32_t)(~ref);
    const DevInstance I = s.inst[li];
    const DevObject ob = s.obj[I.object];
    if (ob.kind == K_BILINEAR || ob.kind == K_BICUBIC) {
        L.wo = L.r.o;
        L.wdinv = L.r.dinv;
        L.ww = L.r.w;
        if (I.xf_count > 0) {
            const M44 x = m_lerp_seq(L.r.time, s.xf + I.xf_offset, I.xf_count);
            L.wo = m_pos(x, L.r.o);
            const V3 d = m_dir(x, L.r.d);
            L.wdinv = divv(splat(1.0f), d);
            if (WORLD)
                L.ww = ray_widths(L.wo, d, m_dir(x, L.r.odx), m_dir(x, L.r.ody), m_dir(x, L.r.ddx),
                                  m_dir(x, L.r.ddy));
        }
        L.nodes = s.dice + 2 * ob.dice;
        L.levels = ob.dice_levels;
        const float4 a = L.nodes[2], c = L.nodes[3];
        float h0, h1;
        if (!bbox_ray(xyz(a), xyz(c), L.wo, L.wdinv, L.b.t, &h0, &h1)) return false;
        const float maxd = longest_axis(sub(xyz(c), xyz(a)));
        if (maxd <= smax(min_width(L.ww, h0, h1) * s.dice_rate, s.min_upoly))
            return accept_leaf(h0, h1, a.w, c.w, maxd, li, L.b) && L.r.occl;
        if (L.levels == 0) {
            L.need_deep = true;
            return false;
        }
        L.walking = true;
        L.expand = true;
        L.winst = li;
        L.k = 1;
        L.depth = 0;
        L.trail = 0;
        return false;
    }
    bool unused = false;
    return intersect_leaf<WORLD, false>(s, li, L.r, L.b, unused);
}

// Writes the lane's hit; true if the ray goes to the second pass.
__device__ __forceinline__ bool lane_finish(const DevScene& s, Lane& L, const DevHits& h) {
    const bool hit = L.b.inst != 0xffffffffu;
    if (hit && !L.r.occl) sphere_uv(s, L.b, L.r);
    const int64_t i = L.idx;
    h.t[i] = hit ? L.b.t : INFINITY;
    h.inst[i] = L.b.inst;
    h.u[i] = L.b.u;
    h.v[i] = L.b.v;
    if (h.aux) h.aux[i] = L.b.aux;
    return (L.need_deep || L.retrace) && !(L.r.occl && hit);
}

__device__ __forceinline__ void enqueue_retrace(uint32_t* __restrict__ queue, uint32_t* __restrict__ counters,
                                                int64_t i) {
    const uint32_t slot = atomicAdd(&counters[0], 1u);
    queue[slot] = (uint32_t)i;
}

// Main pass, one thread per ray.
template <bool WORLD>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace(DevScene s, const float* __restrict__ rays, int64_t n,
                                                        DevHits h, uint32_t flags, uint32_t* __restrict__ queue,
                                                        uint32_t* __restrict__ counters) {
    __shared__ int32_t stk_ref[MAIN_STACK][TRACE_BLOCK];
    __shared__ float stk_near[MAIN_STACK][TRACE_BLOCK];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * TRACE_BLOCK + tid;
    if (i >= n) return;
    Lane L;
    lane_init<WORLD>(s, rays, n, i, flags, L);
    while (!lane_step<WORLD>(s, L, stk_ref, stk_near, tid)) {
    }
    if (lane_finish(s, L, h)) enqueue_retrace(queue, counters, i);
}

// Main pass, persistent: a grid that fills the GPU once; every wave takes
// rays from a global counter, and refills its finished lanes (in one
// atomic per wave) whenever at least `refill_min` of them are idle, so the
// wave's lanes stay busy through the long tail of incoherent rays.  Every
// wave leaves once the counter has passed n and its lanes are idle.
template <bool WORLD>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_persist(DevScene s, const float* __restrict__ rays,
                                                                int64_t n, DevHits h, uint32_t flags,
                                                                uint32_t* __restrict__ queue,
                                                                uint32_t* __restrict__ counters, int refill_min) {
    __shared__ int32_t stk_ref[MAIN_STACK][TRACE_BLOCK];
    __shared__ float stk_near[MAIN_STACK][TRACE_BLOCK];
    const int tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    unsigned long long* next_ray = reinterpret_cast<unsigned long long*>(counters + 2);
    Lane L;
    bool active = false;
    bool exhausted = false;  // wave-uniform
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (!exhausted && nidle >= refill_min) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(next_ray, (unsigned long long)nidle);
            base = ((unsigned long long)(uint32_t)__shfl((int)(uint32_t)(base >> 32), 0) << 32) |
                   (unsigned long long)(uint32_t)__shfl((int)(uint32_t)base, 0);
            if ((int64_t)base >= n) exhausted = true;
            if (!active) {
                const int64_t i = (int64_t)base + __popcll(idle & ((1ull << lane) - 1ull));
                if (i < n) {
                    lane_init<WORLD>(s, rays, n, i, flags, L);
                    active = true;
                }
            }
        }
        if (exhausted && __ballot(active) == 0ull) break;
        if (active && lane_step<WORLD>(s, L, stk_ref, stk_near, tid)) {
            active = false;
            if (lane_finish(s, L, h)) enqueue_retrace(queue, counters, L.idx);
        }
    }
}

// Second pass: retrace the queued rays with subdivision below the cache and
// a stack of STACK >= the scene's worst case.
template <bool WORLD, int STACK>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_deep(DevScene s, const float* __restrict__ rays, int64_t n,
                                                             DevHits h, uint32_t flags,
                                                             const uint32_t* __restrict__ queue,
                                                             const uint32_t* __restrict__ counters) {
    __shared__ int32_t stk_ref[STACK][TRACE_BLOCK];
    __shared__ float stk_near[STACK][TRACE_BLOCK];
    const int tid = threadIdx.x;
    const uint32_t count = counters[0];
    for (uint32_t k = blockIdx.x * TRACE_BLOCK + tid; k < count; k += gridDim.x * TRACE_BLOCK)
        (void)trace_one<WORLD, STACK, true>(s, rays, n, (int64_t)queue[k], h, flags, stk_ref, stk_near, tid);
}

// counters[0]: retrace queue length; counters[2..3]: next ray (u64)
__global__ void k_zero_counters(uint32_t* p) {
    if (threadIdx.x < 4 && blockIdx.x == 0) p[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------------------
// WorldRay::to_ray over a stream (basics/ray.hpp:207-252, :310-323)
// ---------------------------------------------------------------------------
__global__ void k_world_to_rays(const float* __restrict__ w, int64_t n, float* __restrict__ r) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 o = mk(w[0 * n + i], w[1 * n + i], w[2 * n + i]);
    const V3 d = mk(w[3 * n + i], w[4 * n + i], w[5 * n + i]);
    const V3 odx = mk(w[6 * n + i], w[7 * n + i], w[8 * n + i]);
    const V3 ody = mk(w[9 * n + i], w[10 * n + i], w[11 * n + i]);
    const V3 ddx = mk(w[12 * n + i], w[13 * n + i], w[14 * n + i]);
    const V3 ddy = mk(w[15 * n + i], w[16 * n + i], w[17 * n + i]);
    const float time = w[18 * n + i];
    const uint32_t type = __float_as_uint(w[19 * n + i]);
    const bool occl = type == PSY_RAY_OCCLUSION;
    const RayW rw = ray_widths(o, d, odx, ody, ddx, ddy);
    r[0 * n + i] = o.x; r[1 * n + i] = o.y; r[2 * n + i] = o.z;
    r[3 * n + i] = d.x; r[4 * n + i] = d.y; r[5 * n + i] = d.z;
    r[6 * n + i] = occl ? 1.0f : INFINITY;
    r[7 * n + i] = time;
    r[8 * n + i] = rw.owx; r[9 * n + i] = rw.owy;
    r[10 * n + i] = rw.dwx; r[11 * n + i] = rw.dwy;
    r[12 * n + i] = rw.fwx; r[13 * n + i] = rw.fwy;
    r[14 * n + i] = __uint_as_float(occl ? (1u << 30) : 0u);
}

// ---------------------------------------------------------------------------
// Camera rays: init_path + next_ray_for_path step 0
// (integrator/path_trace_integrator.cpp:91-125, basics/camera.hpp:72-104,
// utils/monte_carlo.hpp:17-53, utils/utils.hpp:335-343)
// ---------------------------------------------------------------------------
#define PSY_QPI (3.1415926535897932384626433 / 4)

__device__ __forceinline__ float faster_ln(float x) {
    float y = (float)__float_as_uint(x);
    y *= 8.2629582881927490e-8f;
    return y - 87.989971088f;
}
__device__ __forceinline__ float fast_logit(float p, float width) {
    const float n = 0.001f + (p * 0.998f);
    return faster_ln((n / (1.0f - n))) * width * (0.6266f / 4.0f);
}
__device__ __forceinline__ void square_to_circle(float* x, float* y) {
    if (*x == 0.0f && *y == 0.0f) return;
    float radius, angle;
    if (*x > fabsf(*y)) {
        radius = *x;
        angle = (float)(PSY_QPI * (double)(*y / *x));
    } else if (*y > fabsf(*x)) {
        radius = *y;
        angle = (float)(PSY_QPI * (double)(2 - (*x / *y)));
    } else if (*x < -fabsf(*y)) {
        radius = -*x;
        angle = (float)(PSY_QPI * (double)(4 + (*y / *x)));
    } else {
        radius = -*y;
        angle = (float)(PSY_QPI * (double)(6 - (*x / *y)));
    }
    *x = radius * cosf(angle);
    *y = radius * sinf(angle);
}

__device__ __forceinline__ float lerp_seq_f(float alpha, const float* seq, int n) {
    float na = 0.0f;
    const int i = lerp_seq_index(alpha, n, &na);
    if (i == -1) return seq[0];
    if (i == -2) return seq[n - 1];
    return lerpf(na, seq[i], seq[i + 1]);
}

__device__ __forceinline__ void store_wray(float* w, int64_t n, int64_t i, V3 o, V3 d, V3 odx, V3 ody, V3 ddx,
                                           V3 ddy, float time, uint32_t type) {
    w[0 * n + i] = o.x; w[1 * n + i] = o.y; w[2 * n + i] = o.z;
    w[3 * n + i] = d.x; w[4 * n + i] = d.y; w[5 * n + i] = d.z;
    w[6 * n + i] = odx.x; w[7 * n + i] = odx.y; w[8 * n + i] = odx.z;
    w[9 * n + i] = ody.x; w[10 * n + i] = ody.y; w[11 * n + i] = ody.z;
    w[12 * n + i] = ddx.x; w[13 * n + i] = ddx.y; w[14 * n + i] = ddx.z;
    w[15 * n + i] = ddy.x; w[16 * n + i] = ddy.y; w[17 * n + i] = ddy.z;
    w[18 * n + i] = time;
    w[19 * n + i] = __uint_as_float(type);
}

__global__ void k_camera_rays(DevCamera cam, const HaltonDim* dims, const uint16_t* perm, int width, int height,
                              int spp, uint32_t seed, float* wrays, float* wavelength) {
    const int64_t n = (int64_t)width * height * spp;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = (int)(i % spp);
    const int64_t pix = i / spp;
    const int px = (int)(pix % width);
    const int py = (int)(pix / width);
    const uint32_t off = sampler_offset((uint32_t)px, (uint32_t)py, (uint32_t)s, seed);
    const float min_x = -1.0f, max_x = 1.0f;
    const float max_y = ((float)height) / (float)width;
    const float min_y = -max_y;
    uint32_t dim = 0;
    const float wl = (sampler_get(dims, perm, off, dim++) * (700.0f - 380.0f)) + 380.0f;
    const float samp_x = (fast_logit(sampler_get(dims, perm, off, dim++), 1.5f) + 0.5f + (float)(short)px) / (float)width;
    const float samp_y = (fast_logit(sampler_get(dims, perm, off, dim++), 1.5f) + 0.5f + (float)(short)py) / (float)height;
    const float rx = (float)(((double)samp_x - 0.5) * (double)(max_x - min_x));
    const float ry = (float)((0.5 - (double)samp_y) * (double)(max_y - min_y));
    const float dx = (max_x - min_x) / (float)width;
    const float dy = (max_y - min_y) / (float)height;
    const float sa = sampler_get(dims, perm, off, dim++);
    const float sb = sampler_get(dims, perm, off, dim++);
    const float time = sampler_get(dims, perm, off, dim++);
    // Camera::generate_ray(rx, ry, dx, dy, time, sa, sb)
    const M44 xf = m_lerp_seq(time, cam.xf, cam.n_xf);
    const float tfov = lerp_seq_f(time, cam.tfov, cam.n_tfov);
    const float ap = lerp_seq_f(time, cam.aperture, cam.n_ap);
    const float fd = lerp_seq_f(time, cam.focus, cam.n_fd);
    float ox = ap * ((sa * 2) - 1);
    float oy = ap * ((sb * 2) - 1);
    square_to_circle(&ox, &oy);
    const V3 o = mk(ox, oy, 0.0f);
    V3 d = mk((rx * tfov) - (o.x / fd), (ry * tfov) - (o.y / fd), 1.0f);
    d = normalized(d);
    const V3 ddx = mk(dx * tfov, 0.0f, 0.0f);
    const V3 ddy = mk(0.0f, dy * tfov, 0.0f);
    store_wray(wrays, n, i, m_pos(xf, o), m_dir(xf, d), m_dir(xf, splat(0.0f)), m_dir(xf, splat(0.0f)),
               m_dir(xf, ddx), m_dir(xf, ddy), time, 1u /* WorldRay::CAMERA */);
    if (wavelength) wavelength[i] = wl;
}

// ---------------------------------------------------------------------------
// Lambert bounce rays from closest hits
// (integrator/path_trace_integrator.cpp:183-209,
//  shading/surface_closure.hpp:240-326, utils/utils.hpp:214-248,
//  basics/differential_geometry.hpp:20-40, :83-98)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void coordinate_system_from_vec3(V3 v1, V3* v2, V3* v3o) {
    if (fabsf(v1.x) > fabsf(v1.y)) {
        const float invlen = 1.f / sqrtf(v1.x * v1.x + v1.z * v1.z);
        *v2 = mk(-v1.z * invlen, 0.f, v1.x * invlen);
    } else {
        const float invlen = 1.f / sqrtf(v1.y * v1.y + v1.z * v1.z);
        *v2 = mk(0.f, v1.z * invlen, -v1.y * invlen);
    }
    *v3o = cross(v1, *v2);
}
__device__ __forceinline__ V3 zup_to_vec(V3 from, V3 toz) {
    toz = normalized(toz);
    V3 tox, toy;
    coordinate_system_from_vec3(toz, &tox, &toy);
    return add(add(scale(tox, from.x), scale(toy, from.y)), scale(toz, from.z));
}
__device__ __forceinline__ V3 transfer_origin_diff(float t, V3 normal, V3 d, V3 od, V3 dd) {
    const V3 temp = add(od, scale(dd, t));
    const float td = -dot(temp, normal) / dot(d, normal);
    const V3 real_projected = add(temp, scale(d, td));
    return scale(normalized(real_projected), length(temp));
}

// Object-space hit geometry (n, offset) of a closest hit; false for lights
// and misses.
__device__ bool hit_geometry(const DevScene& s, uint32_t li, float t, float u, float v, float aux, V3 wo, V3 wd,
                             float time, M44* space, V3* p_local, V3* n_local, V3* offset, bool* backfacing) {
    const DevInstance I = s.inst[li];
    const DevObject ob = s.obj[I.object];
    V3 o = wo, d = wd;
    *space = m_identity();
    if (I.xf_count > 0) {
        *space = m_lerp_seq(time, s.xf + I.xf_offset, I.xf_count);
        o = m_pos(*space, wo);
        d = m_dir(*space, wd);
    }
    const float* od = s.odata + ob.data;
    const V3 p = add(o, scale(d, t));
    V3 n;
    if (ob.kind == K_BILINEAR || ob.kind == K_BICUBIC) {
        V3 dpdu, dpdv, dndu, dndv;
        if (ob.kind == K_BILINEAR) {
            Patch<4> pp;
            load_patch<4>(s.odata, ob.data, pp);
            patch_dg<4>(pp, u, v, &n, &dpdu, &dpdv, &dndu, &dndv);
        } else {
            Patch<16> pp;
            load_patch<16>(s.odata, ob.data, pp);
            patch_dg<16>(pp, u, v, &n, &dpdu, &dpdv, &dndu, &dndv);
        }
        *offset = scale(n, aux * 1.74f);
    } else if (ob.kind == K_SPHERE) {
        V3 c; float rad;
        sphere_at(od, ob.ts, 4, time, &c, &rad);
        n = normalized(sub(p, c));
        *offset = scale(n, 0.000001f);
    } else {
        return false;
    }
    *backfacing = dot(n, normalized(d)) > 0.0f;
    *p_local = p;
    *n_local = n;
    return true;
}

__global__ void k_bounce_rays(DevScene s, const HaltonDim* dims, const uint16_t* perm, const float* __restrict__ in,
                              DevHits hits, int64_t n, int width, int spp, uint32_t seed, uint32_t dim_u,
                              uint32_t dim_v, float* out, uint32_t* valid) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t li = hits.inst[i];
    const V3 zero = splat(0.0f);
    bool ok = li != 0xffffffffu;
    int shader = -1;
    if (ok) {
        shader = s.inst[li].shader;
        ok = shader >= 0 && s.shaders[shader].kind == PSY_SHADER_LAMBERT;
    }
    const V3 wo = mk(in[0 * n + i], in[1 * n + i], in[2 * n + i]);
    const V3 wd = mk(in[3 * n + i], in[4 * n + i], in[5 * n + i]);
    const float time = in[18 * n + i];
    M44 space;
    V3 pl, nl, offl;
    bool backfacing = false;
    if (ok) ok = hit_geometry(s, li, hits.t[i], hits.u[i], hits.v[i], hits.aux ? hits.aux[i] : 0.0f, wo, wd,
                              time, &space, &pl, &nl, &offl, &backfacing);
    if (!ok) {
        valid[i] = 0;
        store_wray(out, n, i, zero, zero, zero, zero, zero, zero, time, 2u);
        return;
    }
    valid[i] = 1;
    // DifferentialGeometry::transformed_from(space)
    const M44 inv = m_inverse(space);
    const V3 gp = m_pos(inv, pl);
    V3 gn = m_dir_transpose(space, nl);
    const float il = 1.0f / length(gn);
    gn = scale(gn, il);
    V3 pos_offset = m_dir(inv, offl);
    if (backfacing) pos_offset = scale(pos_offset, -1.0f);
    // LambertClosure::sample
    V3 nn = normalized(gn);
    if (dot(nn, wd) > 0.0f) nn = scale(nn, -1.0f);
    const int sidx = (int)(i % spp);
    const int64_t pix = i / spp;
    const uint32_t off = sampler_offset((uint32_t)(pix % width), (uint32_t)(pix / width), (uint32_t)sidx, seed);
    float su = sampler_get(dims, perm, off, dim_u);
    float sv = sampler_get(dims, perm, off, dim_v);
    su = (su * 2) - 1;
    sv = (sv * 2) - 1;
    square_to_circle(&su, &sv);
    const double zz = 1.0 - (double)((su * su) + (sv * sv));
    const float z = (float)sqrt(zz > 0.0 ? zz : 0.0);
    const V3 outd = zup_to_vec(mk(su, sv, z), nn);
    // LambertClosure::propagate_differentials
    const float len = length(outd);
    const V3 gnn = normalized(gn);
    V3 x, y;
    coordinate_system_from_vec3(outd, &x, &y);
    x = normalized(x);
    y = normalized(y);
    const V3 podx = mk(in[6 * n + i], in[7 * n + i], in[8 * n + i]);
    const V3 pody = mk(in[9 * n + i], in[10 * n + i], in[11 * n + i]);
    const V3 pddx = mk(in[12 * n + i], in[13 * n + i], in[14 * n + i]);
    const V3 pddy = mk(in[15 * n + i], in[16 * n + i], in[17 * n + i]);
    const float t = hits.t[i];
    store_wray(out, n, i, add(gp, pos_offset), outd, transfer_origin_diff(t, gnn, wd, podx, pddx),
               transfer_origin_diff(t, gnn, wd, pody, pddy), sdiv(scale(x, 0.15f), len),
               sdiv(scale(y, 0.15f), len), time, 2u /* WorldRay::R_DIFFUSE */);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

static int check_launch() { return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP; }

int launch_build_dice(const float* odata, const DevObject* objs, const int32_t* patch_ids, int n_patches,
                      int levels, float4* dice, hipStream_t st) {
    const int64_t per = (int64_t)1 << (levels + 1);
    const int64_t total = per * n_patches;
    if (total <= 0) return 0;
    hipLaunchKernelGGL(k_build_dice, dim3(grid_for(total, 256)), dim3(256), 0, st, odata, objs, patch_ids,
                       n_patches, levels, dice);
    return check_launch();
}

template <bool W, int S>
static void launch_deep(const DevScene& s, const float* rays, int64_t n, const DevHits& h, uint32_t flags,
                        const uint32_t* queue, const uint32_t* counters, hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<int64_t>(grid_for(n, TRACE_BLOCK), 1024);
    hipLaunchKernelGGL((k_trace_deep<W, S>), dim3(blocks), dim3(TRACE_BLOCK), 0, st, s, rays, n, h, flags, queue,
                       counters);
}

template <bool W>
static int launch_trace_w(const DevScene& s, int stack_need, const float* rays, int64_t n, const DevHits& h,
                          uint32_t flags, uint32_t* queue, uint32_t* counters, const TraceTuning& tune,
                          hipStream_t st) {
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, counters);
    const int64_t plain_blocks = grid_for(n, TRACE_BLOCK);
    const int64_t persist_blocks = (int64_t)tune.persist_blocks;
    if (tune.persistent && persist_blocks > 0 && persist_blocks < plain_blocks) {
        hipLaunchKernelGGL((k_trace_persist<W>), dim3((unsigned)persist_blocks), dim3(TRACE_BLOCK), 0, st, s, rays,
                           n, h, flags, queue, counters, std::max(1, std::min(64, tune.refill_min)));
    } else {
        hipLaunchKernelGGL((k_trace<W>), dim3((unsigned)plain_blocks), dim3(TRACE_BLOCK), 0, st, s, rays, n, h,
                           flags, queue, counters);
    }
    if (stack_need <= 16) launch_deep<W, 16>(s, rays, n, h, flags, queue, counters, st);
    else if (stack_need <= 32) launch_deep<W, 32>(s, rays, n, h, flags, queue, counters, st);
    else if (stack_need <= 64) launch_deep<W, 64>(s, rays, n, h, flags, queue, counters, st);
    else return PSY_ERR_UNSUPPORTED;
    return check_launch();
}

int launch_trace(const DevScene& s, bool world, bool occl_only, int stack_need, const float* rays, int64_t n,
                 const DevHits& h, uint32_t* queue, uint32_t* counters, const TraceTuning& tune, hipStream_t st) {
    if (n <= 0) return 0;
    if (stack_need > 64) return PSY_ERR_UNSUPPORTED;
    const uint32_t flags = occl_only ? PSY_TRACE_OCCLUSION_ONLY : 0u;
    return world ? launch_trace_w<true>(s, stack_need, rays, n, h, flags, queue, counters, tune, st)
                 : launch_trace_w<false>(s, stack_need, rays, n, h, flags, queue, counters, tune, st);
}

int trace_persist_occupancy(bool world, int* blocks_per_cu) {
    int nb = 0;
    hipError_t e = world ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace_persist<true>, TRACE_BLOCK, 0)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace_persist<false>, TRACE_BLOCK, 0);
    *blocks_per_cu = nb;
    return e == hipSuccess ? 0 : PSY_ERR_HIP;
}

int launch_world_to_rays(const float* wrays, int64_t n, float* rays, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_world_to_rays, dim3(grid_for(n, 256)), dim3(256), 0, st, wrays, n, rays);
    return check_launch();
}

int launch_halton(const HaltonDim* dims, const uint16_t* perm, uint32_t dim, const uint32_t* idx, int64_t n,
                  float* out, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_halton, dim3(grid_for(n, 256)), dim3(256), 0, st, dims, perm, dim, idx, n, out);
    return check_launch();
}

int launch_pixel_samples(const HaltonDim* dims, const uint16_t* perm, int x0, int y0, int w, int h, int spp,
                         uint32_t seed, uint32_t dim0, int n_dims, float* out, hipStream_t st) {
    const int64_t total = (int64_t)w * h * spp;
    if (total <= 0) return 0;
    hipLaunchKernelGGL(k_pixel_samples, dim3(grid_for(total, 256)), dim3(256), 0, st, dims, perm, x0, y0, w, h,
                       spp, seed, dim0, n_dims, out);
    return check_launch();
}

int launch_camera_rays(const DevCamera& cam, const HaltonDim* dims, const uint16_t* perm, int width, int height,
                       int spp, uint32_t seed, float* wrays, float* wavelength, hipStream_t st) {
    const int64_t n = (int64_t)width * height * spp;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_camera_rays, dim3(grid_for(n, 256)), dim3(256), 0, st, cam, dims, perm, width, height,
                       spp, seed, wrays, wavelength);
    return check_launch();
}

int launch_bounce_rays(const DevScene& s, const HaltonDim* dims, const uint16_t* perm, const float* in_wrays,
                       const DevHits& hits, int64_t n, int width, int spp, uint32_t seed, uint32_t dim_u,
                       uint32_t dim_v, float* out_wrays, uint32_t* valid, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_bounce_rays, dim3(grid_for(n, 256)), dim3(256), 0, st, s, dims, perm, in_wrays, hits, n,
                       width, spp, seed, dim_u, dim_v, out_wrays, valid);
    return check_launch();
}

}  // namespace psy
