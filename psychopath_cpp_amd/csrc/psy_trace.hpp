// psy_trace.hpp -- the device functions of the ray tracer: sampling, ray
// set-up, BVH4 + split-dice traversal (main pass and second pass), leaf
// intersection, hit geometry.  psy_kernels.hip wraps them in kernels; they
// are also host-callable when PSY_DEV is defined as host+device before this
// header is included (a CPU build of the same code for tests).
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
// Arithmetic follows the reference's operation order; compile with
// -ffp-contract=off and IEEE division / square root.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "psy_device.hpp"
#include "psy_patch.hpp"
#include "../../include/psy_gpu.h"

#ifndef PSY_DEV
#define PSY_DEV __device__ __forceinline__
#define PSY_DEV_NOINLINE __device__ __noinline__
#define PSY_DEV_FN __device__
#endif

namespace psy {

// ---------------------------------------------------------------------------
// Halton + Sampler (sampling/halton.cpp, sampling/image_sampler.hpp:16-38,
// utils/hash.hpp:6-16)
// ---------------------------------------------------------------------------
PSY_DEV float halton_sample(const HaltonDim* dims, const uint16_t* perm, uint32_t dim,
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

PSY_DEV uint32_t hash_u32(uint32_t n, uint32_t seed) {
    uint32_t hash = n;
    for (int i = 0; i < 3; ++i) {
        hash *= 1936502639u;
        hash ^= hash >> 16;
        hash += seed;
    }
    return hash;
}

PSY_DEV uint32_t sampler_offset(uint32_t x, uint32_t y, uint32_t n, uint32_t seed) {
    return hash_u32(x ^ ((y >> 16) | (y << 16)), seed) + n;
}

PSY_DEV float sampler_get(const HaltonDim* dims, const uint16_t* perm, uint32_t off,
                                             uint32_t dim) {
    // dimension reorder (sampling/image_sampler.hpp:26)
    uint32_t d = dim;
    if (dim < 11) {
        const uint8_t tbl[11] = {10, 7, 6, 5, 4, 2, 9, 8, 3, 1, 0};
        d = tbl[dim];
    }
    return halton_sample(dims, perm, d, off);
}

// Sobol::sample (sampling/sobol.hpp:40-53): the index's bits select matrix
// columns of dimension `dim` (52 columns a dimension; columns past the end
// of the table read as zero), XORed into the 52-bit scramble
PSY_DEV double sobol_value(const uint64_t* m, uint32_t n_cols, uint32_t dim, uint64_t index, uint64_t scramble) {
    uint64_t r = scramble & ((1ull << 52) - 1ull);
    uint32_t c = dim * 52u;
    while (index) {
        if ((index & 1ull) && c < n_cols) r ^= m[c];
        index >>= 1;
        ++c;
    }
    return (double)r * (1.0 / 4503599627370496.0);
}
// The Sobol form of a path's sample (PSY_RENDER_SOBOL): dimension `dim` at
// the Sampler's index, unscrambled, as a float below 1
PSY_DEV float sobol_sampler_get(const uint64_t* m, uint32_t n_cols, uint32_t offset, uint32_t dim) {
    const float v = (float)sobol_value(m, n_cols, dim, (uint64_t)offset, 0ull);
    return v < 1.0f ? v : 0.99999994f;
}



// ---------------------------------------------------------------------------
// Dice cache builder: one lane per split-tree node, replaying the splits
// from the patch root (object/patch_utils.hpp:38-46, :146-192).  Node k of a
// patch (1-based heap) lives at entries 2k, 2k+1 of the patch's slot, so the
// two children of a node form one aligned 64-byte pair.
// ---------------------------------------------------------------------------
template <int N>
PSY_DEV void load_patch(const float* odata, int32_t off, Patch<N>& p) {
    for (int i = 0; i < N; ++i) p.p[i] = mk(odata[off + 3 * i], odata[off + 3 * i + 1], odata[off + 3 * i + 2]);
}

template <int N>
PSY_DEV_FN void build_node(const float* odata, const DevObject& ob, uint32_t k, float4* out) {
    Patch<N> p;
    load_patch<N>(odata, ob.data, p);
    float uv[4];
    patch_at_node<N>(p, k, uv);
    V3 lo, hi;
    patch_bound<N>(p, &lo, &hi);
    out[0] = make_float4(lo.x, lo.y, lo.z, (uv[0] + uv[1]) * 0.5f);
    out[1] = make_float4(hi.x, hi.y, hi.z, (uv[2] + uv[3]) * 0.5f);
}


// ---------------------------------------------------------------------------
// Traversal
// ---------------------------------------------------------------------------
struct Best {
    float t;
    uint32_t inst;
    float u, v, aux;
};

PSY_DEV V3 xyz(float4 a) { return mk(a.x, a.y, a.z); }

// The leaf rule of intersect_rays_with_patch (object/patch_utils.hpp:76-133):
// a hit micro-bbox no larger than the ray footprint (times dice_rate) ends
// the descent; its hit distance is the middle of the slab interval.
PSY_DEV bool accept_leaf(float h0, float h1, float u, float v, float maxd, uint32_t inst, Best& b) {
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

// bbox_ray for a box of a patch's split tree, taking a NaN far distance as a
// miss.  Under bbox_ray's min/max order the far distance is NaN only when
// the x slab gives 0 * inf: the ray's direction has x = +-0 and the box's
// low x plane holds the ray origin (y and z slabs never make it NaN).  Such
// a box counts as hit at near distance 0 whatever its y and z extent, and it
// never takes a hit itself, as (near + far) / 2 is NaN.  Halving at
// midpoints never moves a half's low x plane below its parent's
// ((a + b) * 0.5 >= min(a, b) in floating point), so each box below has its
// low x plane at the origin (NaN again) or beyond it, where it misses: with
// d.x = +0 its near distance is +inf while y or z (one of d.y, d.z is
// nonzero) gives a finite far one; with d.x = -0 its far distance is -inf.
// Skipping the subtree therefore changes no result, while walking it dices
// the whole column of the patch along the plane to micropolygons.
PSY_DEV bool patch_box_ray(V3 lo, V3 hi, V3 o, V3 dinv, float t, float* h0, float* h1) {
    return bbox_ray(lo, hi, o, dinv, t, h0, h1) && *h1 == *h1;
}

// How a trace handles a patch node that needs splits below the dice cache:
// D_FLAG flags the ray for the second pass (main pass), D_SERIAL splits on
// in the lane within a node budget (second pass; a ray that runs out of
// budget is flagged for the third pass), D_COOP splits with the 64 lanes of
// the wave together (third pass: every lane runs the same ray).
constexpr int D_FLAG = 0, D_SERIAL = 1, D_COOP = 2;
struct DeepState {
    bool need;          // flagged for the next pass
    int32_t budget;     // D_SERIAL: node expansions left
    uint64_t* coop;     // D_COOP: the wave's node stack in LDS
};
constexpr int COOP_CAP = 2048;         // node stack entries per wave in the third pass

// Split-dice below the cached depth, over the subtree of a node that is
// already known to need splitting (`base`: its patch, `uvb`: its parameter
// rectangle, `sidx0`: its split-stack index), nearest half first: a node's
// two halves are bounded and tested together, the nearer one that needs
// splitting is descended into (the second half on ties, as the reference
// visits it first) and the other is left pending, to be re-tested against
// the best hit found meanwhile.  Visiting near to far matters for rays that
// graze a surface: their slab overlaps long runs of micro-bboxes, and a
// fixed order can improve the hit at every one of them, while near-to-far
// prunes all but the first.  The walk keeps no patch stack: a node is its
// path of halves (`pb`, deepest half in bit 0), replayed from `base`, and
// `trail` has one bit per depth whose sibling is pending.  The reference's
// split-stack index (object/patch_utils.hpp:80: a node at index 63 is a
// leaf) grows by one for every second half on the path.
template <int N>
PSY_DEV bool dice_deep_from(const DevScene& s, const Patch<N>& base, const float uvb[4], int sidx0, uint32_t inst,
                            V3 o, V3 dinv, const RayW& w, bool occl, Best& b, DeepState& ds) {
    int k = 0;           // depth below the start node
    uint64_t pb = 0;     // path of the current node
    uint64_t trail = 0;  // bit d: the sibling of the path's node at depth d is pending
    bool expand = true;  // the current node is hit and needs splitting
    for (;;) {
        if (expand) {
            if (--ds.budget < 0) {
                ds.need = true;
                return true;
            }
            Patch<N> p = base;
            float uv[4] = {uvb[0], uvb[1], uvb[2], uvb[3]};
            for (int l = k - 1; l >= 0; --l) patch_step<N>(p, ((pb >> l) & 1ull) != 0, uv);
            Patch<N> p1 = p;
            float uv1[4] = {uv[0], uv[1], uv[2], uv[3]};
            patch_step<N>(p1, true, uv1);
            patch_step<N>(p, false, uv);
            const int sidx = sidx0 + __builtin_popcountll(pb);  // of the first half; the second adds one
            bool go0 = false, go1 = false;
            float n0 = 0.0f, n1 = 0.0f;
            {
                V3 lo, hi;
                patch_bound<N>(p1, &lo, &hi);
                float h0, h1;
                if (patch_box_ray(lo, hi, o, dinv, b.t, &h0, &h1)) {
                    const float width = dice_width(min_width(w, h0, h1) * s.dice_rate, s.min_upoly);
                    const float max_dim = longest_axis(sub(hi, lo));
                    if (max_dim <= width || sidx + 1 == 63 || k + 1 == 62) {
                        if (accept_leaf(h0, h1, (uv1[0] + uv1[1]) * 0.5f, (uv1[2] + uv1[3]) * 0.5f, max_dim, inst,
                                        b) && occl)
                            return true;
                    } else {
                        go1 = true;
                        n1 = h0;
                    }
                }
            }
            {
                V3 lo, hi;
                patch_bound<N>(p, &lo, &hi);
                float h0, h1;
                if (patch_box_ray(lo, hi, o, dinv, b.t, &h0, &h1)) {
                    const float width = dice_width(min_width(w, h0, h1) * s.dice_rate, s.min_upoly);
                    const float max_dim = longest_axis(sub(hi, lo));
                    if (max_dim <= width || sidx == 63 || k + 1 == 62) {
                        if (accept_leaf(h0, h1, (uv[0] + uv[1]) * 0.5f, (uv[2] + uv[3]) * 0.5f, max_dim, inst, b) &&
                            occl)
                            return true;
                    } else {
                        go0 = true;
                        n0 = h0;
                    }
                }
            }
            if (go0 || go1) {
                const bool take1 = go1 && (!go0 || n1 <= n0);
                if (go0 && go1) trail |= 1ull << (k + 1);
                pb = (pb << 1) | (take1 ? 1ull : 0ull);
                ++k;
                continue;
            }
        }
        // the deepest pending sibling, re-tested against the best hit so far
        if (trail == 0) return false;
        const int d = 63 - __builtin_clzll(trail);
        trail &= ~(1ull << d);
        pb = (pb >> (k - d)) ^ 1ull;
        k = d;
        Patch<N> p = base;
        float uv[4] = {uvb[0], uvb[1], uvb[2], uvb[3]};
        for (int l = k - 1; l >= 0; --l) patch_step<N>(p, ((pb >> l) & 1ull) != 0, uv);
        V3 lo, hi;
        patch_bound<N>(p, &lo, &hi);
        float h0, h1;
        expand = false;
        if (patch_box_ray(lo, hi, o, dinv, b.t, &h0, &h1)) {
            const float width = dice_width(min_width(w, h0, h1) * s.dice_rate, s.min_upoly);
            const float max_dim = longest_axis(sub(hi, lo));
            if (max_dim <= width || sidx0 + __builtin_popcountll(pb) == 63 || k == 62) {
                if (accept_leaf(h0, h1, (uv[0] + uv[1]) * 0.5f, (uv[2] + uv[3]) * 0.5f, max_dim, inst, b) && occl)
                    return true;
            } else {
                expand = true;
            }
        }
    }
}

// dice_deep_from below heap node k0 of a patch's split tree.
template <int N>
PSY_DEV bool dice_deep(const DevScene& s, const DevObject& ob, uint32_t k0, uint32_t inst, V3 o, V3 dinv,
                       const RayW& w, bool occl, Best& b, DeepState& ds) {
    Patch<N> base;
    load_patch<N>(s.odata, ob.data, base);
    float uvb[4];
    patch_at_node<N>(base, k0, uvb);
    return dice_deep_from<N>(s, base, uvb, __builtin_popcount(k0) - 1, inst, o, dinv, w, occl, b, ds);
}

// The wave's best of the lanes' hits: least t, the lowest lane on ties.
// (Device only, as is everything of the cooperative walk.)
__device__ inline void wave_best(Best& b) {
    for (int off = 32; off >= 1; off >>= 1) {
        const float t2 = __shfl_xor(b.t, off);
        const float u2 = __shfl_xor(b.u, off);
        const float v2 = __shfl_xor(b.v, off);
        const float a2 = __shfl_xor(b.aux, off);
        const uint32_t i2 = (uint32_t)__shfl_xor((int)b.inst, off);
        const bool other_low = (__lane_id() & (uint32_t)off) != 0;  // the partner has the lower lane
        if (t2 < b.t || (t2 == b.t && other_low)) {
            b.t = t2; b.u = u2; b.v = v2; b.aux = a2; b.inst = i2;
        }
    }
}

// LDS order between the lanes of one wave (no other wave shares the memory).
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// A node of the cooperative walk: its path below k0 under a leading one bit.
template <int N>
__device__ void coop_node_patch(const Patch<N>& base, const float uvb[4], uint64_t node, Patch<N>& p, float uv[4]) {
    const int k = 63 - __builtin_clzll(node);
    p = base;
    uv[0] = uvb[0]; uv[1] = uvb[1]; uv[2] = uvb[2]; uv[3] = uvb[3];
    for (int l = k - 1; l >= 0; --l) patch_step<N>(p, ((node >> l) & 1ull) != 0, uv);
}

// dice_deep with the 64 lanes of a wave on one ray (every lane calls it with
// the same arguments): each round the lanes take up to 64 nodes off the
// wave's LDS stack, split them, test both halves against the wave's best
// hit, accept leaves and push the halves that need splitting (the nearer
// one last, so it comes off first).  The closest hit is the same in any
// visit order.  A half that finds the stack full is walked by its lane
// alone.  Every lane returns the same result.
template <int N>
__device__ bool dice_deep_coop(const DevScene& s, const DevObject& ob, uint32_t k0, uint32_t inst, V3 o, V3 dinv,
                            const RayW& w, bool occl, Best& b, uint64_t* stk) {
    Patch<N> base;
    load_patch<N>(s.odata, ob.data, base);
    float uvb[4];
    patch_at_node<N>(base, k0, uvb);
    const int sidx0 = __builtin_popcount(k0) - 1;
    const uint32_t lane = __lane_id();
    const unsigned long long lt = (1ull << lane) - 1ull;
    if (lane == 0) stk[0] = 1ull;  // k0 itself
    wave_lds_sync();
    int count = 1;
    while (count > 0) {
        const int take = count < 64 ? count : 64;
        const int base_i = count - take;
        const bool active = (int)lane < take;
        const uint64_t node = active ? stk[base_i + (int)lane] : 1ull;
        count = base_i;
        wave_lds_sync();
        bool push0 = false, push1 = false, accepted = false;
        uint64_t c0 = 0, c1 = 0;
        uint64_t lower = 0, upper = 0;  // with both halves: the farther below, the nearer on top
        if (active) {
            const int k = 63 - __builtin_clzll(node);
            Patch<N> p;
            float uv[4];
            coop_node_patch<N>(base, uvb, node, p, uv);
            Patch<N> p1 = p;
            float uv1[4] = {uv[0], uv[1], uv[2], uv[3]};
            patch_step<N>(p1, true, uv1);
            patch_step<N>(p, false, uv);
            const int sidx = sidx0 + __builtin_popcountll(node) - 1;
            float n0 = 0.0f, n1 = 0.0f;
            {
                V3 lo, hi;
                patch_bound<N>(p1, &lo, &hi);
                float h0, h1;
                if (patch_box_ray(lo, hi, o, dinv, b.t, &h0, &h1)) {
                    const float width = dice_width(min_width(w, h0, h1) * s.dice_rate, s.min_upoly);
                    const float max_dim = longest_axis(sub(hi, lo));
                    if (max_dim <= width || sidx + 1 == 63 || k + 1 == 62) {
                        accepted |= accept_leaf(h0, h1, (uv1[0] + uv1[1]) * 0.5f, (uv1[2] + uv1[3]) * 0.5f, max_dim,
                                                inst, b);
                    } else {
                        push1 = true;
                        n1 = h0;
                        c1 = (node << 1) | 1ull;
                    }
                }
            }
            {
                V3 lo, hi;
                patch_bound<N>(p, &lo, &hi);
                float h0, h1;
                if (patch_box_ray(lo, hi, o, dinv, b.t, &h0, &h1)) {
                    const float width = dice_width(min_width(w, h0, h1) * s.dice_rate, s.min_upoly);
                    const float max_dim = longest_axis(sub(hi, lo));
                    if (max_dim <= width || sidx == 63 || k + 1 == 62) {
                        accepted |= accept_leaf(h0, h1, (uv[0] + uv[1]) * 0.5f, (uv[2] + uv[3]) * 0.5f, max_dim,
                                                inst, b);
                    } else {
                        push0 = true;
                        n0 = h0;
                        c0 = node << 1;
                    }
                }
            }
            // the second half counts as nearer on ties
            if (push0 && push1) {
                const bool second_nearer = n1 <= n0;
                lower = second_nearer ? c0 : c1;
                upper = second_nearer ? c1 : c0;
            } else {
                lower = push0 ? c0 : c1;
            }
        }
        const bool pa = push0 || push1;
        const bool pd = push0 && push1;
        const unsigned long long ma = __ballot(pa), md = __ballot(pd);
        const int na = __popcll(ma);
        const int ia = count + __popcll(ma & lt);
        const int ib = count + na + __popcll(md & lt);
        DeepState solo{false, INT32_MAX, nullptr};
        if (pa) {
            if (ia < COOP_CAP) {
                stk[ia] = lower;
            } else {
                Patch<N> q;
                float uq[4];
                coop_node_patch<N>(base, uvb, lower, q, uq);
                accepted |= dice_deep_from<N>(s, q, uq, sidx0 + __builtin_popcountll(lower) - 1, inst, o, dinv, w,
                                              occl, b, solo);
            }
        }
        if (pd) {
            if (ib < COOP_CAP) {
                stk[ib] = upper;
            } else {
                Patch<N> q;
                float uq[4];
                coop_node_patch<N>(base, uvb, upper, q, uq);
                accepted |= dice_deep_from<N>(s, q, uq, sidx0 + __builtin_popcountll(upper) - 1, inst, o, dinv, w,
                                              occl, b, solo);
            }
        }
        count = min(count + na + __popcll(md), COOP_CAP);
        wave_lds_sync();
        wave_best(b);
        if (occl && __ballot(accepted) != 0ull) return true;
    }
    return false;
}

template <int DEEP>
PSY_DEV bool deep_or_flag(const DevScene& s, const DevObject& ob, uint32_t k, uint32_t inst, V3 o,
                                             V3 dinv, const RayW& w, bool occl, Best& b, DeepState& ds) {
    if constexpr (DEEP == D_FLAG) {
        ds.need = true;
        return false;
    } else if constexpr (DEEP == D_COOP) {
        return ob.kind == K_BILINEAR ? dice_deep_coop<4>(s, ob, k, inst, o, dinv, w, occl, b, ds.coop)
                                     : dice_deep_coop<16>(s, ob, k, inst, o, dinv, w, occl, b, ds.coop);
    } else {
        return ob.kind == K_BILINEAR ? dice_deep<4>(s, ob, k, inst, o, dinv, w, occl, b, ds)
                                     : dice_deep<16>(s, ob, k, inst, o, dinv, w, occl, b, ds);
    }
}

// Walk one patch's cached split tree, nearest child first, without a
// per-ray stack: `trail` keeps one bit per depth whose other child is still
// pending, and a node's sibling and ancestors follow from its heap index.
// Any visit order finds the reference's closest hit (the leaf t does not
// depend on max_t); within a node the second half is taken on ties, as the
// reference visits it first.  Returns true when an occlusion ray is done.
template <int DEEP>
PSY_DEV bool dice_walk(const DevScene& s, const DevObject& ob, uint32_t inst, V3 o, V3 dinv,
                                          const RayW& w, bool occl, Best& b, DeepState& need_deep) {
    const float4* nodes = s.dice + 2 * ob.dice;
    const int levels = ob.dice_levels;
    {
        const float4 a = nodes[2], c = nodes[3];
        float h0, h1;
        if (!patch_box_ray(xyz(a), xyz(c), o, dinv, b.t, &h0, &h1)) return false;
        const float maxd = longest_axis(sub(xyz(c), xyz(a)));
        if (maxd <= dice_width(min_width(w, h0, h1) * s.dice_rate, s.min_upoly))
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
        const bool hit0 = patch_box_ray(xyz(a0), xyz(e0), o, dinv, b.t, &h00, &h01);
        const bool hit1 = patch_box_ray(xyz(a1), xyz(e1), o, dinv, b.t, &h10, &h11);
        bool split0 = false, split1 = false;
        if (hit1) {
            const float maxd = longest_axis(sub(xyz(e1), xyz(a1)));
            if (maxd <= dice_width(min_width(w, h10, h11) * s.dice_rate, s.min_upoly)) {
                if (accept_leaf(h10, h11, a1.w, e1.w, maxd, inst, b) && occl) return true;
            } else {
                split1 = true;
            }
        }
        if (hit0) {
            const float maxd = longest_axis(sub(xyz(e0), xyz(a0)));
            if (maxd <= dice_width(min_width(w, h00, h01) * s.dice_rate, s.min_upoly)) {
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
            if (!patch_box_ray(xyz(a), xyz(c), o, dinv, b.t, &h0, &h1)) continue;
            const float maxd = longest_axis(sub(xyz(c), xyz(a)));
            if (maxd <= dice_width(min_width(w, h0, h1) * s.dice_rate, s.min_upoly)) {
                if (accept_leaf(h0, h1, a.w, c.w, maxd, inst, b) && occl) return true;
                continue;
            }
            break;
        }
    }
}

// Sphere::intersect_ray / SphereLight::intersect_ray quadratic
// (object/sphere.cpp:67-140)
PSY_DEV bool sphere_t(V3 ro, V3 rd, float max_t, V3 cent, float radi, float* t_out) {
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

PSY_DEV void sphere_at(const float* d, int ts, int stride, float time, V3* cent, float* radi) {
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

// RectangleLight::intersect_ray (lights/rectangle_light.hpp:110-147);
// `fix`: test y against the height (PSY_SCENE_RECT_LIGHT_FIX)
PSY_DEV bool rect_t(const float* d, int ts, float time, V3 o, V3 dir, V3 dinv, float max_t, bool fix,
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
    // the y test compares against the width, as the reference does, unless fixed
    if (x >= (wd * -0.5f) && x <= (wd * 0.5f) && y >= (ht * -0.5f) && y <= ((fix ? ht : wd) * 0.5f)) {
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

// A BVH4 node's child boxes at the ray's time: the stored boxes for a
// static scene; for a moving one, the time-0 boxes plus the ray's time
// (clamped to the shutter, as the transforms are) times their change up to
// time 1.  Each endpoint box is the union of its subtree's, padded on the
// host so the interpolation holds every leaf at every time (and the pad
// covers the rounding of the two operations here).
//
// The planes come back as near / far per axis for the ray's direction
// signs (ds*: that component of the inverse direction is negative, so the
// max plane is met first): each lane loads them from its own offsets
// instead of loading min and max and selecting per child.
PSY_DEV void node_boxes(const DevScene& s, int32_t ref, float time, bool ds0, bool ds1, bool ds2, float4& nx,
                        float4& fx, float4& ny, float4& fy, float4& nz, float4& fz) {
    // 32-bit byte offsets from the (wave-uniform) array base
    const uint32_t o = (uint32_t)ref * (uint32_t)sizeof(BvhNode);
    const uint32_t ix = ds0 ? 16u : 0u, iy = ds1 ? 48u : 32u, iz = ds2 ? 80u : 64u;
    const char* b = reinterpret_cast<const char*>(s.bvh);
#define PSY_AT(B, OFF) (*reinterpret_cast<const float4*>((B) + (OFF)))
    nx = PSY_AT(b, o + ix); fx = PSY_AT(b, o + (ix ^ 16u)); ny = PSY_AT(b, o + iy); fy = PSY_AT(b, o + (iy ^ 16u));
    nz = PSY_AT(b, o + iz); fz = PSY_AT(b, o + (iz ^ 16u));
    if (s.bvh1) {
        const char* b1 = reinterpret_cast<const char*>(s.bvh1);
        const float t = fminf(fmaxf(time, 0.0f), 1.0f);
        const float4 a0 = PSY_AT(b1, o + ix), a1 = PSY_AT(b1, o + (ix ^ 16u)), a2 = PSY_AT(b1, o + iy),
                     a3 = PSY_AT(b1, o + (iy ^ 16u)), a4 = PSY_AT(b1, o + iz), a5 = PSY_AT(b1, o + (iz ^ 16u));
#undef PSY_AT
#define PSY_MIX(V, W) V = make_float4(V.x + W.x * t, V.y + W.y * t, V.z + W.z * t, V.w + W.w * t)
        PSY_MIX(nx, a0); PSY_MIX(fx, a1); PSY_MIX(ny, a2); PSY_MIX(fy, a3); PSY_MIX(nz, a4); PSY_MIX(fz, a5);
#undef PSY_MIX
    }
}

template <bool WORLD, int DEEP>
PSY_DEV bool intersect_leaf(const DevScene& s, uint32_t li, const DevInstance& I, const TRay& r,
                                               Best& b, DeepState& need_deep) {
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
            if (rect_t(od, ob.ts, r.time, o, d, dinv, b.t, (s.flags & PSY_SCENE_RECT_LIGHT_FIX) != 0u, &t)) {
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
PSY_DEV void sphere_uv(const DevScene& s, Best& b, const TRay& r) {
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

PSY_DEV void cmpswap(float& ka, int32_t& ra, float& kb, int32_t& rb) {
    if (kb < ka) {
        const float tk = ka; ka = kb; kb = tk;
        const int32_t tr = ra; ra = rb; rb = tr;
    }
}

#ifndef PSY_TRACE_BLOCK
#define PSY_TRACE_BLOCK 64
#endif
constexpr int TRACE_BLOCK = PSY_TRACE_BLOCK;  // threads per block of the trace kernels
// The second pass has two builds, both launched on every trace; each reads
// the queue length and only one of them walks the queue.  The narrow one
// runs 1 wave per SIMD (256 VGPRs, no spills): the lowest latency per ray,
// for short queues (a primary stream hands on a few hundred rays).  The wide
// one runs 2 waves per SIMD (128 VGPRs, paid for with spills): 1.4x the
// throughput on long queues (a ground plane diced below the cache hands on
// half the camera rays).  Each launches one resident round of 64-thread
// blocks over 256 CUs, as does the third pass (1 wave per SIMD).  The
// queue length that picks the wide build (DEEP_WIDE_MIN by default) and
// the second pass's node budget (DEEP_BUDGET) are launch parameters.
constexpr int DEEP_WIDE_WAVES = 2;
constexpr int DEEP_ROUND = 1024;          // blocks per wave per SIMD over 256 CUs

// A ray of the stream: world form (PSY_WRAY_PLANES, footprint derived here,
// basics/ray.hpp:310-323) or object form (PSY_RAY_PLANES).
template <bool WORLD>
PSY_DEV void load_ray(const float* __restrict__ rays, int64_t n, int64_t i, uint32_t flags,
                                         TRay& r, float& max_t) {
    if (WORLD) {
        r.o = mk(rays[0 * n + i], rays[1 * n + i], rays[2 * n + i]);
        r.d = mk(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        r.odx = mk(rays[6 * n + i], rays[7 * n + i], rays[8 * n + i]);
        r.ody = mk(rays[9 * n + i], rays[10 * n + i], rays[11 * n + i]);
        r.ddx = mk(rays[12 * n + i], rays[13 * n + i], rays[14 * n + i]);
        r.ddy = mk(rays[15 * n + i], rays[16 * n + i], rays[17 * n + i]);
        r.time = rays[18 * n + i];
        const uint32_t type = f2u(rays[19 * n + i]);
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
        r.occl = (f2u(rays[14 * n + i]) & (1u << 30)) != 0;
        r.odx = r.ody = r.ddx = r.ddy = splat(0.0f);
    }
    if (flags & PSY_TRACE_OCCLUSION_ONLY) r.occl = true;
    r.dinv = divv(splat(1.0f), r.d);
}


// One ray: BVH4 nearest-first with an LDS stack, leaves through
// intersect_leaf, below-cache splits per DEEP (D_SERIAL within `budget`
// node expansions, D_COOP with the wave's stack `coop`).  Returns true if
// the ray is flagged for the next pass.
template <bool WORLD, int STACK, int DEEP>
PSY_DEV bool trace_one(const DevScene& s, const float* __restrict__ rays, int64_t n, int64_t i,
                                          const DevHits& h, uint32_t flags, int32_t (*stk_ref)[TRACE_BLOCK],
                                          float (*stk_near)[TRACE_BLOCK], int tid, int32_t budget = INT32_MAX,
                                          uint64_t* coop = nullptr) {
    TRay r;
    float max_t;
    load_ray<WORLD>(rays, n, i, flags, r, max_t);

    Best b;
    b.t = max_t;
    b.inst = 0xffffffffu;
    b.u = b.v = b.aux = 0.0f;
    DeepState need_deep{false, budget, coop};

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
            float4 nx, fx, ny, fy, nz, fz;
            node_boxes(s, ref, r.time, ds0, ds1, ds2, nx, fx, ny, fy, nz, fz);
            const int4 ch = reinterpret_cast<const int4*>(s.bvh + ref)[6];
            // BBox4::intersect_ray lanes (basics/bbox.hpp:566-595); keys are
            // the near distances of hit children, +inf otherwise
#define PSY_LANE(L, KEY)                                                                             \
    float KEY;                                                                                       \
    {                                                                                                \
        const float xlo = (nx.L - r.o.x) * r.dinv.x;                                                 \
        const float xhi = (fx.L - r.o.x) * r.dinv.x;                                                 \
        const float ylo = (ny.L - r.o.y) * r.dinv.y;                                                 \
        const float yhi = (fy.L - r.o.y) * r.dinv.y;                                                 \
        const float zlo = (nz.L - r.o.z) * r.dinv.z;                                                 \
        const float zhi = (fz.L - r.o.z) * r.dinv.z;                                                 \
        const float mins = sse_max(sse_max(xlo, ylo), sse_max(zlo, 0.0f));                           \
        const float maxs = sse_max(sse_min(sse_min(xhi, yhi), zhi), -INFINITY) * PSY_BBOX_MAXT_ADJUST; \
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
            const uint32_t li = (uint32_t)(~ref);
            if (intersect_leaf<WORLD, DEEP>(s, li, s.inst[li], r, b, need_deep)) break;
        }
    }
    const bool hit = b.inst != 0xffffffffu;
    if (hit && !r.occl) sphere_uv(s, b, r);
    h.t[i] = hit ? b.t : INFINITY;
    h.inst[i] = b.inst;
    h.u[i] = b.u;
    h.v[i] = b.v;
    if (h.aux) h.aux[i] = b.aux;
    return need_deep.need && !(r.occl && hit);
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
#ifndef PSY_MAIN_STACK
#define PSY_MAIN_STACK 16
#endif
#ifndef PSY_MAIN_WAVES_PER_SIMD
#define PSY_MAIN_WAVES_PER_SIMD 5
#endif
constexpr int MAIN_STACK = PSY_MAIN_STACK;
// The main pass also comes with a 32-entry stack at lower occupancy (fewer
// rays handed to the second pass on deep hierarchies); waves per SIMD of
// each form, as launch bounds.
constexpr int main_waves(bool world, int stack, bool coherent = false) {
    return stack <= 16 ? ((world || coherent) ? 4 : PSY_MAIN_WAVES_PER_SIMD) : 2;
}
// The persistent form keeps its refill state beside the lane's and needs a
// few more registers: at 5 waves per SIMD it spilled to scratch.
#ifndef PSY_PERSIST_WAVES_PER_SIMD
#define PSY_PERSIST_WAVES_PER_SIMD 4
#endif
constexpr int persist_waves(bool world, int stack) { return stack <= 16 ? PSY_PERSIST_WAVES_PER_SIMD : 2; }

struct Lane {
    TRay r;
    Best b;
    uint32_t idx;
    // split-tree walk (valid while walking): the patch's dice-cache slot,
    // the heap node k being expanded at `depth`, and one `trail` bit per
    // depth whose other child is still to be expanded; world rays also keep
    // the patch-space ray and footprint (object rays never meet transformed
    // patches, so theirs is the ray itself)
    uint32_t slot;
    V3 wo, wdinv;
    RayW ww;
    float wfx, wfy;  // flip distances of the patch-space footprint
    // the leaf transforms wo .. wfy were derived for: the first transform's
    // index, -2 for an untransformed leaf, -1 for none yet
    int32_t wframe;
    uint32_t winst, k, trail;
    int levels, depth;
    // BVH: the nearest hit child, visited next without a stack round trip
    // (CHILD_EMPTY: pop the stack)
    int32_t next;
    int sp;
    uint32_t state;  // LANE_* bits
    // the split-tree nodes at the bottom of the dice cache that still need
    // splitting, as a list in the trace's deep-node pool (DEEP_NONE: empty);
    // LANE_DEEP_MANY once the pool had no room for one
    uint32_t deep_head;
};
enum : uint32_t { LANE_WALKING = 1u, LANE_NEED_DEEP = 2u, LANE_RETRACE = 4u, LANE_DEEP_MANY = 8u };
constexpr uint32_t DEEP_NONE = 0xffffffffu;

// The deep-node pool of one trace call: entry e is 3 words (leaf instance,
// heap node, next entry of the same ray's list or DEEP_NONE); `count` hands
// out entries, `cap` bounds them (0: no pool, every such ray is retraced).
struct DeepPool {
    uint32_t* e;
    uint32_t* count;
    uint32_t cap;
};

PSY_DEV uint32_t atomic_add_u32(uint32_t* p, uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return atomicAdd(p, v);
#else
    return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
#endif
}

// `want` consecutive entries of a counter for each calling lane: one atomic
// per wave for the lanes of the wave that call it together.
PSY_DEV uint32_t wave_alloc(uint32_t* counter, uint32_t want) {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned long long b1 = __ballot(want >= 1u), b2 = __ballot(want >= 2u);
    const unsigned long long act = __ballot(1);
    const uint32_t lane = __lane_id();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const uint32_t off = (uint32_t)(__popcll(b1 & lt) + __popcll(b2 & lt));
    const uint32_t total = (uint32_t)(__popcll(b1) + __popcll(b2));
    const int leader = __ffsll((long long)act) - 1;
    uint32_t base = 0;
    if ((int)lane == leader && total) base = atomicAdd(counter, total);
    base = (uint32_t)__shfl((int)base, leader);
    return base + off;
#else
    return __atomic_fetch_add(counter, want, __ATOMIC_RELAXED);
#endif
}

template <bool WORLD>
PSY_DEV V3 walk_o(const Lane& L) { return WORLD ? L.wo : L.r.o; }
template <bool WORLD>
PSY_DEV V3 walk_dinv(const Lane& L) { return WORLD ? L.wdinv : L.r.dinv; }
template <bool WORLD>
PSY_DEV const RayW& walk_w(const Lane& L) { return WORLD ? L.ww : L.r.w; }

PSY_DEV void lane_start(const DevScene& s, Lane& L, float max_t, uint32_t i);

template <bool WORLD>
PSY_DEV void lane_init(const DevScene& s, const float* __restrict__ rays, int64_t n, uint32_t i,
                                          uint32_t flags, Lane& L) {
    float max_t;
    load_ray<WORLD>(rays, n, (int64_t)i, flags, L.r, max_t);
    lane_start(s, L, max_t, i);
}

// Traversal state of a lane whose ray L.r is loaded.
PSY_DEV void lane_start(const DevScene& s, Lane& L, float max_t, uint32_t i) {
    L.b.t = max_t;
    L.b.inst = 0xffffffffu;
    L.b.u = L.b.v = L.b.aux = 0.0f;
    L.idx = i;
    L.slot = 0;
    L.levels = 0;
    L.winst = 0;
    L.k = 0;
    L.trail = 0;
    L.depth = 0;
    L.next = s.root;
    L.sp = 0;
    L.state = 0;
    L.deep_head = DEEP_NONE;
    L.wframe = -1;  // read by world rays only
    // the main pass's box tests assume a finite inverse direction; the
    // second pass follows the reference's NaN behaviour exactly
    if (!finite3(L.r.dinv)) {
        L.state = LANE_RETRACE;
        L.next = CHILD_EMPTY;
    }
}

// The patch-space ray of a walk through leaf I's split tree (world rays:
// through the leaf's transform at the ray's time) and its footprint's flip
// distances; false when its inverse direction is not finite.  They depend
// on the leaf's transforms alone, so the patches of one instance (a
// subdivision surface's, an instanced assembly's) reuse them: a ray that
// meets several patches of a transformed instance derives them once.
template <bool WORLD>
PSY_DEV bool walk_ray(const DevScene& s, Lane& L, const DevInstance& I) {
    // object rays keep only the two flip distances, cheaper to re-derive
    // than to hold in a register through the BVH walk
    const int32_t frame = I.xf_count > 0 ? I.xf_offset : -2;
    if (WORLD && frame == L.wframe) return true;
    bool finite = true;
    if (WORLD) {
        L.wo = L.r.o;
        L.wdinv = L.r.dinv;
        L.ww = L.r.w;
        if (I.xf_count > 0) {
            const M44 x = m_lerp_seq(L.r.time, s.xf + I.xf_offset, I.xf_count);
            L.wo = m_pos(x, L.r.o);
            const V3 d = m_dir(x, L.r.d);
            L.wdinv = divv(splat(1.0f), d);
            finite = finite3(L.wdinv);
            L.ww = ray_widths(L.wo, d, m_dir(x, L.r.odx), m_dir(x, L.r.ody), m_dir(x, L.r.ddx), m_dir(x, L.r.ddy));
        }
    }
    L.wfx = flip_x(walk_w<WORLD>(L));
    L.wfy = flip_y(walk_w<WORLD>(L));
    if (WORLD) L.wframe = finite ? frame : -1;
    return finite;
}

// One unit of work; true when the ray is finished (or must be retraced).
//   walking : expand heap node k of the patch's cached split tree -- test
//             its two children (k = 0 stands for the root alone), accept
//             leaves, descend into the nearer child to split; with nothing
//             to descend into, move to the deepest pending sibling (whose
//             box was already found hit and too large when its parent was
//             expanded; its children lie inside it, so expanding it without
//             a re-test finds the same hits)
//   else    : one BVH4 node, or one leaf instance
template <bool WORLD, int STACK = MAIN_STACK>
PSY_DEV bool lane_step(const DevScene& s, Lane& L, int32_t (*stk_ref)[TRACE_BLOCK],
                                          float (*stk_near)[TRACE_BLOCK], int tid, const DeepPool& P) {
    if (L.state & LANE_WALKING) {
        const uint32_t c0 = 2 * L.k;
        const float4* q = s.dice + 2 * ((size_t)L.slot + c0);
        const V3 wo = walk_o<WORLD>(L), wdinv = walk_dinv<WORLD>(L);
        const RayW& ww = walk_w<WORLD>(L);
        const float4 a0 = q[0], e0 = q[1], a1 = q[2], e1 = q[3];
        float h00, h01, h10, h11;
        const bool hit0 = bbox_ray_fast(xyz(a0), xyz(e0), wo, wdinv, L.b.t, &h00, &h01) && L.k != 0;
        const bool hit1 = bbox_ray_fast(xyz(a1), xyz(e1), wo, wdinv, L.b.t, &h10, &h11);
        bool split0 = false, split1 = false;
        // the second half first: the reference visits it first on ties
        if (hit1) {
            const float maxd = longest_axis(sub(xyz(e1), xyz(a1)));
            if (maxd <= dice_width(min_width_flip(ww, L.wfx, L.wfy, h10, h11) * s.dice_rate, s.min_upoly)) {
                if (accept_leaf(h10, h11, a1.w, e1.w, maxd, L.winst, L.b) && L.r.occl) return true;
            } else {
                split1 = true;
            }
        }
        if (hit0) {
            const float maxd = longest_axis(sub(xyz(e0), xyz(a0)));
            if (maxd <= dice_width(min_width_flip(ww, L.wfx, L.wfy, h00, h01) * s.dice_rate, s.min_upoly)) {
                if (accept_leaf(h00, h01, a0.w, e0.w, maxd, L.winst, L.b) && L.r.occl) return true;
            } else {
                split0 = true;
            }
        }
        split0 = split0 && h00 <= L.b.t;
        split1 = split1 && h10 <= L.b.t;
        const int cd = L.depth + 1;
        if (cd == L.levels) {
            // children that need splits below the cache go onto the lane's
            // list for the second pass to resume from
            const uint32_t nd = (split0 ? 1u : 0u) + (split1 ? 1u : 0u);
            if (nd) {
                L.state |= LANE_NEED_DEEP;
                // rare, so one atomic per lane (a wave-aggregated one costs the main
                // pass registers)
                const uint32_t e0 = P.cap ? atomic_add_u32(P.count, nd) : 0u;
                if (e0 + nd > P.cap) {
                    L.state |= LANE_DEEP_MANY;
                } else {
                    uint32_t e = e0;
                    if (split1) {
                        P.e[3 * e] = L.winst; P.e[3 * e + 1] = c0 + 1; P.e[3 * e + 2] = L.deep_head;
                        L.deep_head = e++;
                    }
                    if (split0) {
                        P.e[3 * e] = L.winst; P.e[3 * e + 1] = c0; P.e[3 * e + 2] = L.deep_head;
                        L.deep_head = e;
                    }
                }
            }
            split0 = split1 = false;
        }
        if (split0 && split1) {
            L.k = (h00 < h10) ? c0 : c0 + 1;
            L.trail |= 1u << cd;
            L.depth = cd;
        } else if (split0 || split1) {
            L.k = split0 ? c0 : c0 + 1;
            L.depth = cd;
        } else if (L.trail == 0) {
            L.state &= ~LANE_WALKING;
        } else {
            const int dd = 31 - __builtin_clz(L.trail);
            L.k = (L.k >> (L.depth - dd)) ^ 1u;
            L.depth = dd;
            L.trail &= ~(1u << dd);
        }
        return false;
    }

    int32_t ref;
    if (L.next != CHILD_EMPTY) {
        ref = L.next;
        L.next = CHILD_EMPTY;
    } else {
        if (L.sp == 0) return true;
        --L.sp;
        ref = stk_ref[L.sp][tid];
        if (!(stk_near[L.sp][tid] < L.b.t)) return false;
    }
    if (ref >= 0) {
        const TRay& r = L.r;
        const bool ds0 = r.dinv.x < 0.0f, ds1 = r.dinv.y < 0.0f, ds2 = r.dinv.z < 0.0f;
        float4 nx, fx, ny, fy, nz, fz;
        node_boxes(s, ref, r.time, ds0, ds1, ds2, nx, fx, ny, fy, nz, fz);
        const int4 ch = reinterpret_cast<const int4*>(s.bvh + ref)[6];
        // BBox4::intersect_ray lanes (basics/bbox.hpp:566-595) for a finite
        // inverse direction (no NaN); keys are the near distances of hit
        // children, +inf otherwise
#define PSY_LANE(LN, KEY)                                                                            \
    float KEY;                                                                                       \
    {                                                                                                \
        const float xlo = (nx.LN - r.o.x) * r.dinv.x;                                                \
        const float xhi = (fx.LN - r.o.x) * r.dinv.x;                                                \
        const float ylo = (ny.LN - r.o.y) * r.dinv.y;                                                \
        const float yhi = (fy.LN - r.o.y) * r.dinv.y;                                                \
        const float zlo = (nz.LN - r.o.z) * r.dinv.z;                                                \
        const float zhi = (fz.LN - r.o.z) * r.dinv.z;                                                \
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
        if (L.sp + npush > STACK) {
            L.state |= LANE_RETRACE;
            return true;
        }
        if (k3 < INFINITY) { stk_ref[L.sp][tid] = r3; stk_near[L.sp][tid] = k3; ++L.sp; }
        if (k2 < INFINITY) { stk_ref[L.sp][tid] = r2; stk_near[L.sp][tid] = k2; ++L.sp; }
        if (k1 < INFINITY) { stk_ref[L.sp][tid] = r1; stk_near[L.sp][tid] = k1; ++L.sp; }
        if (k0 < INFINITY) L.next = r0;
        return false;
    }
    // leaf instance: a patch starts a split-tree walk at its root
    const uint32_t li = (uint32_t)(~ref);
    const DevInstance I = s.inst[li];
    if (I.kind == K_BILINEAR || I.kind == K_BICUBIC) {
        // Walk state first, then the patch-space ray.  The transform block
        // has no exit of its own: with a `return` inside it, hipcc (ROCm
        // 7.2, gfx950) reused the register holding the dice slot inside the
        // block for lanes that then walked with slot 0.
        L.slot = (uint32_t)I.dice;
        L.levels = I.dice_levels;
        L.winst = li;
        L.k = 0;
        L.depth = -1;
        L.trail = 0;
        if (!walk_ray<WORLD>(s, L, I)) {
            L.state |= LANE_RETRACE;
            return true;
        }
        L.state |= LANE_WALKING;
        return false;
    }
    DeepState unused{false, 0, nullptr};
    return intersect_leaf<WORLD, D_FLAG>(s, li, I, L.r, L.b, unused);
}

// The lane's final hit (t, inst, u, v, aux) in place in L.b; true if the
// ray goes to the second pass.
PSY_DEV bool lane_result(const DevScene& s, Lane& L) {
    const bool hit = L.b.inst != 0xffffffffu;
    if (hit && !L.r.occl) sphere_uv(s, L.b, L.r);
    if (!hit) L.b.t = INFINITY;
    return (L.state & (LANE_NEED_DEEP | LANE_RETRACE)) != 0 && !(L.r.occl && hit);
}

// Writes the lane's hit; true if the ray goes to the second pass.
PSY_DEV bool lane_finish(const DevScene& s, Lane& L, const DevHits& h) {
    const bool again = lane_result(s, L);
    const uint32_t i = L.idx;
    h.t[i] = L.b.t;
    h.inst[i] = L.b.inst;
    h.u[i] = L.b.u;
    h.v[i] = L.b.v;
    if (h.aux) h.aux[i] = L.b.aux;
    return again;
}

// Hands a finished lane's ray to the second pass: queue entry = ray index,
// with bit 31 set when the second pass can resume it from the cache nodes
// it listed (heads[slot] = its list) instead of retracing it from the BVH
// root.
constexpr uint32_t QUEUE_RESUME = 0x80000000u;
PSY_DEV void enqueue_second(uint32_t* __restrict__ queue, uint32_t* __restrict__ heads,
                            uint32_t* __restrict__ counters, const Lane& L) {
    const uint32_t slot = wave_alloc(&counters[0], 1u);
    const bool listed = (L.state & (LANE_RETRACE | LANE_DEEP_MANY)) == 0;
    queue[slot] = L.idx | (listed ? QUEUE_RESUME : 0u);
    heads[slot] = L.deep_head;
}

// Continuations.  Once fewer than a set number of a main-pass wave's lanes
// are still tracing, the wave writes those lanes' traversal state to a
// continuation buffer and ends; a follow-on launch packs the continued
// rays 64 to a wave and finishes them.  A lane's state is CONT_FIXED words
// (best hit, ray index, walk state, BVH next / stack depth, flags, deep
// list) plus its stack entries (ref, near); the ray itself is reloaded from
// the stream and its walk ray re-derived from the walked leaf, so each ray
// takes the same steps as it would have in one launch.  Entry e's word w is
// at cont[w * cap + e].
constexpr int CONT_FIXED = 16;
// One level: the continued rays' own launch runs them to the end (a second
// level, continuing that launch's tails again, was slower: 2.21 against
// 2.01 ms on the 2^22-ray heightfield stream).
#ifndef PSY_CONT_LEVELS
#define PSY_CONT_LEVELS 1
#endif
constexpr int CONT_LEVELS = PSY_CONT_LEVELS;  // counted in the trace's counters[5 + level]
// A continuation buffer: `cap` entries of cont_words(stack) words; a wave
// continues its lanes into it once fewer than `below` of them are tracing
// (0: never); *count counts its entries, taken a wave at a time.  (Packing
// the continued rays in wave order instead, through a prefix sum of
// per-wave counts, left the follow-on launch's L2 hit rate as it was and
// cost the scan: 2.07 against 1.99 ms on the 2^22-ray heightfield stream.)
struct ContBuf {
    uint32_t* c;
    uint32_t cap;
    int below;
    uint32_t* count;
};

constexpr int cont_words(int stack) { return CONT_FIXED + 2 * stack; }

template <int STACK>
PSY_DEV void cont_save(uint32_t* __restrict__ cont, uint32_t cap, uint32_t e, const Lane& L,
                       int32_t (*stk_ref)[TRACE_BLOCK], float (*stk_near)[TRACE_BLOCK], int tid) {
    uint32_t* c = cont + e;
    const size_t m = cap;
    c[0 * m] = f2u(L.b.t);
    c[1 * m] = L.b.inst;
    c[2 * m] = f2u(L.b.u);
    c[3 * m] = f2u(L.b.v);
    c[4 * m] = f2u(L.b.aux);
    c[5 * m] = L.idx;
    c[6 * m] = L.slot;
    c[7 * m] = L.winst;
    c[8 * m] = L.k;
    c[9 * m] = L.trail;
    c[10 * m] = (uint32_t)L.levels;
    c[11 * m] = (uint32_t)L.depth;
    c[12 * m] = (uint32_t)L.next;
    c[13 * m] = (uint32_t)L.sp;
    c[14 * m] = L.state;
    c[15 * m] = L.deep_head;
    for (int j = 0; j < L.sp; ++j) {
        c[(CONT_FIXED + 2 * j) * m] = (uint32_t)stk_ref[j][tid];
        c[(CONT_FIXED + 2 * j + 1) * m] = f2u(stk_near[j][tid]);
    }
}

template <bool WORLD, int STACK>
PSY_DEV void cont_load(const DevScene& s, const float* __restrict__ rays, int64_t rs, uint32_t flags,
                       const uint32_t* __restrict__ cont, uint32_t cap, uint32_t e, Lane& L,
                       int32_t (*stk_ref)[TRACE_BLOCK], float (*stk_near)[TRACE_BLOCK], int tid) {
    const uint32_t* c = cont + e;
    const size_t m = cap;
    lane_init<WORLD>(s, rays, rs, c[5 * m], flags, L);
    L.b.t = u2f(c[0 * m]);
    L.b.inst = c[1 * m];
    L.b.u = u2f(c[2 * m]);
    L.b.v = u2f(c[3 * m]);
    L.b.aux = u2f(c[4 * m]);
    L.slot = c[6 * m];
    L.winst = c[7 * m];
    L.k = c[8 * m];
    L.trail = c[9 * m];
    L.levels = (int)c[10 * m];
    L.depth = (int)c[11 * m];
    L.next = (int32_t)c[12 * m];
    L.sp = (int)c[13 * m];
    L.state = c[14 * m];
    L.deep_head = c[15 * m];
    for (int j = 0; j < L.sp; ++j) {
        stk_ref[j][tid] = (int32_t)c[(CONT_FIXED + 2 * j) * m];
        stk_near[j][tid] = u2f(c[(CONT_FIXED + 2 * j + 1) * m]);
    }
    // a walking lane's leaf gave a finite walk ray when the walk started
    if (L.state & LANE_WALKING) (void)walk_ray<WORLD>(s, L, s.inst[L.winst]);
}

// Runs lane L to the end of its ray or, once fewer than `out.below` of the
// wave's lanes are left, writes it to `out` (all of the wave's remaining
// lanes together, if the buffer has room for them).  True when continued.
template <bool WORLD, int STACK>
PSY_DEV bool lane_run(const DevScene& s, Lane& L, int32_t (*stk_ref)[TRACE_BLOCK],
                      float (*stk_near)[TRACE_BLOCK], int tid, const DeepPool& pool, const ContBuf& out) {
    int below = out.below;  // wave-uniform
#if !defined(__HIP_DEVICE_COMPILE__)
    (void)below;  // the host pass has no wave intrinsics
#endif
    for (;;) {
        if (lane_step<WORLD, STACK>(s, L, stk_ref, stk_near, tid, pool)) return false;
#if defined(__HIP_DEVICE_COMPILE__)
        const int live = __popcll(__ballot(1));
        if (live < below) {
            const uint32_t e = wave_alloc(out.count, 1u);
            const uint32_t first = (uint32_t)__shfl((int)e, __ffsll((long long)__ballot(1)) - 1);
            if ((uint64_t)first + (uint64_t)live <= (uint64_t)out.cap) {
                cont_save<STACK>(out.c, out.cap, e, L, stk_ref, stk_near, tid);
                return true;
            }
            below = 0;  // no room: these lanes finish here
        }
#endif
    }
}

// Second pass for a ray the main pass left with a list of cache nodes to
// split further: its hit so far (as the main pass wrote it), then the
// split-dice below each listed node within a shared budget of node
// expansions.  The main pass explored everything else with the same or a
// looser bound on t, so the result is the full trace's.  Returns true when
// the budget ran out (the ray then takes the third pass, which retraces it
// whole).
template <bool WORLD>
PSY_DEV bool resume_list(const DevScene& s, const float* __restrict__ rays, int64_t n, int64_t i, uint32_t head,
                         const uint32_t* __restrict__ pool, const DevHits& h, uint32_t flags, int32_t budget) {
    TRay r;
    float max_t;
    load_ray<WORLD>(rays, n, i, flags, r, max_t);
    Best b;
    b.inst = h.inst[i];
    const bool hit = b.inst != 0xffffffffu;
    b.t = hit ? h.t[i] : max_t;
    b.u = h.u[i];
    b.v = h.v[i];
    b.aux = h.aux ? h.aux[i] : 0.0f;
    const uint32_t before = b.inst;
    const float t_before = b.t;
    DeepState ds{false, budget, nullptr};
    for (uint32_t e = head; e != DEEP_NONE; e = pool[3 * e + 2]) {
        const uint32_t li = pool[3 * e], k = pool[3 * e + 1];
        const DevInstance I = s.inst[li];
        const DevObject ob = s.obj[I.object];
        V3 o = r.o, dinv = r.dinv;
        RayW w = r.w;
        if (I.xf_count > 0) {
            const M44 x = m_lerp_seq(r.time, s.xf + I.xf_offset, I.xf_count);
            o = m_pos(x, r.o);
            const V3 d = m_dir(x, r.d);
            dinv = divv(splat(1.0f), d);
            if (WORLD) w = ray_widths(o, d, m_dir(x, r.odx), m_dir(x, r.ody), m_dir(x, r.ddx), m_dir(x, r.ddy));
        }
        bool done;
        if (ob.kind == K_BILINEAR) done = dice_deep<4>(s, ob, k, li, o, dinv, w, r.occl, b, ds);
        else done = dice_deep<16>(s, ob, k, li, o, dinv, w, r.occl, b, ds);
        if (ds.need) return true;
        if (done) break;  // an occlusion ray found its hit
    }
    if (b.inst != before || b.t != t_before) {
        h.t[i] = b.t;
        h.inst[i] = b.inst;
        h.u[i] = b.u;
        h.v[i] = b.v;
        if (h.aux) h.aux[i] = b.aux;
    }
    return false;
}

// Main pass, one thread per ray.

// Main pass, persistent: a grid that fills the GPU once; every wave claims
// rays from a global counter 256 at a time (one atomic per claim) and
// refills its finished lanes from its claim whenever at least `refill_min`
// of them are idle, so the wave's lanes stay busy through the long tail of
// incoherent rays.  Every wave leaves once the counter has passed n and its
// lanes are idle.

// Second pass: retrace the queued rays with subdivision below the cache and
// a stack of STACK >= the scene's worst case.

// counters[0]: retrace queue length; counters[2..3]: next ray (u64)

// ---------------------------------------------------------------------------
// WorldRay::to_ray over a stream (basics/ray.hpp:207-252, :310-323)
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Camera rays: init_path + next_ray_for_path step 0
// (integrator/path_trace_integrator.cpp:91-125, basics/camera.hpp:72-104,
// utils/monte_carlo.hpp:17-53, utils/utils.hpp:335-343)
// ---------------------------------------------------------------------------
#define PSY_QPI (3.1415926535897932384626433 / 4)

PSY_DEV float faster_ln(float x) {
    float y = (float)f2u(x);
    y *= 8.2629582881927490e-8f;
    return y - 87.989971088f;
}
PSY_DEV float fast_logit(float p, float width) {
    const float n = 0.001f + (p * 0.998f);
    return faster_ln((n / (1.0f - n))) * width * (0.6266f / 4.0f);
}
PSY_DEV void square_to_circle(float* x, float* y) {
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

PSY_DEV float lerp_seq_f(float alpha, const float* seq, int n) {
    float na = 0.0f;
    const int i = lerp_seq_index(alpha, n, &na);
    if (i == -1) return seq[0];
    if (i == -2) return seq[n - 1];
    return lerpf(na, seq[i], seq[i + 1]);
}

PSY_DEV void store_wray(float* w, int64_t n, int64_t i, V3 o, V3 d, V3 odx, V3 ody, V3 ddx,
                                           V3 ddy, float time, uint32_t type) {
    w[0 * n + i] = o.x; w[1 * n + i] = o.y; w[2 * n + i] = o.z;
    w[3 * n + i] = d.x; w[4 * n + i] = d.y; w[5 * n + i] = d.z;
    w[6 * n + i] = odx.x; w[7 * n + i] = odx.y; w[8 * n + i] = odx.z;
    w[9 * n + i] = ody.x; w[10 * n + i] = ody.y; w[11 * n + i] = ody.z;
    w[12 * n + i] = ddx.x; w[13 * n + i] = ddx.y; w[14 * n + i] = ddx.z;
    w[15 * n + i] = ddy.x; w[16 * n + i] = ddy.y; w[17 * n + i] = ddy.z;
    w[18 * n + i] = time;
    w[19 * n + i] = u2f(type);
}

// Camera ray of a path's step 0 (next_ray_for_path, :110-124).
PSY_DEV void camera_ray(const DevCamera& cam, int width, int height, int px, int py, float sx,
                                           float sy, float sa, float sb, float time, V3* o, V3* d, V3* ddx,
                                           V3* ddy, V3* zero) {
    const float min_x = -1.0f, max_x = 1.0f;
    const float max_y = ((float)height) / (float)width;
    const float min_y = -max_y;
    const float samp_x = (fast_logit(sx, 1.5f) + 0.5f + (float)(short)px) / (float)width;
    const float samp_y = (fast_logit(sy, 1.5f) + 0.5f + (float)(short)py) / (float)height;
    const float rx = (float)(((double)samp_x - 0.5) * (double)(max_x - min_x));
    const float ry = (float)((0.5 - (double)samp_y) * (double)(max_y - min_y));
    const float dx = (max_x - min_x) / (float)width;
    const float dy = (max_y - min_y) / (float)height;
    // Camera::generate_ray(rx, ry, dx, dy, time, sa, sb)
    const M44 xf = m_lerp_seq(time, cam.xf, cam.n_xf);
    const float tfov = lerp_seq_f(time, cam.tfov, cam.n_tfov);
    const float ap = lerp_seq_f(time, cam.aperture, cam.n_ap);
    const float fd = lerp_seq_f(time, cam.focus, cam.n_fd);
    float ox = ap * ((sa * 2) - 1);
    float oy = ap * ((sb * 2) - 1);
    square_to_circle(&ox, &oy);
    const V3 lo = mk(ox, oy, 0.0f);
    V3 ld = mk((rx * tfov) - (lo.x / fd), (ry * tfov) - (lo.y / fd), 1.0f);
    ld = normalized(ld);
    *o = m_pos(xf, lo);
    *d = m_dir(xf, ld);
    *zero = m_dir(xf, splat(0.0f));
    *ddx = m_dir(xf, mk(dx * tfov, 0.0f, 0.0f));
    *ddy = m_dir(xf, mk(0.0f, dy * tfov, 0.0f));
}


// ---------------------------------------------------------------------------
// Lambert bounce rays from closest hits
// (integrator/path_trace_integrator.cpp:183-209,
//  shading/surface_closure.hpp:240-326, utils/utils.hpp:214-248,
//  basics/differential_geometry.hpp:20-40, :83-98)
// ---------------------------------------------------------------------------
PSY_DEV void coordinate_system_from_vec3(V3 v1, V3* v2, V3* v3o) {
    if (fabsf(v1.x) > fabsf(v1.y)) {
        const float invlen = 1.f / sqrtf(v1.x * v1.x + v1.z * v1.z);
        *v2 = mk(-v1.z * invlen, 0.f, v1.x * invlen);
    } else {
        const float invlen = 1.f / sqrtf(v1.y * v1.y + v1.z * v1.z);
        *v2 = mk(0.f, v1.z * invlen, -v1.y * invlen);
    }
    *v3o = cross(v1, *v2);
}
PSY_DEV V3 zup_to_vec(V3 from, V3 toz) {
    toz = normalized(toz);
    V3 tox, toy;
    coordinate_system_from_vec3(toz, &tox, &toy);
    return add(add(scale(tox, from.x), scale(toy, from.y)), scale(toz, from.z));
}
PSY_DEV V3 transfer_origin_diff(float t, V3 normal, V3 d, V3 od, V3 dd) {
    const V3 temp = add(od, scale(dd, t));
    const float td = -dot(temp, normal) / dot(d, normal);
    const V3 real_projected = add(temp, scale(d, td));
    return scale(normalized(real_projected), length(temp));
}

// Object-space hit geometry (n, offset) of a closest hit; false for lights
// and misses.
PSY_DEV_FN bool hit_geometry(const DevScene& s, uint32_t li, float t, float u, float v, float aux, V3 wo, V3 wd,
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

}  // namespace psy
