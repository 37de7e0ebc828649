/*
 * ora_scene.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Scene set-up restated from the reference:
 *   object bounds       object/sphere.cpp:51-62, object/bilinear.cpp:19-31,
 *                       object/bicubic (finalize), lights/sphere_light.hpp:20-52,
 *                       lights/rectangle_light.hpp:17-33
 *   instance bounds     scene/assembly.hpp:261-312, basics/bbox.hpp:223-251,
 *                       :629-669 (transform_from)
 *   BVH build           accel/bvh.cpp:14-188 (centroid-midpoint split)
 *   BVH4 collapse       accel/bvh4.cpp:17-159
 *   id bits             scene/assembly.hpp:245-254, utils/utils.hpp:254-287
 */
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include "ora_internal.h"

uint32_t ora_intlog2(uint32_t v) {
    uint32_t r, shift;
    r = (uint32_t)(v > 0xFFFF) << 4; v >>= r;
    shift = (uint32_t)(v > 0xFF) << 3; v >>= shift; r |= shift;
    shift = (uint32_t)(v > 0xF) << 2; v >>= shift; r |= shift;
    shift = (uint32_t)(v > 0x3) << 1; v >>= shift; r |= shift;
    r |= (v >> 1);
    return r;
}

uint32_t ora_upper_pow2(uint32_t v) {
    v--; v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16; v++;
    return v;
}

int ora_kind_floats(int kind) {
    switch (kind) {
        case PSY_OBJ_SPHERE: return 4;
        case PSY_OBJ_BILINEAR: return 12;
        case PSY_OBJ_BICUBIC: return 48;
        case PSY_OBJ_SPHERE_LIGHT: return 7;
        case PSY_OBJ_RECTANGLE_LIGHT: return 5;
        default: return 0;
    }
}

/* Bilinear::bound / Bicubic::bound: min/max over control points */
static bbox bound_points(const float* p, int n) {
    bbox b;
    b.min = v3m(p[0], p[1], p[2]);
    b.max = b.min;
    for (int i = 1; i < n; ++i) {
        v3 q = v3m(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        b.min = v3min(b.min, q);
        b.max = v3max(b.max, q);
    }
    return b;
}

/* SubdivisionSurface::build_bvh_recursive (object/subdivision_surface.cpp:
 * 295-354): split the nodes' centres at the middle of the axis picked by
 * "y if longer than x, then z if longer than y", std::partition order */
typedef struct {
    ora_sd_node* n;
    int count;
} sd_build;

static int sd_center_below(const ora_sd_node* a, int axis, float pmid) {
    const v3 c = v3scale(v3add(a->b.min, a->b.max), 0.5f);
    return v3get(c, axis) < pmid;
}

static int sd_recurse(sd_build* B, int* order, int begin, int end) {
    if (begin + 1 == end) return order[begin];
    bbox cb = bbox_empty();
    for (int i = begin; i < end; ++i) {
        const ora_sd_node* q = &B->n[order[i]];
        const v3 c = v3scale(v3add(q->b.min, q->b.max), 0.5f);
        cb.min = v3m(cb.min.x < c.x ? cb.min.x : c.x, cb.min.y < c.y ? cb.min.y : c.y, cb.min.z < c.z ? cb.min.z : c.z);
        cb.max = v3m(cb.max.x > c.x ? cb.max.x : c.x, cb.max.y > c.y ? cb.max.y : c.y, cb.max.z > c.z ? cb.max.z : c.z);
    }
    const v3 ext = v3sub(cb.max, cb.min);
    int axis = 0;
    if (ext.y > ext.x) axis = 1;
    if (ext.z > ext.y) axis = 2;
    const float pmid = (v3get(cb.min, axis) + v3get(cb.max, axis)) * 0.5f;
    /* std::partition for bidirectional iterators */
    int first = begin, last = end;
    for (;;) {
        while (first != last && sd_center_below(&B->n[order[first]], axis, pmid)) ++first;
        if (first == last) break;
        --last;
        while (first != last && !sd_center_below(&B->n[order[last]], axis, pmid)) --last;
        if (first == last) break;
        const int t = order[first]; order[first] = order[last]; order[last] = t;
        ++first;
    }
    int mid = first;
    if (mid == begin) ++mid;
    const int me = B->count++;
    const int c0 = sd_recurse(B, order, begin, mid);
    const int c1 = sd_recurse(B, order, mid, end);
    B->n[me].child[0] = c0;
    B->n[me].child[1] = c1;
    B->n[me].patch = -1;
    B->n[me].b = bbox_merge(B->n[c0].b, B->n[c1].b);
    return me;
}

static void object_finalize(ora_object* o);

/* SubdivisionSurface::finalize (object/subdivision_surface.cpp:85-233):
 * patches, bounds of the control points, patch BVH (:237-292) */
static void subdiv_finalize(ora_object* o) {
    o->n_bounds = 1;
    o->bounds = (bbox*)malloc(sizeof(bbox));
    bbox bb = bbox_empty();
    for (int i = 0; i < o->n_verts; ++i) {
        const v3 q = v3m(o->data[3 * i], o->data[3 * i + 1], o->data[3 * i + 2]);
        bb.min = v3min(bb.min, q);
        bb.max = v3max(bb.max, q);
    }
    o->bounds[0] = bb;
    o->sd_root = -1;
    if (o->n_patches == 0) return;
    sd_build B;
    B.n = (ora_sd_node*)calloc(2 * (size_t)o->n_patches, sizeof(ora_sd_node));
    B.count = o->n_patches;
    int* order = (int*)malloc(sizeof(int) * (size_t)o->n_patches);
    for (int i = 0; i < o->n_patches; ++i) {
        object_finalize(&o->patches[i]);
        B.n[i].b = o->patches[i].bounds[0];
        B.n[i].child[0] = B.n[i].child[1] = -1;
        B.n[i].patch = i;
        order[i] = i;
    }
    o->sd_root = sd_recurse(&B, order, 0, o->n_patches);
    o->sd_nodes = B.n;
    free(order);
}

static void object_finalize(ora_object* o) {
    if (o->kind == PSY_OBJ_SUBDIVISION) {
        subdiv_finalize(o);
        return;
    }
    const int nf = ora_kind_floats(o->kind);
    if (o->kind == PSY_OBJ_SPHERE_LIGHT) {
        /* positions and radii share the time-sample count here, so the
         * reference's lerp_seq(i/s, radii) reduces to radii[i] */
        o->n_bounds = o->ts;
        o->bounds = (bbox*)malloc(sizeof(bbox) * o->ts);
        for (int t = 0; t < o->ts; ++t) {
            const float* d = o->data + t * nf;
            v3 pos = v3m(d[0], d[1], d[2]);
            v3 rad3 = v3s(d[3]);
            o->bounds[t].min = v3sub(pos, rad3);
            o->bounds[t].max = v3add(pos, rad3);
        }
        return;
    }
    o->n_bounds = o->ts;
    o->bounds = (bbox*)malloc(sizeof(bbox) * o->ts);
    for (int t = 0; t < o->ts; ++t) {
        const float* d = o->data + t * nf;
        bbox b;
        switch (o->kind) {
            case PSY_OBJ_SPHERE:
                b.min = v3m(d[0] - d[3], d[1] - d[3], d[2] - d[3]);
                b.max = v3m(d[0] + d[3], d[1] + d[3], d[2] + d[3]);
                break;
            case PSY_OBJ_BILINEAR:
                b = bound_points(d, 4);
                /* displace_distance = 0 (config.cpp:14) */
                b.min = v3sub(b.min, v3s(0.0f));
                b.max = v3add(b.max, v3s(0.0f));
                break;
            case PSY_OBJ_BICUBIC:
                b = bound_points(d, 16);
                b.min = v3sub(b.min, v3s(0.0f));
                b.max = v3add(b.max, v3s(0.0f));
                break;
            case PSY_OBJ_RECTANGLE_LIGHT: {
                v3 dim3 = v3m(d[0] * 0.5f, d[1] * 0.5f, 0.0f);
                b.min = v3sub(v3s(0.0f), dim3);
                b.max = v3add(v3s(0.0f), dim3);
                break;
            }
            default:
                b = bbox_empty();
        }
        o->bounds[t] = b;
    }
}

/* BBox::inverse_transformed (basics/bbox.hpp:223-251): corners through
 * pos_from, which inverts the matrix for every corner. */
bbox ora_bbox_inverse_transformed(bbox b, const mat4* xf) {
    const v3 vs[8] = {
        v3m(b.min.x, b.min.y, b.min.z), v3m(b.min.x, b.min.y, b.max.z),
        v3m(b.min.x, b.max.y, b.min.z), v3m(b.min.x, b.max.y, b.max.z),
        v3m(b.max.x, b.min.y, b.min.z), v3m(b.max.x, b.min.y, b.max.z),
        v3m(b.max.x, b.max.y, b.min.z), v3m(b.max.x, b.max.y, b.max.z)};
    const mat4 inv = ora_mat4_inverse(xf);
    bbox r = bbox_empty();
    for (int k = 0; k < 8; ++k) {
        v3 v = mat4_pos(&inv, vs[k]);
        for (int i = 0; i < 3; ++i) {
            float vi = v3get(v, i);
            v3set(&r.min, i, vi < v3get(r.min, i) ? vi : v3get(r.min, i));
            v3set(&r.max, i, vi > v3get(r.max, i) ? vi : v3get(r.max, i));
        }
    }
    return r;
}

/* raw bounds of what instance i points at (object bounds / accel bounds) */
static const bbox* inst_raw_bounds(const ora_assembly* a, size_t i, int* n) {
    const psy_instance_desc* in = &a->inst[i];
    if (in->type == PSY_INSTANCE_OBJECT) {
        *n = a->obj[in->data_index].n_bounds;
        return a->obj[in->data_index].bounds;
    }
    const ora_assembly* s = a->sub[in->data_index];
    *n = s->n_bounds;
    return s->bounds;
}

/* Assembly::instance_bounds_at (scene/assembly.hpp:284-305) */
bbox ora_instance_bounds_at(const ora_assembly* a, float t, size_t i) {
    int n;
    const bbox* bs = inst_raw_bounds(a, i, &n);
    bbox bb = bbox_lerp_seq(t, bs, n);
    const psy_instance_desc* in = &a->inst[i];
    if (in->transform_count > 0) {
        mat4 x = mat4_lerp_seq(t, a->xf + in->transform_index, in->transform_count);
        bb = ora_bbox_inverse_transformed(bb, &x);
    }
    return bb;
}

/* Assembly::instance_bounds + transform_from (basics/bbox.hpp:641-669) */
bbox* ora_instance_bounds_vec(const ora_assembly* a, size_t i, int* out_n) {
    int n;
    const bbox* bs = inst_raw_bounds(a, i, &n);
    const psy_instance_desc* in = &a->inst[i];
    const int tc = in->transform_count;
    const mat4* xs = a->xf + in->transform_index;
    bbox* r;
    if (tc == 0) {
        r = (bbox*)malloc(sizeof(bbox) * n);
        memcpy(r, bs, sizeof(bbox) * n);
        *out_n = n;
    } else if (n == tc) {
        r = (bbox*)malloc(sizeof(bbox) * n);
        for (int k = 0; k < n; ++k) r[k] = ora_bbox_inverse_transformed(bs[k], &xs[k]);
        *out_n = n;
    } else if (n > tc) {
        const float s = (float)(n - 1);
        r = (bbox*)malloc(sizeof(bbox) * n);
        for (int k = 0; k < n; ++k) {
            mat4 x = mat4_lerp_seq((float)k / s, xs, tc);
            r[k] = ora_bbox_inverse_transformed(bs[k], &x);
        }
        *out_n = n;
    } else {
        const float s = (float)(tc - 1);
        r = (bbox*)malloc(sizeof(bbox) * tc);
        for (int k = 0; k < tc; ++k)
            r[k] = ora_bbox_inverse_transformed(bbox_lerp_seq((float)k / s, bs, n), &xs[k]);
        *out_n = tc;
    }
    return r;
}

/* ---------------- BVH (accel/bvh.cpp) ------------------------------------ */
typedef struct { size_t inst; v3 bmin, bmax, c; } bvh_prim;

typedef struct {
    const ora_assembly* a;
    bvh_prim* bag;
    ora_bvh_node* nodes; size_t n_nodes, cap_nodes;
    bbox* bboxes; size_t n_bb, cap_bb;
} bvh_builder;

#define BVH_IS_LEAF 1u
#define BVH4_IS_SKIP (1u << 8)
#define BVH4_IS_2ND (1u << 9)
#define BVH4_IS_3RD (1u << 10)
#define BVH4_IS_4TH (1u << 11)

static size_t push_node(bvh_builder* b) {
    if (b->n_nodes == b->cap_nodes) {
        b->cap_nodes = b->cap_nodes ? b->cap_nodes * 2 : 64;
        b->nodes = (ora_bvh_node*)realloc(b->nodes, sizeof(ora_bvh_node) * b->cap_nodes);
    }
    memset(&b->nodes[b->n_nodes], 0, sizeof(ora_bvh_node));
    return b->n_nodes++;
}
static void push_bbox(bvh_builder* b, bbox x) {
    if (b->n_bb == b->cap_bb) {
        b->cap_bb = b->cap_bb ? b->cap_bb * 2 : 64;
        b->bboxes = (bbox*)realloc(b->bboxes, sizeof(bbox) * b->cap_bb);
    }
    b->bboxes[b->n_bb++] = x;
}

/* std::partition (libstdc++ bidirectional variant) on bag[first, last) by
 * c[axis] < mid; returns the split point. */
static size_t partition_prims(bvh_prim* bag, size_t first, size_t last, int axis, float mid) {
    for (;;) {
        for (;;) {
            if (first == last) return first;
            if (v3get(bag[first].c, axis) < mid) ++first; else break;
        }
        --last;
        for (;;) {
            if (first == last) return first;
            if (!(v3get(bag[last].c, axis) < mid)) --last; else break;
        }
        bvh_prim t = bag[first]; bag[first] = bag[last]; bag[last] = t;
        ++first;
    }
}

/* BVH::split_primitives (accel/bvh.cpp:89-122) */
static size_t split_primitives(bvh_builder* b, size_t first, size_t last) {
    v3 mn = b->bag[first].c, mx = b->bag[first].c;
    for (size_t i = first + 1; i <= last; ++i) {
        for (int d = 0; d < 3; ++d) {
            const float ci = v3get(b->bag[i].c, d);
            v3set(&mn, d, v3get(mn, d) < ci ? v3get(mn, d) : ci);
            v3set(&mx, d, v3get(mx, d) > ci ? v3get(mx, d) : ci);
        }
    }
    int axis = 0;
    if ((mx.y - mn.y) > (mx.x - mn.x)) axis = 1;
    if ((mx.z - mn.z) > (mx.y - mn.y)) axis = 2;
    const float pmid = .5f * (v3get(mn, axis) + v3get(mx, axis));
    size_t split = partition_prims(b->bag, first, last + 1, axis, pmid);
    if (split > 0) split--;
    if (split < first) split = first;
    return split;
}

/* BVH::recursive_build (accel/bvh.cpp:130-188) */
static size_t recursive_build(bvh_builder* b, size_t parent, size_t first, size_t last) {
    const size_t me = push_node(b);
    b->nodes[me].flags = 0;
    b->nodes[me].parent_index = parent;
    if (first == last) {
        b->nodes[me].flags |= BVH_IS_LEAF;
        b->nodes[me].child_or_data = b->bag[first].inst;
        int n;
        bbox* bbs = ora_instance_bounds_vec(b->a, b->bag[first].inst, &n);
        b->nodes[me].bbox_index = b->n_bb;
        b->nodes[me].ts = (uint16_t)n;
        for (int k = 0; k < n; ++k) push_bbox(b, bbs[k]);
        free(bbs);
    } else {
        const size_t split = split_primitives(b, first, last);
        const size_t c1 = recursive_build(b, me, first, split);
        const size_t c2 = recursive_build(b, me, split + 1, last);
        b->nodes[me].child_or_data = c2;
        b->nodes[me].bbox_index = b->n_bb;
        if (b->nodes[c1].ts == b->nodes[c2].ts) {
            b->nodes[me].ts = b->nodes[c1].ts;
            for (size_t i = 0; i < b->nodes[me].ts; ++i) {
                bbox x = bbox_merge(b->bboxes[b->nodes[c1].bbox_index + i],
                                    b->bboxes[b->nodes[c2].bbox_index + i]);
                push_bbox(b, x);
            }
        } else {
            b->nodes[me].ts = 1;
            bbox x = b->bboxes[b->nodes[c1].bbox_index];
            for (size_t i = 1; i < b->nodes[c1].ts; ++i)
                x = bbox_merge(x, b->bboxes[b->nodes[c1].bbox_index + i]);
            for (size_t i = 0; i < b->nodes[c2].ts; ++i)
                x = bbox_merge(x, b->bboxes[b->nodes[c2].bbox_index + i]);
            push_bbox(b, x);
        }
    }
    return me;
}

typedef struct { ora_bvh4_node* v; size_t n, cap; } node4_vec;
static ora_bvh4_node* n4_push(node4_vec* nv) {
    if (nv->n == nv->cap) {
        nv->cap = nv->cap ? nv->cap * 2 : 64;
        nv->v = (ora_bvh4_node*)realloc(nv->v, sizeof(ora_bvh4_node) * nv->cap);
    }
    ora_bvh4_node* n = &nv->v[nv->n++];
    memset(n, 0, sizeof *n);
    const bbox e = bbox_empty();
    n->bounds = bbox4_from(e, e, e, e);
    return n;
}

/* BVH4::build (accel/bvh4.cpp:17-159) */
static void bvh4_build(ora_assembly* a) {
    bvh_builder b;
    memset(&b, 0, sizeof b);
    b.a = a;
    a->nodes = NULL; a->n_nodes = 0;
    a->n_bounds = 1;
    a->bounds = (bbox*)malloc(sizeof(bbox));
    a->bounds[0] = bbox_empty();
    if (a->n_inst == 0) return;
    b.bag = (bvh_prim*)malloc(sizeof(bvh_prim) * a->n_inst);
    for (int i = 0; i < a->n_inst; ++i) {
        bbox bb = ora_instance_bounds_at(a, 0.5f, (size_t)i);
        b.bag[i].inst = (size_t)i;
        b.bag[i].bmin = bb.min;
        b.bag[i].bmax = bb.max;
        b.bag[i].c = lerpv(0.5f, bb.min, bb.max);
    }
    recursive_build(&b, 0, 0, (size_t)a->n_inst - 1);
    free(b.bag);

    node4_vec nv = {0, 0, 0};
    n4_push(&nv);
    for (size_t bni = 0; bni < b.n_nodes; ++bni) {
        ora_bvh_node* bn = &b.nodes[bni];
        const size_t ni = nv.n - 1;
        if (bn->flags & BVH4_IS_SKIP) continue;
        if (bn->flags & BVH4_IS_2ND) nv.v[bn->parent_index].child[0] = ni;
        else if (bn->flags & BVH4_IS_3RD) nv.v[bn->parent_index].child[1] = ni;
        else if (bn->flags & BVH4_IS_4TH) nv.v[bn->parent_index].child[2] = ni;
        if (bn->flags & BVH_IS_LEAF) {
            nv.v[ni].child[0] = 0;
            nv.v[ni].data_index = bn->child_or_data;
            n4_push(&nv);
        } else {
            int cc = 0;
            ora_bvh_node* ch[4] = {0, 0, 0, 0};
            const size_t ci1 = bni + 1, ci2 = bn->child_or_data;
            if (b.nodes[ci1].flags & BVH_IS_LEAF) {
                ch[cc++] = &b.nodes[ci1];
            } else {
                b.nodes[ci1].flags |= BVH4_IS_SKIP;
                ch[cc++] = &b.nodes[ci1 + 1];
                ch[cc++] = &b.nodes[b.nodes[ci1].child_or_data];
            }
            if (b.nodes[ci2].flags & BVH_IS_LEAF) {
                ch[cc++] = &b.nodes[ci2];
            } else {
                b.nodes[ci2].flags |= BVH4_IS_SKIP;
                ch[cc++] = &b.nodes[ci2 + 1];
                ch[cc++] = &b.nodes[b.nodes[ci2].child_or_data];
            }
            if (ch[1]) ch[1]->flags |= BVH4_IS_2ND;
            if (ch[2]) ch[2]->flags |= BVH4_IS_3RD;
            if (ch[3]) ch[3]->flags |= BVH4_IS_4TH;
            for (int i = 0; i < cc; ++i) ch[i]->parent_index = ni;
            int equal = 1;
            int most = ch[0]->ts;
            for (int i = 1; i < cc; ++i) {
                equal = equal && (ch[i - 1]->ts == ch[i]->ts);
                if (ch[i]->ts > most) most = ch[i]->ts;
            }
            const bbox e = bbox_empty();
            if (equal) {
                nv.v[ni].ts = ch[0]->ts;
                for (int i = 0; i < ch[0]->ts; ++i) {
                    bbox bb[4] = {e, e, e, e};
                    for (int c = 0; c < cc; ++c) bb[c] = b.bboxes[ch[c]->bbox_index + i];
                    nv.v[nv.n - 1].bounds = bbox4_from(bb[0], bb[1], bb[2], bb[3]);
                    n4_push(&nv);
                }
            } else {
                nv.v[ni].ts = (uint32_t)most;
                const float s = (float)(most - 1);
                for (int i = 0; i < most; ++i) {
                    bbox bb[4] = {e, e, e, e};
                    for (int c = 0; c < cc; ++c) {
                        if (ch[c]->ts == most) bb[c] = b.bboxes[ch[c]->bbox_index + i];
                        else bb[c] = bbox_lerp_seq((float)i / s, b.bboxes + ch[c]->bbox_index, ch[c]->ts);
                    }
                    nv.v[nv.n - 1].bounds = bbox4_from(bb[0], bb[1], bb[2], bb[3]);
                    n4_push(&nv);
                }
            }
        }
    }
    nv.n -= 1;  /* nodes.pop_back() */
    a->nodes = nv.v;
    a->n_nodes = nv.n;
    free(a->bounds);
    a->n_bounds = b.nodes[0].ts;
    a->bounds = (bbox*)malloc(sizeof(bbox) * a->n_bounds);
    memcpy(a->bounds, b.bboxes + b.nodes[0].bbox_index, sizeof(bbox) * a->n_bounds);
    free(b.nodes);
    free(b.bboxes);
}

void ora_light_tree_build(ora_assembly* a);

/* Assembly::finalize (scene/assembly.hpp:206-238): children first */
static void assembly_finalize(ora_assembly* a) {
    if (a->finalized) return;
    for (int i = 0; i < a->n_sub; ++i) assembly_finalize(a->sub[i]);
    for (int i = 0; i < a->n_obj; ++i) object_finalize(&a->obj[i]);
    bvh4_build(a);
    ora_light_tree_build(a);
    a->finalized = 1;
}

static mat4 load_mat(const float* f) {
    mat4 m;
    memcpy(&m.m[0][0], f, sizeof(float) * 16);
    return m;
}

ora_scene* ora_scene_create(const psy_scene_desc* d) {
    if (!d || d->n_assemblies < 1) return NULL;
    int ok = 1;
    ora_scene* s = (ora_scene*)calloc(1, sizeof(ora_scene));
    s->n_asm = d->n_assemblies;
    s->asms = (ora_assembly*)calloc((size_t)s->n_asm, sizeof(ora_assembly));
    s->flags = d->flags;
    s->n_shaders = d->n_shaders;
    s->shaders = (psy_shader_desc*)malloc(sizeof(psy_shader_desc) * (d->n_shaders + 1));
    if (d->n_shaders) memcpy(s->shaders, d->shaders, sizeof(psy_shader_desc) * d->n_shaders);
    for (int k = 0; k < s->n_asm; ++k) {
        const psy_assembly_desc* ad = &d->assemblies[k];
        ora_assembly* a = &s->asms[k];
        a->scene = s;
        a->n_inst = ad->n_instances;
        a->inst = (psy_instance_desc*)malloc(sizeof(psy_instance_desc) * (ad->n_instances + 1));
        if (ad->n_instances) memcpy(a->inst, ad->instances, sizeof(psy_instance_desc) * ad->n_instances);
        a->n_xf = ad->n_transforms;
        a->xf = (mat4*)malloc(sizeof(mat4) * (ad->n_transforms + 1));
        for (int i = 0; i < ad->n_transforms; ++i) a->xf[i] = load_mat(ad->transforms + 16 * i);
        a->n_obj = ad->n_objects;
        a->obj = (ora_object*)calloc((size_t)ad->n_objects + 1, sizeof(ora_object));
        for (int i = 0; i < ad->n_objects; ++i) {
            const psy_object_desc* od = &ad->objects[i];
            ora_object* o = &a->obj[i];
            o->kind = od->kind;
            o->escale = 1.0f;
            o->ts = od->time_samples;
            if (od->kind == PSY_OBJ_SUBDIVISION) {
                float* pp = NULL;
                const int64_t np = od->time_samples == 1 ? ora_subdiv_patches(od->data, od->n_vertices,
                    od->face_vert_counts, od->n_faces, od->face_vert_indices, &pp) : -1;
                if (np < 0) { ok = 0; continue; }
                o->n_verts = od->n_vertices;
                o->data = (float*)malloc(sizeof(float) * (3 * (size_t)od->n_vertices + 1));
                memcpy(o->data, od->data, sizeof(float) * 3 * (size_t)od->n_vertices);
                o->n_patches = (int)np;
                o->patches = (ora_object*)calloc((size_t)np + 1, sizeof(ora_object));
                for (int64_t k = 0; k < np; ++k) {
                    ora_object* q = &o->patches[k];
                    q->kind = PSY_OBJ_BICUBIC;
                    q->ts = 1;
                    q->escale = 1.0f;
                    q->data = (float*)malloc(sizeof(float) * 48);
                    memcpy(q->data, pp + 48 * k, sizeof(float) * 48);
                }
                ora_free(pp);
                continue;
            }
            const int nf = ora_kind_floats(od->kind) * od->time_samples;
            o->data = (float*)malloc(sizeof(float) * (nf + 1));
            memcpy(o->data, od->data, sizeof(float) * nf);
        }
        a->n_sub = ad->n_assemblies;
        a->sub = (ora_assembly**)malloc(sizeof(ora_assembly*) * (ad->n_assemblies + 1));
        for (int i = 0; i < ad->n_assemblies; ++i) a->sub[i] = &s->asms[ad->assemblies[i]];
    }
    /* camera (basics/camera.hpp:23-69): tfov = sin(fov/2)/cos(fov/2) */
    const psy_camera_desc* c = &d->camera;
    s->n_cam_xf = c->n_transforms;
    s->cam_xf = (mat4*)malloc(sizeof(mat4) * (c->n_transforms + 1));
    for (int i = 0; i < c->n_transforms; ++i) s->cam_xf[i] = load_mat(c->transforms + 16 * i);
    s->n_tfov = c->n_fovs;
    s->tfovs = (float*)malloc(sizeof(float) * (c->n_fovs + 1));
    for (int i = 0; i < c->n_fovs; ++i) {
        const float f = c->fovs[i];
        s->tfovs[i] = (float)(sin((double)(f / 2)) / cos((double)(f / 2)));
    }
    if (c->n_aperture_radii == 0 || c->n_focus_distances == 0) {
        s->n_ap = 1; s->aperture = (float*)malloc(sizeof(float)); s->aperture[0] = 0.0f;
        s->n_fd = 1; s->focus = (float*)malloc(sizeof(float)); s->focus[0] = 1.0f;
    } else {
        s->n_ap = c->n_aperture_radii;
        s->aperture = (float*)malloc(sizeof(float) * s->n_ap);
        memcpy(s->aperture, c->aperture_radii, sizeof(float) * s->n_ap);
        s->n_fd = c->n_focus_distances;
        s->focus = (float*)malloc(sizeof(float) * s->n_fd);
        memcpy(s->focus, c->focus_distances, sizeof(float) * s->n_fd);
        for (int i = 0; i < s->n_fd; ++i) {
            if (s->focus[i] <= 0.0f) {
                s->n_ap = 1; s->aperture[0] = 0.0f;
                s->n_fd = 1; s->focus[0] = 1.0f;
                break;
            }
        }
    }
    memcpy(s->background, d->background, sizeof(float) * 3);
    s->dice_rate = d->dice_rate;
    s->min_upoly = d->min_upoly_size;
    if (!ok) {
        ora_scene_destroy(s);
        return NULL;
    }
    assembly_finalize(&s->asms[0]);
    for (int k = 1; k < s->n_asm; ++k) assembly_finalize(&s->asms[k]);
    return s;
}

void ora_scene_destroy(ora_scene* s) {
    if (!s) return;
    for (int k = 0; k < s->n_asm; ++k) {
        ora_assembly* a = &s->asms[k];
        for (int i = 0; i < a->n_obj; ++i) {
            ora_object* o = &a->obj[i];
            for (int k = 0; k < o->n_patches; ++k) { free(o->patches[k].data); free(o->patches[k].bounds); }
            free(o->patches); free(o->sd_nodes);
            free(o->data); free(o->bounds);
        }
        for (size_t i = 0; i < a->n_lnodes; ++i) free(a->lnodes[i].bounds);
        free(a->lnodes); free(a->lbounds);
        free(a->obj); free(a->inst); free(a->xf); free(a->sub); free(a->nodes); free(a->bounds);
    }
    free(s->asms); free(s->shaders); free(s->cam_xf); free(s->tfovs); free(s->aperture); free(s->focus);
    free(s);
}

int64_t ora_scene_bvh4_nodes(const ora_scene* s, int assembly) {
    if (!s || assembly < 0 || assembly >= s->n_asm) return -1;
    return (int64_t)s->asms[assembly].n_nodes;
}

/* Differentiable parameters (the product's psy_scene_set_params): grey
 * albedo per shader (Lambert shaders only) and an emission scale per light
 * object in scene order.  Light trees keep the energies of the scene as
 * created. */
int ora_scene_set_params(ora_scene* s, const float* albedo, int n_albedo, const float* escale, int n_escale) {
    if (albedo) {
        if (n_albedo != s->n_shaders) return -1;
        for (int i = 0; i < n_albedo; ++i)
            if (s->shaders[i].kind == PSY_SHADER_LAMBERT)
                for (int k = 0; k < 3; ++k) s->shaders[i].color[k] = albedo[i];
    }
    if (escale) {
        int j = 0;
        for (int a = 0; a < s->n_asm; ++a)
            for (int i = 0; i < s->asms[a].n_obj; ++i) {
                ora_object* o = &s->asms[a].obj[i];
                if (o->kind != PSY_OBJ_SPHERE_LIGHT && o->kind != PSY_OBJ_RECTANGLE_LIGHT) continue;
                if (j >= n_escale) return -1;
                o->escale = escale[j++];
            }
        if (j != n_escale) return -1;
    }
    return 0;
}
