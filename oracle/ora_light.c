/*
 * ora_light.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * LightTree::build restated (accel/light_tree.cpp:99-238): a binary tree
 * over the emitting instances of an assembly, centroid-midpoint splits,
 * node energy = sum of child energies, node bounds = time-sample merge.
 */
#include <stdlib.h>
#include <string.h>
#include "ora_internal.h"

typedef struct { size_t inst; v3 center; bbox bb; float energy; } lbuild;

/* Color::energy (color/color.hpp:369-375) */
static float color_energy(const float* c) {
    float acc = 0.0f;
    for (int i = 0; i < 3; ++i) acc += c[i];
    acc /= 3.0f;
    return acc;
}

static float object_emitted_energy(const ora_object* ob) {
    /* total_emitted_color = lerp_seq(0, colors) = colors[0] for lights */
    if (ob->kind == PSY_OBJ_SPHERE_LIGHT) return color_energy(ob->data + 4);
    if (ob->kind == PSY_OBJ_RECTANGLE_LIGHT) return color_energy(ob->data + 2);
    return 0.0f;
}

/* Transform::get_inv_scale (math/transform.hpp:104-111) */
static v3 inv_scale(const mat4* m) {
    v3 s;
    s.x = sqrtf((m->m[0][0] * m->m[0][0]) + (m->m[0][1] * m->m[0][1]) + (m->m[0][2] * m->m[0][2]));
    s.y = sqrtf((m->m[1][0] * m->m[1][0]) + (m->m[1][1] * m->m[1][1]) + (m->m[1][2] * m->m[1][2]));
    s.z = sqrtf((m->m[2][0] * m->m[2][0]) + (m->m[2][1] * m->m[2][1]) + (m->m[2][2] * m->m[2][2]));
    return s;
}

static mat4 instance_xform_at(const ora_assembly* a, float t, size_t i) {
    const psy_instance_desc* in = &a->inst[i];
    if (in->transform_count > 0) return mat4_lerp_seq(t, a->xf + in->transform_index, in->transform_count);
    return mat4_identity();
}

/* merge(vector<BBox>, vector<BBox>) (basics/bbox.hpp:603-626) */
static bbox* merge_bounds(const bbox* a, int na, const bbox* b, int nb, int* nout) {
    bbox* c;
    if (na == 0) { c = (bbox*)malloc(sizeof(bbox) * (nb + 1)); memcpy(c, b, sizeof(bbox) * nb); *nout = nb; }
    else if (nb == 0) { c = (bbox*)malloc(sizeof(bbox) * (na + 1)); memcpy(c, a, sizeof(bbox) * na); *nout = na; }
    else if (na == nb) {
        c = (bbox*)malloc(sizeof(bbox) * na);
        for (int i = 0; i < na; ++i) c[i] = bbox_merge(a[i], b[i]);
        *nout = na;
    } else if (na > nb) {
        const float s = (float)(na - 1);
        c = (bbox*)malloc(sizeof(bbox) * na);
        for (int i = 0; i < na; ++i) c[i] = bbox_merge(a[i], bbox_lerp_seq((float)i / s, b, nb));
        *nout = na;
    } else {
        const float s = (float)(nb - 1);
        c = (bbox*)malloc(sizeof(bbox) * nb);
        for (int i = 0; i < nb; ++i) c[i] = bbox_merge(b[i], bbox_lerp_seq((float)i / s, a, na));
        *nout = nb;
    }
    return c;
}

typedef struct { ora_ltree_node* v; size_t n, cap; } lnode_vec;

static size_t lpush(lnode_vec* nv) {
    if (nv->n == nv->cap) {
        nv->cap = nv->cap ? nv->cap * 2 : 16;
        nv->v = (ora_ltree_node*)realloc(nv->v, sizeof(ora_ltree_node) * nv->cap);
    }
    memset(&nv->v[nv->n], 0, sizeof(ora_ltree_node));
    return nv->n++;
}

/* split_lights (accel/light_tree.cpp:159-188) with libstdc++ partition */
static size_t split_lights(lbuild* v, size_t start, size_t end) {
    v3 mn = v[start].center, mx = v[start].center;
    for (size_t i = start + 1; i < end; ++i) {
        for (int d = 0; d < 3; ++d) {
            const float c = v3get(v[i].center, d);
            v3set(&mn, d, v3get(mn, d) < c ? v3get(mn, d) : c);
            v3set(&mx, d, v3get(mx, d) > c ? v3get(mx, d) : c);
        }
    }
    int axis = 0;
    if ((mx.y - mn.y) > (mx.x - mn.x)) axis = 1;
    if ((mx.z - mn.z) > (mx.y - mn.y)) axis = 2;
    const float pmid = .5f * (v3get(mn, axis) + v3get(mx, axis));
    size_t first = start, last = end;
    for (;;) {
        for (;;) {
            if (first == last) goto done;
            if (v3get(v[first].center, axis) < pmid) ++first; else break;
        }
        --last;
        for (;;) {
            if (first == last) goto done;
            if (!(v3get(v[last].center, axis) < pmid)) --last; else break;
        }
        { lbuild t = v[first]; v[first] = v[last]; v[last] = t; }
        ++first;
    }
done:
    if (first == start) first++;
    return first;
}

bbox* ora_instance_bounds_vec(const ora_assembly* a, size_t i, int* n);

static size_t lrecursive(ora_assembly* a, lnode_vec* nv, lbuild* v, size_t start, size_t end) {
    const size_t me = lpush(nv);
    if (start + 1 == end) {
        const psy_instance_desc* in = &a->inst[v[start].inst];
        nv->v[me].is_leaf = 1;
        nv->v[me].instance_index = v[start].inst;
        if (in->type == PSY_INSTANCE_OBJECT) {
            nv->v[me].bounds = ora_instance_bounds_vec(a, v[start].inst, &nv->v[me].n_bounds);
        } else {
            const ora_assembly* s = a->sub[in->data_index];
            int n = s->n_lbounds;
            bbox* bs = (bbox*)malloc(sizeof(bbox) * (n + 1));
            if (in->transform_count > 0) {
                /* transform_from(sub light bounds, xforms) */
                const mat4* xs = a->xf + in->transform_index;
                const int tc = in->transform_count;
                if (n == tc) for (int k = 0; k < n; ++k) bs[k] = ora_bbox_inverse_transformed(s->lbounds[k], &xs[k]);
                else if (n > tc) {
                    const float sc = (float)(n - 1);
                    for (int k = 0; k < n; ++k) {
                        mat4 x = mat4_lerp_seq((float)k / sc, xs, tc);
                        bs[k] = ora_bbox_inverse_transformed(s->lbounds[k], &x);
                    }
                } else {
                    free(bs);
                    bs = (bbox*)malloc(sizeof(bbox) * tc);
                    const float sc = (float)(tc - 1);
                    for (int k = 0; k < tc; ++k)
                        bs[k] = ora_bbox_inverse_transformed(bbox_lerp_seq((float)k / sc, s->lbounds, n), &xs[k]);
                    n = tc;
                }
            } else {
                memcpy(bs, s->lbounds, sizeof(bbox) * n);
            }
            nv->v[me].bounds = bs;
            nv->v[me].n_bounds = n;
        }
        nv->v[me].energy = v[start].energy;
    } else {
        nv->v[me].is_leaf = 0;
        const size_t split = split_lights(v, start, end);
        const size_t c1 = lrecursive(a, nv, v, start, split);
        const size_t c2 = lrecursive(a, nv, v, split, end);
        nv->v[me].index1 = c1;
        nv->v[me].index2 = c2;
        nv->v[me].bounds = merge_bounds(nv->v[c1].bounds, nv->v[c1].n_bounds, nv->v[c2].bounds,
                                        nv->v[c2].n_bounds, &nv->v[me].n_bounds);
        nv->v[me].energy = nv->v[c1].energy + nv->v[c2].energy;
    }
    return me;
}

void ora_light_tree_build(ora_assembly* a) {
    lbuild* v = (lbuild*)malloc(sizeof(lbuild) * (a->n_inst + 1));
    size_t n = 0;
    a->total_lights = 0;
    for (int i = 0; i < a->n_inst; ++i) {
        const psy_instance_desc* in = &a->inst[i];
        if (in->type == PSY_INSTANCE_OBJECT) {
            const float e = object_emitted_energy(&a->obj[in->data_index]);
            if (e > 0.0f) {
                v[n].inst = (size_t)i;
                v[n].bb = ora_instance_bounds_at(a, 0.5f, (size_t)i);
                v[n].center = v3scale(v3add(v[n].bb.min, v[n].bb.max), 0.5f);
                mat4 x = instance_xform_at(a, 0.5f, (size_t)i);
                const v3 sc = inv_scale(&x);
                const float ss = ((sc.x * sc.y) + (sc.x * sc.z) + (sc.y * sc.z)) * 0.33333333f;
                v[n].energy = e / ss;
                ++n;
                a->total_lights++;
            }
        } else {
            const ora_assembly* s = a->sub[in->data_index];
            const size_t count = s->total_lights;
            const float te[3] = {s->total_energy, s->total_energy, s->total_energy};
            const float e = color_energy(te);
            if (count > 0 && e > 0.0f) {
                v[n].inst = (size_t)i;
                bbox lb = bbox_lerp_seq(0.5f, s->lbounds, s->n_lbounds);
                if (in->transform_count > 0) {
                    mat4 x = mat4_lerp_seq(0.5f, a->xf + in->transform_index, in->transform_count);
                    lb = ora_bbox_inverse_transformed(lb, &x);
                }
                v[n].bb = lb;
                v[n].center = v3scale(v3add(lb.min, lb.max), 0.5f);
                mat4 x = instance_xform_at(a, 0.5f, (size_t)i);
                const v3 sc = inv_scale(&x);
                const float ss = ((sc.x * sc.y) + (sc.x * sc.z) + (sc.y * sc.z)) * 0.33333333f;
                v[n].energy = e / ss;
                ++n;
                a->total_lights += count;
            }
        }
    }
    lnode_vec nv = {0, 0, 0};
    if (n > 0) {
        lrecursive(a, &nv, v, 0, n);
        a->n_lbounds = nv.v[0].n_bounds;
        a->lbounds = (bbox*)malloc(sizeof(bbox) * a->n_lbounds);
        memcpy(a->lbounds, nv.v[0].bounds, sizeof(bbox) * a->n_lbounds);
        a->total_energy = nv.v[0].energy;
    } else {
        a->n_lbounds = 1;
        a->lbounds = (bbox*)malloc(sizeof(bbox));
        a->lbounds[0] = bbox_empty();
        a->total_energy = 0.0f;
    }
    a->lnodes = nv.v;
    a->n_lnodes = nv.n;
    free(v);
}
