/*
 * ora_trace.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Breadth-first ray-stream tracer restated from the reference:
 *   Tracer::trace / trace_assembly / trace_surface / trace_lightsource
 *                                tracer/tracer.cpp:38-313
 *   BVH4StreamTraverser::next_object   accel/bvh4.cpp:163-236
 *   mutable_partition                  utils/utils.hpp:122-146
 *   intersect_rays_with_patch<PATCH>   object/patch_utils.hpp:19-200
 *   Bilinear / Bicubic operations      object/bilinear.hpp:44-141,
 *                                      object/bicubic.hpp:55-189
 *   Sphere::intersect_ray              object/sphere.cpp:67-189
 *   SphereLight / RectangleLight       lights/sphere_light.hpp:124-246,
 *                                      lights/rectangle_light.hpp:76-147
 *   WorldRay::update_ray, Ray widths   basics/ray.hpp:122-153, :207-323
 *   closest_ray_t2                     utils/utils.hpp:443-456
 *   merge(Transform)                   math/transform.hpp:241-272
 *   InstanceID                         basics/instance_id.hpp:9-40
 */
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#ifdef ORA_DEBUG
#include <stdio.h>
#endif
#include "ora_internal.h"

#ifndef M_PI
#define M_PI 3.1415926535897932384626433
#endif

/* ---------------- rays --------------------------------------------------- */
static inline v3 va(const float* f) { return v3m(f[0], f[1], f[2]); }
static inline void set3(float* d, v3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

/* closest_ray_t2 (utils/utils.hpp:443-456) */
static void closest_ray_t2(v3 o1, v3 d1, v3 o2, v3 d2, float* t_out, float* dist_out) {
    const v3 dd = v3sub(d1, d2);
    const float dd2 = v3dot(dd, dd);
    if (dd2 < 0.00001f) { *t_out = -1.0f; *dist_out = -1.0f; return; }
    const v3 w = v3sub(o1, o2);
    const float t = -v3dot(w, dd) / dd2;
    *t_out = t;
    *dist_out = v3len(v3sub(v3add(o1, v3scale(d1, t)), v3add(o2, v3scale(d2, t))));
}

static void set_widths(ora_ray* r, v3 odx, v3 ody, v3 ddx, v3 ddy) {
    const v3 orx = v3add(r->o, odx), drx = v3add(r->d, ddx);
    const v3 ory = v3add(r->o, ody), dry = v3add(r->d, ddy);
    float tdx, lx, tdy, ly;
    closest_ray_t2(r->o, r->d, orx, drx, &tdx, &lx);
    closest_ray_t2(r->o, r->d, ory, dry, &tdy, &ly);
    r->owx = v3len(odx);
    if (tdx <= 0.0f) { r->dwx = v3len(ddx); r->fwx = 0.0f; }
    else { r->dwx = (lx - r->owx) / tdx; r->fwx = lx; }
    r->owy = v3len(ody);
    if (tdy <= 0.0f) { r->dwy = v3len(ddy); r->fwy = 0.0f; }
    else { r->dwy = (ly - r->owy) / tdy; r->fwy = ly; }
    r->d_inv = v3div(v3s(1.0f), r->d);   /* Ray::finalize -> update_accel */
}

/* WorldRay::update_ray(Ray*) */
void ora_update_ray(const ora_wray* w, ora_ray* r) {
    r->o = va(w->o);
    r->d = va(w->d);
    set_widths(r, va(w->odx), va(w->ody), va(w->ddx), va(w->ddy));
}

/* WorldRay::update_ray(Ray*, const Transform&) */
void ora_update_ray_xf(const ora_wray* w, ora_ray* r, const mat4* xf) {
    r->o = mat4_pos(xf, va(w->o));
    r->d = mat4_dir(xf, va(w->d));
    set_widths(r, mat4_dir(xf, va(w->odx)), mat4_dir(xf, va(w->ody)),
               mat4_dir(xf, va(w->ddx)), mat4_dir(xf, va(w->ddy)));
}

/* WorldRay::to_ray */
static void to_ray(const ora_wray* w, ora_ray* r) {
    memset(r, 0, sizeof *r);
    r->time = w->time;
    r->max_t = INFINITY;
    if (w->type == PSY_RAY_OCCLUSION) {
        r->max_t = 1.0f;
        r->id_and_flags |= RAY_OCCL;
    }
    ora_update_ray(w, r);
}

/* Ray::min_width (basics/ray.hpp:132-153) */
float ora_ray_min_width(const ora_ray* r, float tnear, float tfar) {
    const float tflipx = (r->owx - r->fwx) / r->dwx;
    const float tflipy = (r->owy - r->fwy) / r->dwy;
    float minx, miny;
    if (tnear < tflipx && tfar > tflipx) minx = r->fwx;
    else minx = smin(fabsf((r->owx - r->fwx) + (r->dwx * tnear)) + r->fwx,
                     fabsf((r->owx - r->fwx) + (r->dwx * tfar)) + r->fwx);
    if (tnear < tflipy && tfar > tflipy) miny = r->fwy;
    else miny = smin(fabsf((r->owy - r->fwy) + (r->dwy * tnear)) + r->fwy,
                     fabsf((r->owy - r->fwy) + (r->dwy * tfar)) + r->fwy);
    return smin(minx, miny);
}

void ora_world_to_rays(const ora_wray* w, int64_t n, ora_ray_out* out) {
    for (int64_t i = 0; i < n; ++i) {
        ora_ray r;
        to_ray(&w[i], &r);
        ora_ray_out* o = &out[i];
        o->o[0] = r.o.x; o->o[1] = r.o.y; o->o[2] = r.o.z;
        o->d[0] = r.d.x; o->d[1] = r.d.y; o->d[2] = r.d.z;
        o->max_t = r.max_t; o->time = r.time;
        o->owx = r.owx; o->owy = r.owy; o->dwx = r.dwx; o->dwy = r.dwy; o->fwx = r.fwx; o->fwy = r.fwy;
        o->flags = r.id_and_flags;
    }
}

/* ---------------- KAT helpers ------------------------------------------- */
void ora_matrix_inverse(const float in[16], float out[16]) {
    mat4 m;
    memcpy(&m.m[0][0], in, sizeof(float) * 16);
    mat4 r = ora_mat4_inverse(&m);
    memcpy(out, &r.m[0][0], sizeof(float) * 16);
}

int ora_bbox_ray(const float bmin[3], const float bmax[3], const float o[3],
                 const float d[3], float max_t, float* h0, float* h1) {
    bbox b; b.min = va(bmin); b.max = va(bmax);
    const v3 dinv = v3div(v3s(1.0f), va(d));
    return bbox_intersect(&b, va(o), dinv, max_t, h0, h1);
}

unsigned ora_bbox4_ray(const float bounds[24], const float o[3], const float d[3],
                       float max_t, float near_out[4]) {
    bbox4 b;
    memcpy(b.b, bounds, sizeof(float) * 24);
    const v3 dinv = v3div(v3s(1.0f), va(d));
    return bbox4_intersect(&b, va(o), dinv, max_t, near_out);
}

/* ---------------- tracer state ------------------------------------------ */
typedef struct {
    const ora_scene* scene;
    const ora_wray* w;          /* world rays of this batch */
    ora_inter* inter;           /* intersections of this batch */
    ora_ray* rays;
    /* transform frames (Stack xform_stack) */
    mat4* xf_pool; size_t xf_top, xf_cap;
    size_t frame_start[256]; size_t frame_count[256]; int n_frames;
    int shader_stack[256]; int n_shader;
    uint64_t element_id; int element_pos;
    /* patch frames */
    float* patch_pool; size_t patch_cap;
    int brute;   /* 1: skip the BVH, test every instance (oracle self-check) */
} tracer;

static void push_frame(tracer* tr, size_t count) {
    if (tr->xf_top + count > tr->xf_cap) {
        tr->xf_cap = (tr->xf_top + count) * 2 + 16;
        tr->xf_pool = (mat4*)realloc(tr->xf_pool, sizeof(mat4) * tr->xf_cap);
    }
    tr->frame_start[tr->n_frames] = tr->xf_top;
    tr->frame_count[tr->n_frames] = count;
    tr->n_frames++;
    tr->xf_top += count;
}
static void pop_frame(tracer* tr) {
    tr->n_frames--;
    tr->xf_top = tr->frame_start[tr->n_frames];
}

/* merge(Transform* dest, a, b) (math/transform.hpp:241-272) */
static void merge_xf(mat4* dest, const mat4* a, size_t na, const mat4* b, size_t nb) {
    if (na == nb) {
        for (size_t i = 0; i < na; ++i) dest[i] = mat4_mul(&a[i], &b[i]);
    } else if (na > nb) {
        if (nb == 0) { for (size_t i = 0; i < na; ++i) dest[i] = a[i]; }
        else {
            const float inv_s = 1.0f / (float)(na - 1);
            for (size_t i = 0; i < na; ++i) {
                mat4 lb = mat4_lerp_seq((float)i * inv_s, b, (int)nb);
                dest[i] = mat4_mul(&a[i], &lb);
            }
        }
    } else {
        if (na == 0) { for (size_t i = 0; i < nb; ++i) dest[i] = b[i]; }
        else {
            const float inv_s = 1.0f / (float)(nb - 1);
            for (size_t i = 0; i < nb; ++i) {
                mat4 la = mat4_lerp_seq((float)i * inv_s, a, (int)na);
                dest[i] = mat4_mul(&b[i], &la);
            }
        }
    }
}

static void id_push(tracer* tr, uint64_t sub, int bits) {
    tr->element_id <<= bits;
    tr->element_id |= sub & (uint64_t)((1 << bits) - 1);
    tr->element_pos += bits;
}
static void id_pop(tracer* tr, int bits) {
    tr->element_id >>= bits;
    tr->element_pos -= bits;
}

static void set_space(tracer* tr, ora_inter* in, float time) {
    const size_t top = tr->n_frames - 1;
    const size_t cnt = tr->frame_count[top];
    mat4 s = cnt > 0 ? mat4_lerp_seq(time, tr->xf_pool + tr->frame_start[top], (int)cnt) : mat4_identity();
    memcpy(in->space, &s.m[0][0], sizeof(float) * 16);
}

/* SurfaceShader::shade (shading/surface_shader.hpp:28-53) and the
 * default magenta emitter (tracer/tracer.cpp:248-250) */
static void shade(const tracer* tr, ora_inter* in, int shader) {
    if (shader >= 0) {
        const psy_shader_desc* sd = &tr->scene->shaders[shader];
        in->closure = sd->kind;
        memcpy(in->closure_col, sd->color, sizeof(float) * 3);
        in->closure_param[0] = sd->roughness;
        in->closure_param[1] = sd->tail_shape;
        in->closure_param[2] = sd->fresnel;
    } else {
        in->closure = PSY_SHADER_EMIT;
        in->closure_col[0] = 1.0f; in->closure_col[1] = 0.0f; in->closure_col[2] = 1.0f;
    }
    in->closure_prob = 1.0f;
}

/* ---------------- sphere ------------------------------------------------- */
static int sphere_hit_t(const ora_ray* ray, v3 cent, float radi, float* t_out) {
    const v3 o = v3sub(ray->o, cent);
    const v3 d = ray->d;
    const float a = v3len2(d);
    const float b = 2.0f * v3dot(d, o);
    const float c = v3len2(o) - radi * radi;
    float disc = b * b - 4.0f * a * c;
    if (disc < 0.0f) return 0;
    disc = sqrtf(disc);
    float q;
    if (b < 0.0f) q = -0.5f * (b - disc);
    else q = -0.5f * (b + disc);
    float t0 = q / a, t1;
    if (q != 0.0f) t1 = c / q; else t1 = ray->max_t;
    if (t0 > t1) { float tmp = t1; t1 = t0; t0 = tmp; }
    if (t0 >= ray->max_t || t1 < 0.0001f) return 0;
    if (t0 >= 0.0001f) *t_out = t0;
    else if (t1 < ray->max_t) *t_out = t1;
    else return 0;
    return 1;
}

static void sphere_center_radius(const ora_object* ob, float time, v3* cent, float* radi, int stride) {
    float na;
    int i = lerp_seq_index(time, ob->ts, &na);
    const float* d = ob->data;
    if (i == -1) { *cent = va(d); *radi = d[3]; return; }
    if (i == -2) { *cent = va(d + (ob->ts - 1) * stride); *radi = d[(ob->ts - 1) * stride + 3]; return; }
    *cent = lerpv(na, va(d + i * stride), va(d + (i + 1) * stride));
    *radi = lerpf(na, d[i * stride + 3], d[(i + 1) * stride + 3]);
}

/* Sphere::intersect_ray (object/sphere.cpp:67-189) */
static int sphere_intersect(const ora_object* ob, const ora_ray* ray, ora_inter* in) {
    v3 cent; float radi;
    sphere_center_radius(ob, ray->time, &cent, &radi, 4);
    float t;
    if (!sphere_hit_t(ray, cent, radi, &t)) return 0;
    if (!ray_is_occl(ray)) {
        in->t = t;
        const v3 p = v3add(ray->o, v3scale(ray->d, t));
        v3 n = v3norm(v3sub(p, cent));
        set3(in->p, p);
        set3(in->n, n);
        in->backfacing = v3dot(n, v3norm(ray->d)) > 0.0f;
        const v3 unit_p = n;
        const v3 pp = v3scale(unit_p, radi);
        const float lat_cos = unit_p.z;
        const float lat_sin = sqrtf((unit_p.x * unit_p.x) + (unit_p.y * unit_p.y));
        const float long_cos = unit_p.x / lat_sin;
        const float long_sin = unit_p.y / lat_sin;
        float latitude = acosf(lat_cos);
        float longitude = 0.0f;
        if (unit_p.x != 0.0f || unit_p.y != 0.0f) {
            longitude = acosf(long_cos);
            if (unit_p.y < 0.0f) longitude = (float)((2.0 * M_PI) - (double)longitude);
        }
        const float pi2 = (float)(M_PI * 2);
        const float fpi = (float)M_PI;
        in->u = longitude / pi2;
        in->v = (float)((double)latitude / M_PI);
        const v3 dpdu = v3scale(v3m(pp.y * -1.0f, pp.x, 0.0f), pi2);
        const v3 dpdv = v3scale(v3m(pp.z * long_cos, pp.z * long_sin, -radi * lat_sin), fpi);
        const v3 d2pduu = v3scale(v3m(pp.x, pp.y, 0.0f), (-pi2 * pi2));
        const v3 d2pduv = v3scale(v3scale(v3scale(v3m(-long_sin, long_cos, 0.0f), fpi), pp.z), pi2);
        const v3 d2pdvv = v3scale(v3m(pp.x, pp.y, pp.z), (float)(-M_PI * M_PI));
        const float E = v3dot(dpdu, dpdu), F = v3dot(dpdu, dpdv), G = v3dot(dpdv, dpdv);
        const float e = v3dot(n, d2pduu), f = v3dot(n, d2pduv), g = v3dot(n, d2pdvv);
        const float invEGF2 = 1.0f / ((E * G) - (F * F));
        const v3 dndu = v3add(v3scale(dpdu, ((f * F) - (e * G)) * invEGF2), v3scale(dpdv, ((e * F) - (f * E)) * invEGF2));
        const v3 dndv = v3add(v3scale(dpdu, ((g * F) - (f * G)) * invEGF2), v3scale(dpdv, ((f * F) - (g * E)) * invEGF2));
        set3(in->dpdu, dpdu); set3(in->dpdv, dpdv); set3(in->dndu, dndu); set3(in->dndv, dndv);
        set3(in->offset, v3scale(n, 0.000001f));
    }
    return 1;
}

/* ---------------- lights ------------------------------------------------- */
/* spherical_triangle_solid_angle (utils/monte_carlo.hpp:110-137) */
static float sph_tri_area(v3 A, v3 B, v3 C) {
    double cos_a = (double)v3dot(B, C), cos_b = (double)v3dot(C, A), cos_c = (double)v3dot(A, B);
    cos_a = cos_a < -1.0 ? -1.0 : (cos_a > 1.0 ? 1.0 : cos_a);
    cos_b = cos_b < -1.0 ? -1.0 : (cos_b > 1.0 ? 1.0 : cos_b);
    cos_c = cos_c < -1.0 ? -1.0 : (cos_c > 1.0 ? 1.0 : cos_c);
    const double sin_a = sqrt(1.0 - cos_a * cos_a), sin_b = sqrt(1.0 - cos_b * cos_b),
                 sin_c = sqrt(1.0 - cos_c * cos_c);
    if (cos_a == 1.0 || cos_b == 1.0 || cos_c == 1.0) return 0.0f;
    double cA = (cos_a - cos_b * cos_c) / (sin_b * sin_c);
    double cB = (cos_b - cos_c * cos_a) / (sin_c * sin_a);
    double cC = (cos_c - cos_a * cos_b) / (sin_a * sin_b);
    cA = cA < -1.0 ? -1.0 : (cA > 1.0 ? 1.0 : cA);
    cB = cB < -1.0 ? -1.0 : (cB > 1.0 ? 1.0 : cB);
    cC = cC < -1.0 ? -1.0 : (cC > 1.0 ? 1.0 : cC);
    return (float)(acos(cA) + acos(cB) + acos(cC) - M_PI);
}

static void rect_dims(const ora_object* ob, float time, float* w, float* h, float col[3]) {
    float na;
    int i = lerp_seq_index(time, ob->ts, &na);
    const float* d = ob->data;
    if (i == -1 || i == -2) {
        const float* s = d + (i == -1 ? 0 : (ob->ts - 1) * 5);
        *w = s[0]; *h = s[1];
        col[0] = s[2]; col[1] = s[3]; col[2] = s[4];
        return;
    }
    const float f = lerp_fac(na);
    const float* a = d + i * 5;
    const float* b = d + (i + 1) * 5;
    *w = (a[0] * f) + (b[0] * na);
    *h = (a[1] * f) + (b[1] * na);
    for (int k = 0; k < 3; ++k) col[k] = (a[2 + k] * f) + (b[2 + k] * na);
}

/* RectangleLight::sample_pdf (lights/rectangle_light.hpp:76-93) */
float ora_rect_light_pdf(float w, float h, v3 arr) {
    const v3 p1 = v3norm(v3sub(v3m(w * 0.5f, h * 0.5f, 0.0f), arr));
    const v3 p2 = v3norm(v3sub(v3m(w * -0.5f, h * 0.5f, 0.0f), arr));
    const v3 p3 = v3norm(v3sub(v3m(w * -0.5f, h * -0.5f, 0.0f), arr));
    const v3 p4 = v3norm(v3sub(v3m(w * 0.5f, h * -0.5f, 0.0f), arr));
    const float a1 = sph_tri_area(p2, p1, p3);
    const float a2 = sph_tri_area(p4, p1, p3);
    return 1.0f / (a1 + a2);
}

/* RectangleLight::intersect_ray (lights/rectangle_light.hpp:110-147);
 * `fix`: test y against the height (PSY_SCENE_RECT_LIGHT_FIX) */
static int rect_light_intersect(const ora_object* ob, const ora_ray* ray, int fix, ora_inter* in) {
    if (ray->d.z == 0.0f) return 0;
    float w, h, col[3];
    rect_dims(ob, ray->time, &w, &h, col);
    const float t = -ray->o.z * ray->d_inv.z;
    if (t <= 0.0f || t > ray->max_t) return 0;
    const float x = ray->o.x + (ray->d.x * t);
    const float y = ray->o.y + (ray->d.y * t);
    /* the y test uses the width, as the reference does (:127), unless fixed */
    if (x >= (w * -0.5f) && x <= (w * 0.5f) && y >= (h * -0.5f) && y <= ((fix ? h : w) * 0.5f)) {
        in->t = t;
        set3(in->p, v3m(x, y, 0.0f));
        set3(in->n, v3m(0.0f, 0.0f, 1.0f));
        in->backfacing = ray->d.z > 0.0f;
        in->light_pdf = ora_rect_light_pdf(w, h, ray->o);
        set3(in->offset, v3m(0.0f, 0.0f, 0.000001f));
        const double area = (double)(w * h);
        in->closure = PSY_SHADER_EMIT;
        for (int k = 0; k < 3; ++k) in->closure_col[k] = ((col[k] * 0.5f) / (float)area) * ob->escale;
        return 1;
    }
    return 0;
}

/* SphereLight::sample_pdf (lights/sphere_light.hpp:124-141) */
float ora_sphere_light_pdf(v3 pos, double radius, v3 arr) {
    const double d2 = (double)v3len2(v3sub(pos, arr));
    const double d = sqrt(d2);
    if (d > radius) {
        const double s2 = (radius * radius) / d2 < 1.0 ? (radius * radius) / d2 : 1.0;
        const double cos_theta_max = sqrt(1.0 - s2);
        return (float)(1.0 / (2.0 * M_PI * (1.0 - cos_theta_max)));
    }
    return (float)(1.0f / (4.0f * M_PI));
}

/* SphereLight::intersect_ray (lights/sphere_light.hpp:159-246) */
static int sphere_light_intersect(const ora_object* ob, const ora_ray* ray, ora_inter* in) {
    v3 cent; float radi;
    sphere_center_radius(ob, ray->time, &cent, &radi, 7);
    const double area = 4.0 * M_PI * radi * radi;
    float t;
    if (!sphere_hit_t(ray, cent, radi, &t)) return 0;
    if (!ray_is_occl(ray)) {
        in->t = t;
        const v3 p = v3add(ray->o, v3scale(ray->d, t));
        const v3 n = v3norm(v3sub(p, cent));
        set3(in->p, p);
        set3(in->n, n);
        in->backfacing = v3dot(n, v3norm(ray->d)) > 0.0f;
        in->light_pdf = ora_sphere_light_pdf(cent, (double)radi, ray->o);
        set3(in->offset, v3scale(n, 0.000001f));
        float col[3];
        float na = 0.0f;
        int i = lerp_seq_index(ray->time, ob->ts, &na);
        for (int k = 0; k < 3; ++k) {
            if (i == -1) col[k] = ob->data[4 + k];
            else if (i == -2) col[k] = ob->data[(ob->ts - 1) * 7 + 4 + k];
            else col[k] = (ob->data[i * 7 + 4 + k] * lerp_fac(na)) + (ob->data[(i + 1) * 7 + 4 + k] * na);
        }
        in->closure = PSY_SHADER_EMIT;
        for (int k = 0; k < 3; ++k) in->closure_col[k] = (col[k] / (float)area) * ob->escale;
    }
    return 1;
}

/* ---------------- patches ------------------------------------------------ */
static inline v3 pt(const float* p, int i) { return v3m(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void spt(float* p, int i, v3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

static void bilinear_split_u(const float* p, float* p1, float* p2) {
    spt(p2, 0, v3scale(v3add(pt(p, 0), pt(p, 1)), 0.5f));
    spt(p2, 1, pt(p, 1));
    spt(p2, 2, v3scale(v3add(pt(p, 2), pt(p, 3)), 0.5f));
    spt(p2, 3, pt(p, 3));
    const v3 m01 = v3scale(v3add(pt(p, 0), pt(p, 1)), 0.5f);
    const v3 m23 = v3scale(v3add(pt(p, 2), pt(p, 3)), 0.5f);
    const v3 a0 = pt(p, 0), a2 = pt(p, 2);
    spt(p1, 0, a0); spt(p1, 1, m01); spt(p1, 2, a2); spt(p1, 3, m23);
}
static void bilinear_split_v(const float* p, float* p1, float* p2) {
    const v3 m02 = v3scale(v3add(pt(p, 0), pt(p, 2)), 0.5f);
    const v3 m13 = v3scale(v3add(pt(p, 1), pt(p, 3)), 0.5f);
    const v3 a0 = pt(p, 0), a1 = pt(p, 1), a2 = pt(p, 2), a3 = pt(p, 3);
    spt(p2, 0, m02); spt(p2, 1, m13); spt(p2, 2, a2); spt(p2, 3, a3);
    spt(p1, 0, a0); spt(p1, 1, a1); spt(p1, 2, m02); spt(p1, 3, m13);
}
static void bicubic_split_u(const float* p, float* p1, float* p2) {
    float in[48];
    memcpy(in, p, sizeof in);
    for (int r = 0; r < 4; ++r) {
        const int rr = r * 4;
        const v3 tmp = v3scale(v3add(pt(in, rr + 1), pt(in, rr + 2)), 0.5f);
        spt(p2, rr + 3, pt(in, rr + 3));
        spt(p2, rr + 2, v3scale(v3add(pt(in, rr + 3), pt(in, rr + 2)), 0.5f));
        spt(p2, rr + 1, v3scale(v3add(tmp, pt(p2, rr + 2)), 0.5f));
        spt(p1, rr + 0, pt(in, rr + 0));
        spt(p1, rr + 1, v3scale(v3add(pt(in, rr + 0), pt(in, rr + 1)), 0.5f));
        spt(p1, rr + 2, v3scale(v3add(tmp, pt(p1, rr + 1)), 0.5f));
        spt(p1, rr + 3, v3scale(v3add(pt(p1, rr + 2), pt(p2, rr + 1)), 0.5f));
        spt(p2, rr + 0, pt(p1, rr + 3));
    }
}
static void bicubic_split_v(const float* p, float* p1, float* p2) {
    float in[48];
    memcpy(in, p, sizeof in);
    for (int c = 0; c < 4; ++c) {
        const v3 tmp = v3scale(v3add(pt(in, c + 4), pt(in, c + 8)), 0.5f);
        spt(p2, c + 12, pt(in, c + 12));
        spt(p2, c + 8, v3scale(v3add(pt(in, c + 12), pt(in, c + 8)), 0.5f));
        spt(p2, c + 4, v3scale(v3add(tmp, pt(p2, c + 8)), 0.5f));
        spt(p1, c + 0, pt(in, c + 0));
        spt(p1, c + 4, v3scale(v3add(pt(in, c + 0), pt(in, c + 4)), 0.5f));
        spt(p1, c + 8, v3scale(v3add(tmp, pt(p1, c + 4)), 0.5f));
        spt(p1, c + 12, v3scale(v3add(pt(p1, c + 8), pt(p2, c + 4)), 0.5f));
        spt(p2, c + 0, pt(p1, c + 12));
    }
}

static bbox patch_bound(const float* p, int npts) {
    bbox b;
    b.min = pt(p, 0); b.max = b.min;
    for (int i = 1; i < npts; ++i) { b.min = v3min(b.min, pt(p, i)); b.max = v3max(b.max, pt(p, i)); }
    return b;
}

/* Bilinear::differential_geometry (object/bilinear.hpp:104-128) */
static v3 bl_eval_p(float u, v3 p0, v3 p1) {
    const float b0 = 1.0f - u, b1 = u;
    return v3add(v3scale(p0, b0), v3scale(p1, b1));
}
static v3 bl_eval_pd(float u, v3 p0, v3 p1) {
    (void)u;
    return v3add(v3scale(p0, -1.0f), v3scale(p1, 1.0f));
}
static void bilinear_dg(const float* p, float u, float v, v3* n, v3* dpdu, v3* dpdv, v3* dndu, v3* dndv) {
    *dpdu = bl_eval_pd(u, bl_eval_p(v, pt(p, 0), pt(p, 2)), bl_eval_p(v, pt(p, 1), pt(p, 3)));
    *dpdv = bl_eval_pd(v, bl_eval_p(u, pt(p, 0), pt(p, 1)), bl_eval_p(u, pt(p, 2), pt(p, 3)));
    *n = v3norm(v3cross(*dpdv, *dpdu));
    const v3 d2pduu = v3s(0.0f);
    const v3 d2pduv = bl_eval_pd(v, bl_eval_pd(u, pt(p, 0), pt(p, 1)), bl_eval_pd(u, pt(p, 2), pt(p, 3)));
    const v3 d2pdvv = v3s(0.0f);
    const float E = v3dot(*dpdu, *dpdu), F = v3dot(*dpdu, *dpdv), G = v3dot(*dpdv, *dpdv);
    const float e = v3dot(*n, d2pduu), f = v3dot(*n, d2pduv), g = v3dot(*n, d2pdvv);
    const float invEGF2 = 1.0f / ((E * G) - (F * F));
    *dndu = v3add(v3scale(*dpdu, ((f * F) - (e * G)) * invEGF2), v3scale(*dpdv, ((e * F) - (f * E)) * invEGF2));
    *dndv = v3add(v3scale(*dpdu, ((g * F) - (f * G)) * invEGF2), v3scale(*dpdv, ((f * F) - (g * E)) * invEGF2));
}

/* Bicubic::differential_geometry (object/bicubic.hpp:134-177) */
static v3 bc_eval_p(float u, v3 p0, v3 p1, v3 p2, v3 p3) {
    const float iu = 1.0f - u;
    const float b0 = iu * iu * iu, b1 = 3.0f * u * iu * iu, b2 = 3.0f * u * u * iu, b3 = u * u * u;
    return v3add(v3add(v3add(v3scale(p0, b0), v3scale(p1, b1)), v3scale(p2, b2)), v3scale(p3, b3));
}
static v3 bc_eval_pd(float u, v3 p0, v3 p1, v3 p2, v3 p3) {
    const float iu = 1.0f - u;
    const float d0 = -3.0f * iu * iu;
    const float d1 = (3.0f * iu * iu) - (6.0f * iu * u);
    const float d2 = (6.0f * iu * u) - (3.0f * u * u);
    const float d3 = 3.0f * u * u;
    return v3add(v3add(v3add(v3scale(p0, d0), v3scale(p1, d1)), v3scale(p2, d2)), v3scale(p3, d3));
}
static v3 bc_eval_pdd(float u, v3 p0, v3 p1, v3 p2, v3 p3) {
    const float iu = 1.0f - u;
    const float dd0 = 6.0f * iu, dd1 = (6.0f * u) - (12.0f * iu), dd2 = (6.0f * iu) - (12.0f * u), dd3 = 6.0f * u;
    return v3add(v3add(v3add(v3scale(p0, dd0), v3scale(p1, dd1)), v3scale(p2, dd2)), v3scale(p3, dd3));
}
static void bicubic_dg(const float* p, float u, float v, v3* n, v3* dpdu, v3* dpdv, v3* dndu, v3* dndv) {
    v3 pu[4], pv[4], pdv[4];
    for (int i = 0; i < 4; ++i) {
        pu[i] = bc_eval_p(v, pt(p, i), pt(p, i + 4), pt(p, i + 8), pt(p, i + 12));
        pv[i] = bc_eval_p(u, pt(p, i * 4), pt(p, i * 4 + 1), pt(p, i * 4 + 2), pt(p, i * 4 + 3));
        pdv[i] = bc_eval_pd(u, pt(p, i * 4), pt(p, i * 4 + 1), pt(p, i * 4 + 2), pt(p, i * 4 + 3));
    }
    *dpdu = bc_eval_pd(u, pu[0], pu[1], pu[2], pu[3]);
    *dpdv = bc_eval_pd(v, pv[0], pv[1], pv[2], pv[3]);
    *n = v3norm(v3cross(*dpdv, *dpdu));
    const v3 d2pduu = bc_eval_pdd(u, pu[0], pu[1], pu[2], pu[3]);
    const v3 d2pduv = bc_eval_pd(v, pdv[0], pdv[1], pdv[2], pdv[3]);
    const v3 d2pdvv = bc_eval_pdd(v, pv[0], pv[1], pv[2], pv[3]);
    const float E = v3dot(*dpdu, *dpdu), F = v3dot(*dpdu, *dpdv), G = v3dot(*dpdv, *dpdv);
    const float e = v3dot(*n, d2pduu), f = v3dot(*n, d2pduv), g = v3dot(*n, d2pdvv);
    const float invEGF2 = 1.0f / ((E * G) - (F * F));
    *dndu = v3add(v3scale(*dpdu, ((f * F) - (e * G)) * invEGF2), v3scale(*dpdv, ((e * F) - (f * E)) * invEGF2));
    *dndv = v3add(v3scale(*dpdu, ((g * F) - (f * G)) * invEGF2), v3scale(*dpdv, ((f * F) - (g * E)) * invEGF2));
}

#define SPLIT_STACK_SIZE 64

/* mutable_partition over ray pointers (utils/utils.hpp:122-146); the
 * predicate runs exactly once per element. */
typedef int (*ray_pred)(void* ctx, ora_ray* r);
static ora_ray* mutable_partition(ora_ray* b, ora_ray* e, ray_pred pred, void* ctx) {
    for (;;) {
        for (;;) {
            if (b == e) return b;
            if (!pred(ctx, b)) break;
            ++b;
        }
        do {
            if (b == --e) return b;
        } while (!pred(ctx, e));
        ora_ray tmp = *b; *b = *e; *e = tmp;
        ++b;
    }
}

typedef struct {
    tracer* tr;
    const ora_object* ob;
    int npts, tsc, stack_i;
    bbox bboxes[8];
    float max_dim;
    const float* uv;   /* current uv rect: u0 u1 v0 v1 */
    int shader;
    const mat4* pxf; size_t npxf;
} patch_ctx;

/* The per-ray predicate of intersect_rays_with_patch (object/patch_utils.hpp:49-143) */
static int patch_pred(void* vctx, ora_ray* ray) {
    patch_ctx* c = (patch_ctx*)vctx;
    if (ray_is_done(ray)) return 1;
    size_t t_index = 0;
    float t_nalpha = 0.0f;
    float h0, h1;
    int hit;
    if (c->tsc == 1) {
        hit = bbox_intersect(&c->bboxes[0], ray->o, ray->d_inv, ray->max_t, &h0, &h1);
    } else {
        const float t_time = ray->time * (float)(c->tsc - 1);
        t_index = (size_t)t_time;
        t_nalpha = t_time - (float)t_index;
        bbox lb = bbox_lerp(t_nalpha, c->bboxes[t_index], c->bboxes[t_index + 1]);
        hit = bbox_intersect(&lb, ray->o, ray->d_inv, ray->max_t, &h0, &h1);
    }
    if (!hit) return 1;
    /* std::max(min_width * dice_rate, min_upoly) (object/patch_utils.hpp:77);
     * a NaN footprint, for which the reference splits until its boxes stop
     * being hit (which need not end), takes min_upoly here */
    const float mw = ora_ray_min_width(ray, h0, h1) * c->tr->scene->dice_rate;
    const float width = mw > c->tr->scene->min_upoly ? mw : c->tr->scene->min_upoly;
    if (c->max_dim <= width || c->stack_i == (SPLIT_STACK_SIZE - 1)) {
        const float tt = (h0 + h1) * 0.5f;
        if (tt > 0.0f && tt < ray->max_t) {
            ora_inter* in = &c->tr->inter[ray_id(ray)];
            in->hit = 1;
            in->id = c->tr->element_id;
            if (ray_is_occl(ray)) {
                ray->id_and_flags |= RAY_DONE;
            } else {
                const int nf = ora_kind_floats(c->ob->kind);
                float ipatch[48];
                if (c->tsc == 1) {
                    memcpy(ipatch, c->ob->data, sizeof(float) * nf);
                } else {
                    const float* pa = c->ob->data + t_index * nf;
                    const float* pb = c->ob->data + (t_index + 1) * nf;
                    for (int k = 0; k < nf / 3; ++k) spt(ipatch, k, lerpv(t_nalpha, pt(pa, k), pt(pb, k)));
                }
                ray->max_t = tt;
                const float u = (c->uv[0] + c->uv[1]) * 0.5f;
                const float v = (c->uv[2] + c->uv[3]) * 0.5f;
                const float offset = c->max_dim * 1.74f;
                in->t = tt;
                mat4 sp = c->npxf > 0 ? mat4_lerp_seq(ray->time, c->pxf, (int)c->npxf) : mat4_identity();
                memcpy(in->space, &sp.m[0][0], sizeof(float) * 16);
                set3(in->p, v3add(ray->o, v3scale(ray->d, tt)));
                in->u = u;
                in->v = v;
                v3 n, dpdu, dpdv, dndu, dndv;
                if (c->ob->kind == PSY_OBJ_BILINEAR) bilinear_dg(ipatch, u, v, &n, &dpdu, &dpdv, &dndu, &dndv);
                else bicubic_dg(ipatch, u, v, &n, &dpdu, &dpdv, &dndu, &dndv);
                set3(in->n, n); set3(in->dpdu, dpdu); set3(in->dpdv, dpdv); set3(in->dndu, dndu); set3(in->dndv, dndv);
                in->backfacing = v3dot(n, v3norm(ray->d)) > 0.0f;
                set3(in->offset, v3scale(n, offset));
                in->patch_dim = c->max_dim;
                shade(c->tr, in, c->shader);
            }
        }
        return 1;
    }
    return 0;
}

/* intersect_rays_with_patch<PATCH> (object/patch_utils.hpp:19-200) */
static void intersect_patch(tracer* tr, const ora_object* ob, ora_ray* rb, ora_ray* re) {
    const int npts = ob->kind == PSY_OBJ_BILINEAR ? 4 : 16;
    const int nf = npts * 3;
    const int tsc = ob->ts;
    patch_ctx c;
    c.tr = tr; c.ob = ob; c.npts = npts; c.tsc = tsc;
    c.shader = tr->shader_stack[tr->n_shader - 1];
    const size_t top = tr->n_frames - 1;
    c.pxf = tr->xf_pool + tr->frame_start[top];
    c.npxf = tr->frame_count[top];
    ora_ray* rs_first[SPLIT_STACK_SIZE];
    ora_ray* rs_second[SPLIT_STACK_SIZE];
    float uv_stack[SPLIT_STACK_SIZE][4];
    const size_t frame = (size_t)tsc * nf;
    if (tr->patch_cap < frame * (SPLIT_STACK_SIZE + 2)) {
        tr->patch_cap = frame * (SPLIT_STACK_SIZE + 2);
        tr->patch_pool = (float*)realloc(tr->patch_pool, sizeof(float) * tr->patch_cap);
    }
    float* pool = tr->patch_pool;
    int stack_i = 0;
    rs_first[0] = rb; rs_second[0] = re;
    memcpy(pool, ob->data, sizeof(float) * frame);
    uv_stack[0][0] = 0.0f; uv_stack[0][1] = 1.0f; uv_stack[0][2] = 0.0f; uv_stack[0][3] = 1.0f;
    int pframe = 0;  /* top patch frame index */
    while (stack_i >= 0) {
        float* cur = pool + (size_t)pframe * frame;
        c.bboxes[0] = patch_bound(cur, npts);
        float max_dim = longest_axis(v3sub(c.bboxes[0].max, c.bboxes[0].min));
        for (int i = 1; i < tsc; ++i) {
            c.bboxes[i] = patch_bound(cur + i * nf, npts);
            max_dim = smax(max_dim, longest_axis(v3sub(c.bboxes[i].max, c.bboxes[i].min)));
        }
        c.max_dim = max_dim;
        c.stack_i = stack_i;
        c.uv = uv_stack[stack_i];
        rs_first[stack_i] = mutable_partition(rs_first[stack_i], rs_second[stack_i], patch_pred, &c);
        if (rs_first[stack_i] != rs_second[stack_i]) {
            float uv[4];
            memcpy(uv, uv_stack[stack_i], sizeof uv);
            float* next = pool + (size_t)(pframe + 1) * frame;
            const float ulen = npts == 4 ? longest_axis(v3sub(pt(cur, 0), pt(cur, 1)))
                                         : longest_axis(v3sub(pt(cur, 0), pt(cur, 3)));
            const float vlen = npts == 4 ? longest_axis(v3sub(pt(cur, 0), pt(cur, 2)))
                                         : longest_axis(v3sub(pt(cur, 0), pt(cur, 12)));
            if (ulen > vlen) {
                for (int i = 0; i < tsc; ++i) {
                    float tmp1[48], tmp2[48];
                    if (npts == 4) bilinear_split_u(cur + i * nf, tmp1, tmp2);
                    else bicubic_split_u(cur + i * nf, tmp1, tmp2);
                    memcpy(cur + i * nf, tmp1, sizeof(float) * nf);
                    memcpy(next + i * nf, tmp2, sizeof(float) * nf);
                }
                const float mid = (float)((double)(uv[0] + uv[1]) * 0.5);
                uv_stack[stack_i][0] = uv[0]; uv_stack[stack_i][1] = mid;
                uv_stack[stack_i][2] = uv[2]; uv_stack[stack_i][3] = uv[3];
                uv_stack[stack_i + 1][0] = mid; uv_stack[stack_i + 1][1] = uv[1];
                uv_stack[stack_i + 1][2] = uv[2]; uv_stack[stack_i + 1][3] = uv[3];
            } else {
                for (int i = 0; i < tsc; ++i) {
                    float tmp1[48], tmp2[48];
                    if (npts == 4) bilinear_split_v(cur + i * nf, tmp1, tmp2);
                    else bicubic_split_v(cur + i * nf, tmp1, tmp2);
                    memcpy(cur + i * nf, tmp1, sizeof(float) * nf);
                    memcpy(next + i * nf, tmp2, sizeof(float) * nf);
                }
                const float mid = (float)((double)(uv[2] + uv[3]) * 0.5);
                uv_stack[stack_i][0] = uv[0]; uv_stack[stack_i][1] = uv[1];
                uv_stack[stack_i][2] = uv[2]; uv_stack[stack_i][3] = mid;
                uv_stack[stack_i + 1][0] = uv[0]; uv_stack[stack_i + 1][1] = uv[1];
                uv_stack[stack_i + 1][2] = mid; uv_stack[stack_i + 1][3] = uv[3];
            }
            rs_first[stack_i + 1] = rs_first[stack_i];
            rs_second[stack_i + 1] = rs_second[stack_i];
            ++stack_i;
            ++pframe;
        } else {
            --stack_i;
            --pframe;
        }
    }
}

/* ---------------- surfaces / lights dispatch ---------------------------- */
static void trace_surface_or_light(tracer* tr, const ora_object* ob, ora_ray* rb, ora_ray* re) {
    const int is_light = ob->kind == PSY_OBJ_SPHERE_LIGHT || ob->kind == PSY_OBJ_RECTANGLE_LIGHT;
    for (ora_ray* r = rb; r != re; ++r) {
        ora_inter* in = &tr->inter[ray_id(r)];
        int hit;
        if (ob->kind == PSY_OBJ_SPHERE) hit = sphere_intersect(ob, r, in);
        else if (ob->kind == PSY_OBJ_SPHERE_LIGHT) hit = sphere_light_intersect(ob, r, in);
        else hit = rect_light_intersect(ob, r, (tr->scene->flags & PSY_SCENE_RECT_LIGHT_FIX) != 0, in);
        if (!hit) continue;
        in->hit = 1;
        in->id = tr->element_id;
        if (ray_is_occl(r)) {
            r->id_and_flags |= RAY_DONE;
        } else {
            r->max_t = in->t;
            set_space(tr, in, r->time);
            in->patch_dim = 0.0f;
            if (!is_light) shade(tr, in, tr->shader_stack[tr->n_shader - 1]);
        }
    }
}

/* ---------------- BVH4 stream traversal --------------------------------- */
#define BVH4_STACK_SIZE 64
typedef struct {
    const ora_assembly* a;
    ora_ray* rays_end;
    int first_call;
    int sp;
    size_t node_stack[BVH4_STACK_SIZE];
    ora_ray* rs_first[BVH4_STACK_SIZE];
    ora_ray* rs_second[BVH4_STACK_SIZE];
    /* per-node predicate state */
    size_t cur_node;
    int num_children, rot_set, rot;
} bvh4_trav;

static int child_count(const ora_assembly* a, size_t ni) {
    if (a->nodes[ni].child[1] == 0) return 2;
    if (a->nodes[ni].child[2] == 0) return 3;
    return 4;
}
static size_t child_of(const ora_assembly* a, size_t ni, int n) {
    if (n == 0) return ni + a->nodes[ni].ts;
    return a->nodes[ni].child[n - 1];
}

static int leaf_pred(void* vt, ora_ray* r) {
    bvh4_trav* t = (bvh4_trav*)vt;
    if (ray_is_done(r)) return 0;
    if (t->first_call) return 1;
    const int b = (int)(r->trav_stack & 1);
    r->trav_stack >>= 1;
    return b;
}

static int inner_pred(void* vt, ora_ray* r) {
    bvh4_trav* t = (bvh4_trav*)vt;
    if (ray_is_done(r)) return 0;
    if (!t->first_call) {
        const int b = (int)(r->trav_stack & 1);
        r->trav_stack >>= 1;
        if (!b) return 0;
    }
    const ora_bvh4_node* nb = &t->a->nodes[t->cur_node];
    const int ts = (int)nb->ts;
    bbox4 b;
    float na;
    int i = lerp_seq_index(r->time, ts, &na);
    if (i == -1) b = nb[0].bounds;
    else if (i == -2) b = nb[ts - 1].bounds;
    else {
        const float f = lerp_fac(na);
        for (int k = 0; k < 6; ++k)
            for (int l = 0; l < 4; ++l) b.b[k][l] = (nb[i].bounds.b[k][l] * f) + (nb[i + 1].bounds.b[k][l] * na);
    }
    float near_hits[4];
    const unsigned hit_mask = bbox4_intersect(&b, r->o, r->d_inv, r->max_t, near_hits);
#ifdef ORA_DEBUG
    fprintf(stderr, "node %zu ts %d time %g o=(%g %g %g) mask %x\n", t->cur_node, ts, r->time, r->o.x, r->o.y, r->o.z, hit_mask);
    for (int l = 0; l < 4; ++l)
        fprintf(stderr, "   lane %d x[%g %g] y[%g %g] z[%g %g]\n", l, b.b[0][l], b.b[1][l], b.b[2][l], b.b[3][l], b.b[4][l], b.b[5][l]);
#endif
    if (hit_mask != 0) {
        const int n = t->num_children;
        if (!t->rot_set) {
            t->rot_set = 1;
            for (int k = 1; k < n; ++k)
                if (near_hits[k] < near_hits[t->rot]) t->rot = k;
        }
        const uint32_t value = (hit_mask >> t->rot) | (hit_mask << (n - t->rot));
        r->trav_stack <<= n;
        r->trav_stack |= value & (uint32_t)((1 << n) - 1);
    }
    return hit_mask != 0;
}

/* BVH4StreamTraverser::next_object (accel/bvh4.cpp:163-236) */
static int next_object(bvh4_trav* t, ora_ray** rb, ora_ray** re, size_t* idx) {
    const ora_assembly* a = t->a;
    while (t->sp >= 0) {
        const size_t ni = t->node_stack[t->sp];
        if (a->nodes[ni].child[0] == 0) {
            t->rs_second[t->sp] = mutable_partition(t->rs_first[t->sp], t->rs_second[t->sp], leaf_pred, t);
            if (t->rs_second[t->sp] != t->rs_first[t->sp]) {
                *rb = t->rs_first[t->sp];
                *re = t->rs_second[t->sp];
                *idx = a->nodes[ni].data_index;
                --t->sp;
                return 1;
            }
            --t->sp;
        } else {
            t->num_children = child_count(a, ni);
            t->cur_node = ni;
            t->rot_set = 0;
            t->rot = 0;
            t->rs_second[t->sp] = mutable_partition(t->rs_first[t->sp], t->rs_second[t->sp], inner_pred, t);
            if (t->first_call) t->first_call = 0;
            if (t->rs_second[t->sp] != t->rs_first[t->sp]) {
                const int n = t->num_children;
                for (int i = 0; i < n; ++i) {
                    t->rs_first[t->sp + i] = t->rs_first[t->sp];
                    t->rs_second[t->sp + i] = t->rs_second[t->sp];
                    t->node_stack[t->sp + i] = child_of(a, ni, n - 1 - ((i + n - t->rot) % n));
                }
                t->sp += n - 1;
            } else {
                --t->sp;
            }
        }
    }
    *rb = t->rays_end; *re = t->rays_end; *idx = 0;
    return 0;
}

/* the node test of SubdivisionSurface::intersect_rays: the node's bounds
 * (one time sample) against the ray */
static int sd_node_pred(void* vctx, ora_ray* ray) {
    const ora_sd_node* n = (const ora_sd_node*)vctx;
    float h0, h1;
    return bbox_intersect(&n->b, ray->o, ray->d_inv, ray->max_t, &h0, &h1);
}

/* SubdivisionSurface::intersect_rays (object/subdivision_surface.cpp:41-83):
 * the patch BVH walked with ray partitions, each leaf's patch diced as a
 * Bicubic under the instance's id and shader */
static void intersect_subdiv(tracer* tr, const ora_object* ob, ora_ray* rb, ora_ray* re) {
    if (ob->sd_root < 0) return;
    int node_stack[64];
    ora_ray* end_stack[64];
    int sp = 0;
    node_stack[0] = ob->sd_root;
    end_stack[0] = re;
    while (sp >= 0) {
        const ora_sd_node* n = &ob->sd_nodes[node_stack[sp]];
        if (n->patch < 0) {
            end_stack[sp] = mutable_partition(rb, end_stack[sp], sd_node_pred, (void*)n);
            if (rb != end_stack[sp] && sp + 1 < 64) {
                node_stack[sp + 1] = n->child[0];
                node_stack[sp] = n->child[1];
                end_stack[sp + 1] = end_stack[sp];
                ++sp;
            } else {
                --sp;
            }
        } else {
            intersect_patch(tr, &ob->patches[n->patch], rb, end_stack[sp]);
            --sp;
        }
    }
}

/* Tracer::trace_assembly (tracer/tracer.cpp:125-219) */
static void trace_assembly(tracer* tr, const ora_assembly* a, ora_ray* rays, ora_ray* end) {
    bvh4_trav t;
    t.a = a;
    t.rays_end = end;
    t.first_call = 1;
    t.sp = a->n_nodes == 0 ? -1 : 0;
    t.node_stack[0] = 0;
    t.rs_first[0] = rays;
    t.rs_second[0] = end;
    ora_ray *hb, *he;
    size_t idx;
    size_t brute_i = 0;
    if (tr->brute) {
        hb = rays; he = a->n_inst > 0 ? end : rays; idx = 0;
    } else {
        next_object(&t, &hb, &he, &idx);
    }
    while (hb != he) {
        const psy_instance_desc* inst = &a->inst[idx];
        const int bits = asm_id_bits(a);
        id_push(tr, idx, bits);
        const size_t ptop = tr->n_frames - 1;
        const size_t pstart = tr->frame_start[ptop];
        const size_t pcount = tr->frame_count[ptop];
        if (inst->transform_count > 0) {
            const size_t larger = (size_t)inst->transform_count > pcount ? (size_t)inst->transform_count : pcount;
            push_frame(tr, larger);
            const size_t fs = tr->frame_start[tr->n_frames - 1];
            merge_xf(tr->xf_pool + fs, tr->xf_pool + pstart, pcount, a->xf + inst->transform_index,
                     (size_t)inst->transform_count);
            for (ora_ray* r = hb; r != he; ++r) {
                mat4 x = mat4_lerp_seq(r->time, tr->xf_pool + fs, (int)larger);
                ora_update_ray_xf(&tr->w[ray_id(r)], r, &x);
            }
        }
        if (inst->shader_index >= 0) tr->shader_stack[tr->n_shader++] = inst->shader_index;
        if (inst->type == PSY_INSTANCE_OBJECT) {
            const ora_object* ob = &a->obj[inst->data_index];
            if (ob->kind == PSY_OBJ_BILINEAR || ob->kind == PSY_OBJ_BICUBIC) intersect_patch(tr, ob, hb, he);
            else if (ob->kind == PSY_OBJ_SUBDIVISION) intersect_subdiv(tr, ob, hb, he);
            else trace_surface_or_light(tr, ob, hb, he);
        } else {
            trace_assembly(tr, a->sub[inst->data_index], hb, he);
        }
        if (inst->shader_index >= 0) tr->n_shader--;
        if (inst->transform_count > 0) {
            if (pcount > 0) {
                for (ora_ray* r = hb; r != he; ++r) {
                    mat4 x = mat4_lerp_seq(r->time, tr->xf_pool + pstart, (int)pcount);
                    ora_update_ray_xf(&tr->w[ray_id(r)], r, &x);
                }
            } else {
                for (ora_ray* r = hb; r != he; ++r) ora_update_ray(&tr->w[ray_id(r)], r);
            }
            pop_frame(tr);
        }
        id_pop(tr, bits);
        if (tr->brute) {
            ++brute_i;
            if (brute_i < (size_t)a->n_inst) { hb = rays; he = end; idx = brute_i; }
            else { hb = he = end; }
        } else {
            next_object(&t, &hb, &he, &idx);
        }
    }
}

/* std::partition (libstdc++ bidirectional) by d[axis] > 0 */
static ora_ray* partition_dir(ora_ray* first, ora_ray* last, int axis) {
    for (;;) {
        for (;;) {
            if (first == last) return first;
            if (v3get(first->d, axis) > 0.0f) ++first; else break;
        }
        --last;
        for (;;) {
            if (first == last) return first;
            if (!(v3get(last->d, axis) > 0.0f)) --last; else break;
        }
        ora_ray t = *first; *first = *last; *last = t;
        ++first;
    }
}

static void init_inter(ora_inter* in) {
    memset(in, 0, sizeof *in);
    in->t = INFINITY;
    in->light_pdf = 9999.0f;
    mat4 id = mat4_identity();
    memcpy(in->space, &id.m[0][0], sizeof(float) * 16);
    in->closure_prob = 1.0f;
}

/* Tracer::trace (tracer/tracer.cpp:38-121) */
static void trace_batch(tracer* tr, const ora_wray* w, int64_t n, ora_inter* out) {
    tr->w = w;
    tr->inter = out;
    tr->rays = (ora_ray*)realloc(tr->rays, sizeof(ora_ray) * (size_t)(n > 0 ? n : 1));
    tr->element_id = 0; tr->element_pos = 0;
    for (int64_t i = 0; i < n; ++i) {
        to_ray(&w[i], &tr->rays[i]);
        tr->rays[i].id_and_flags = (tr->rays[i].id_and_flags & ~RAY_IDMASK) | ((uint32_t)i & RAY_IDMASK);
        init_inter(&out[i]);
    }
    tr->n_shader = 0;
    tr->shader_stack[tr->n_shader++] = -1;
    tr->n_frames = 0; tr->xf_top = 0;
    push_frame(tr, 0);
    ora_ray* b = tr->rays;
    ora_ray* e = tr->rays + n;
    ora_ray* x4 = partition_dir(b, e, 0);
    ora_ray* y2 = partition_dir(b, x4, 1);
    ora_ray* y6 = partition_dir(x4, e, 1);
    ora_ray* z1 = partition_dir(b, y2, 2);
    ora_ray* z3 = partition_dir(y2, x4, 2);
    ora_ray* z5 = partition_dir(x4, y6, 2);
    ora_ray* z7 = partition_dir(y6, e, 2);
    const ora_assembly* root = &tr->scene->asms[0];
    trace_assembly(tr, root, b, z1);
    trace_assembly(tr, root, z1, y2);
    trace_assembly(tr, root, y2, z3);
    trace_assembly(tr, root, z3, x4);
    trace_assembly(tr, root, x4, z5);
    trace_assembly(tr, root, z5, y6);
    trace_assembly(tr, root, y6, z7);
    trace_assembly(tr, root, z7, e);
}

typedef struct {
    const ora_scene* s;
    const ora_wray* rays;
    ora_inter* out;
    int brute;
    int64_t n, chunk;
    int64_t* next;
    pthread_mutex_t* mu;
} worker_arg;

static void* worker(void* p) {
    worker_arg* wa = (worker_arg*)p;
    tracer tr;
    memset(&tr, 0, sizeof tr);
    tr.scene = wa->s;
    tr.brute = wa->brute;
    for (;;) {
        pthread_mutex_lock(wa->mu);
        const int64_t start = *wa->next;
        *wa->next += wa->chunk;
        pthread_mutex_unlock(wa->mu);
        if (start >= wa->n) break;
        const int64_t cnt = (start + wa->chunk > wa->n) ? wa->n - start : wa->chunk;
        trace_batch(&tr, wa->rays + start, cnt, wa->out + start);
    }
    free(tr.rays); free(tr.xf_pool); free(tr.patch_pool);
    return NULL;
}

static int64_t trace_impl(const ora_scene* s, const ora_wray* rays, int64_t n, ora_inter* out,
                          int threads, int64_t chunk, int brute) {
    if (!s || n <= 0) return 0;
    if (chunk <= 0) chunk = 65536;
    if (threads < 1) threads = 1;
    int64_t next = 0;
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    worker_arg wa = {s, rays, out, brute, n, chunk, &next, &mu};
    if (threads == 1) {
        worker(&wa);
        return n;
    }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, worker, &wa);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
    return n;
}

/* ---------------- more KAT helpers --------------------------------------- */
/* Imath Vec3 length / normalize / dot / cross (math/vector_test.cpp:90-137) */
float ora_vec3_length(const float v[3]) { return v3len(va(v)); }
void ora_vec3_normalize(const float v[3], float out[3]) { set3(out, v3norm(va(v))); }
float ora_vec3_dot(const float a[3], const float b[3]) { return v3dot(va(a), va(b)); }
void ora_vec3_cross(const float a[3], const float b[3], float out[3]) { set3(out, v3cross(va(a), va(b))); }

/* mutable_partition on ints with predicate v == key (utils/utils_test.cpp:6-117);
 * returns the split index, *calls the number of predicate evaluations. */
static int g_part_key, g_part_calls;
static int int_pred(const int* v) { ++g_part_calls; return *v == g_part_key; }
int ora_partition_ints(int* v, int n, int key, int* calls) {
    int* b = v;
    int* e = v + n;
    g_part_key = key;
    g_part_calls = 0;
    for (;;) {
        for (;;) {
            if (b == e) goto done;
            if (!int_pred(b)) break;
            ++b;
        }
        do {
            if (b == --e) goto done;
        } while (!int_pred(e));
        { int t = *b; *b = *e; *e = t; }
        ++b;
    }
done:
    *calls = g_part_calls;
    return (int)(b - v);
}

int64_t ora_trace(const ora_scene* s, const ora_wray* rays, int64_t n, ora_inter* out,
                  int threads, int64_t chunk) {
    return trace_impl(s, rays, n, out, threads, chunk, 0);
}

/* Same tracer with the BVH replaced by visiting every instance of every
 * assembly: the closest hit it finds pins the BVH restatement. */
int64_t ora_trace_brute(const ora_scene* s, const ora_wray* rays, int64_t n, ora_inter* out,
                        int threads, int64_t chunk) {
    return trace_impl(s, rays, n, out, threads, chunk, 1);
}

/* ---- entry points of the oracle's own checks (tests/test_oracle_pins.py):
 * the functions above, one at a time, on plain arrays ------------------- */
int ora_kat_sphere_t(const float o[3], const float d[3], const float c[3], float r, float max_t, float* t) {
    ora_ray ray;
    memset(&ray, 0, sizeof ray);
    ray.o = va(o);
    ray.d = va(d);
    ray.max_t = max_t;
    ray.d_inv = v3div(v3s(1.0f), ray.d);
    return sphere_hit_t(&ray, va(c), r, t);
}

void ora_kat_closest_ray_t2(const float o1[3], const float d1[3], const float o2[3], const float d2[3],
                            float* t, float* dist) {
    closest_ray_t2(va(o1), va(d1), va(o2), va(d2), t, dist);
}

float ora_kat_min_width(const float w6[6], float tnear, float tfar) {
    ora_ray ray;
    memset(&ray, 0, sizeof ray);
    ray.owx = w6[0]; ray.owy = w6[1]; ray.dwx = w6[2]; ray.dwy = w6[3]; ray.fwx = w6[4]; ray.fwy = w6[5];
    return ora_ray_min_width(&ray, tnear, tfar);
}

/* kind PSY_OBJ_BILINEAR (4 points) or PSY_OBJ_BICUBIC (16); dir 0: split_u */
void ora_kat_patch_split(int kind, const float* p, int dir, float* p1, float* p2) {
    if (kind == PSY_OBJ_BILINEAR) {
        if (dir == 0) bilinear_split_u(p, p1, p2); else bilinear_split_v(p, p1, p2);
    } else {
        if (dir == 0) bicubic_split_u(p, p1, p2); else bicubic_split_v(p, p1, p2);
    }
}

void ora_kat_patch_dg(int kind, const float* p, float u, float v, float n[3], float dpdu[3], float dpdv[3],
                      float dndu[3], float dndv[3]) {
    v3 vn, vu, vv, nu, nv;
    if (kind == PSY_OBJ_BILINEAR) bilinear_dg(p, u, v, &vn, &vu, &vv, &nu, &nv);
    else bicubic_dg(p, u, v, &vn, &vu, &vv, &nu, &nv);
    set3(n, vn); set3(dpdu, vu); set3(dpdv, vv); set3(dndu, nu); set3(dndv, nv);
}

/* InstanceID push_back sequence as the tracer builds element ids */
uint64_t ora_kat_instance_id(const uint64_t* sub, const int* bits, int n) {
    tracer tr;
    memset(&tr, 0, sizeof tr);
    for (int i = 0; i < n; ++i) id_push(&tr, sub[i], bits[i]);
    return tr.element_id;
}

float ora_kat_sphere_light_pdf(const float pos[3], float radius, const float arr[3]) {
    return ora_sphere_light_pdf(va(pos), (double)radius, va(arr));
}

float ora_kat_rect_light_pdf(float w, float h, const float arr[3]) { return ora_rect_light_pdf(w, h, va(arr)); }
