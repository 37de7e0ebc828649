/*
 * ora_camera.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Ray-stream generation restated from the reference:
 *   init_path / next_ray_for_path step 0   integrator/path_trace_integrator.cpp:91-125
 *   bounce step                            integrator/path_trace_integrator.cpp:183-209
 *   Camera::generate_ray                   basics/camera.hpp:72-104
 *   fast_logit / faster_ln                 utils/monte_carlo.hpp:17-21, utils/utils.hpp:335-343
 *   square_to_circle, cosine_sample_hemisphere
 *                                          utils/monte_carlo.hpp:29-61
 *   zup_to_vec, coordinate_system_from_vec3 utils/utils.hpp:214-248
 *   LambertClosure::sample / propagate_differentials
 *                                          shading/surface_closure.hpp:240-326
 *   DifferentialGeometry::transformed_from basics/differential_geometry.hpp:20-40
 *   transfer_ray_origin_differential       basics/differential_geometry.hpp:83-98
 *   Film window [-1,1] x [-h/w, h/w]       renderer/renderer.cpp:52-55
 */
#include <stdlib.h>
#include <string.h>
#include "ora_internal.h"

#ifndef M_PI
#define M_PI 3.1415926535897932384626433
#endif
#define QPI (M_PI / 4)
#define INV_PI (1.0 / M_PI)

static inline v3 va(const float* f) { return v3m(f[0], f[1], f[2]); }
static inline void set3(float* d, v3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

static float faster_ln(float x) {
    uint32_t i;
    memcpy(&i, &x, 4);
    float y = (float)i;
    y *= 8.2629582881927490e-8f;
    return y - 87.989971088f;
}

static float fast_logit(float p, float width) {
    const float n = 0.001f + (p * 0.998f);
    return faster_ln((n / (1.0f - n))) * width * (0.6266f / 4.0f);
}

static void square_to_circle(float* x, float* y) {
    if (*x == 0.0f && *y == 0.0f) return;
    float radius, angle;
    if (*x > fabsf(*y)) {
        radius = *x;
        angle = (float)(QPI * (double)(*y / *x));
    } else if (*y > fabsf(*x)) {
        radius = *y;
        angle = (float)(QPI * (double)(2 - (*x / *y)));
    } else if (*x < -fabsf(*y)) {
        radius = -*x;
        angle = (float)(QPI * (double)(4 + (*y / *x)));
    } else {
        radius = -*y;
        angle = (float)(QPI * (double)(6 - (*x / *y)));
    }
    *x = radius * cosf(angle);
    *y = radius * sinf(angle);
}

static void coordinate_system_from_vec3(v3 v1, v3* v2, v3* v3o) {
    if (fabsf(v1.x) > fabsf(v1.y)) {
        const float invlen = 1.f / sqrtf(v1.x * v1.x + v1.z * v1.z);
        *v2 = v3m(-v1.z * invlen, 0.f, v1.x * invlen);
    } else {
        const float invlen = 1.f / sqrtf(v1.y * v1.y + v1.z * v1.z);
        *v2 = v3m(0.f, v1.z * invlen, -v1.y * invlen);
    }
    *v3o = v3cross(v1, *v2);
}

static v3 zup_to_vec(v3 from, v3 toz) {
    toz = v3norm(toz);
    v3 tox, toy;
    coordinate_system_from_vec3(toz, &tox, &toy);
    return v3add(v3add(v3scale(tox, from.x), v3scale(toy, from.y)), v3scale(toz, from.z));
}

/* Camera::generate_ray */
void ora_camera_ray_from_dims(const ora_scene* s, float x, float y, float dx, float dy, float time,
                              float u, float v, ora_wray* w);
static void generate_ray(const ora_scene* s, float x, float y, float dx, float dy, float time,
                         float u, float v, ora_wray* w) {
    memset(w, 0, sizeof *w);
    w->type = 1;  /* WorldRay::CAMERA */
    w->time = time;
    const mat4 xf = mat4_lerp_seq(time, s->cam_xf, s->n_cam_xf);
    const float tfov = lerp_seq_f(time, s->tfovs, s->n_tfov, 1);
    const float ap = lerp_seq_f(time, s->aperture, s->n_ap, 1);
    const float fd = lerp_seq_f(time, s->focus, s->n_fd, 1);
    float ox = ap * ((u * 2) - 1);
    float oy = ap * ((v * 2) - 1);
    square_to_circle(&ox, &oy);
    v3 o = v3m(ox, oy, 0.0f);
    v3 d = v3m((x * tfov) - (o.x / fd), (y * tfov) - (o.y / fd), 1.0f);
    d = v3norm(d);
    const v3 ddx = v3m(dx * tfov, 0.0f, 0.0f);
    const v3 ddy = v3m(0.0f, dy * tfov, 0.0f);
    set3(w->o, mat4_pos(&xf, o));
    set3(w->d, mat4_dir(&xf, d));
    set3(w->odx, mat4_dir(&xf, v3s(0.0f)));
    set3(w->ody, mat4_dir(&xf, v3s(0.0f)));
    set3(w->ddx, mat4_dir(&xf, ddx));
    set3(w->ddy, mat4_dir(&xf, ddy));
}

void ora_camera_ray_from_dims(const ora_scene* s, float x, float y, float dx, float dy, float time,
                              float u, float v, ora_wray* w) {
    generate_ray(s, x, y, dx, dy, time, u, v, w);
}

void ora_camera_rays(const ora_scene* s, int width, int height, int spp, uint32_t seed,
                     int64_t first, int64_t count, ora_wray* out, float* wavelength) {
    const float min_x = -1.0f, max_x = 1.0f;
    const float max_y = ((float)height) / (float)width;
    const float min_y = -max_y;
    for (int64_t k = 0; k < count; ++k) {
        const int64_t i = first + k;
        const int sidx = (int)(i % spp);
        const int64_t pix = i / spp;
        const int px = (int)(pix % width);
        const int py = (int)(pix / width);
        const uint32_t off = ora_sampler_offset((uint32_t)px, (uint32_t)py, (uint32_t)sidx, seed);
        uint32_t dim = 0;
        const float wl = (ora_sampler_get(off, dim++) * (700.0f - 380.0f)) + 380.0f;
        const float samp_x = (fast_logit(ora_sampler_get(off, dim++), 1.5f) + 0.5f + (float)(short)px) / (float)width;
        const float samp_y = (fast_logit(ora_sampler_get(off, dim++), 1.5f) + 0.5f + (float)(short)py) / (float)height;
        const float rx = (float)(((double)samp_x - 0.5) * (double)(max_x - min_x));
        const float ry = (float)((0.5 - (double)samp_y) * (double)(max_y - min_y));
        const float dx = (max_x - min_x) / (float)width;
        const float dy = (max_y - min_y) / (float)height;
        const float sa = ora_sampler_get(off, dim++);
        const float sb = ora_sampler_get(off, dim++);
        const float sc = ora_sampler_get(off, dim++);
        generate_ray(s, rx, ry, dx, dy, sc, sa, sb, &out[k]);
        if (wavelength) wavelength[k] = wl;
    }
}

/* DifferentialGeometry::transformed_from(space) -> world p and n */
static void dg_to_world(const ora_inter* h, v3* p, v3* n) {
    mat4 sp;
    memcpy(&sp.m[0][0], h->space, sizeof(float) * 16);
    const mat4 inv = ora_mat4_inverse(&sp);
    *p = mat4_pos(&inv, va(h->p));
    v3 nn = mat4_dir_transpose(&sp, va(h->n));
    const float il = 1.0f / v3len(nn);
    *n = v3scale(nn, il);
}

static v3 transfer_origin_diff(float t, v3 normal, v3 d, v3 od, v3 dd) {
    const v3 temp = v3add(od, v3scale(dd, t));
    const float td = -v3dot(temp, normal) / v3dot(d, normal);
    const v3 real_projected = v3add(temp, v3scale(d, td));
    return v3scale(v3norm(real_projected), v3len(temp));
}

/* LambertClosure::sample + the bounce-ray set-up of next_ray_for_path for
 * one stored hit: writes o, d and differentials of `w`, the pdf, and the
 * world-space hit point / normal. */
void ora_bounce_dir(const ora_inter* h, const ora_wray* prev, float su, float sv, ora_wray* w, float* pdf,
                    v3* world_p, v3* world_n) {
    v3 gp, gn;
    dg_to_world(h, &gp, &gn);
    mat4 sp;
    memcpy(&sp.m[0][0], h->space, sizeof(float) * 16);
    const mat4 inv = ora_mat4_inverse(&sp);
    v3 pos_offset = mat4_dir(&inv, va(h->offset));
    if (h->backfacing) pos_offset = v3scale(pos_offset, -1.0f);
    const v3 pd = va(prev->d);
    v3 nn = v3norm(gn);
    if (v3dot(nn, pd) > 0.0f) nn = v3scale(nn, -1.0f);
    su = (su * 2) - 1;
    sv = (sv * 2) - 1;
    square_to_circle(&su, &sv);
    const double zz = 1.0 - (double)((su * su) + (sv * sv));
    const float z = (float)sqrt(zz > 0.0 ? zz : 0.0);
    const v3 dir = v3m(su, sv, z);
    *pdf = dir.z * (float)INV_PI;
    const v3 outd = zup_to_vec(dir, nn);
    set3(w->o, v3add(gp, pos_offset));
    set3(w->d, outd);
    /* LambertClosure::propagate_differentials */
    const float len = v3len(outd);
    const v3 gnn = v3norm(gn);
    v3 x, y;
    coordinate_system_from_vec3(outd, &x, &y);
    x = v3norm(x);
    y = v3norm(y);
    set3(w->odx, transfer_origin_diff(h->t, gnn, pd, va(prev->odx), va(prev->ddx)));
    set3(w->ody, transfer_origin_diff(h->t, gnn, pd, va(prev->ody), va(prev->ddy)));
    set3(w->ddx, v3sdiv(v3scale(x, 0.15f), len));
    set3(w->ddy, v3sdiv(v3scale(y, 0.15f), len));
    *world_p = gp;
    *world_n = gn;
}

void ora_bounce_rays(const ora_scene* s, const ora_wray* in, const ora_inter* hits,
                     int64_t first, int64_t count, int width, int spp, uint32_t seed,
                     uint32_t dim_u, uint32_t dim_v, ora_wray* out, uint32_t* valid) {
    (void)s;
    for (int64_t k = 0; k < count; ++k) {
        const int64_t i = first + k;
        const ora_inter* h = &hits[k];
        ora_wray* w = &out[k];
        memset(w, 0, sizeof *w);
        if (!h->hit || h->closure != PSY_SHADER_LAMBERT) { valid[k] = 0; continue; }
        valid[k] = 1;
        const int sidx = (int)(i % spp);
        const int64_t pix = i / spp;
        const uint32_t off = ora_sampler_offset((uint32_t)(pix % width), (uint32_t)(pix / width), (uint32_t)sidx, seed);
        float pdf;
        v3 gp, gn;
        ora_bounce_dir(h, &in[k], ora_sampler_get(off, dim_u), ora_sampler_get(off, dim_v), w, &pdf, &gp, &gn);
        w->time = in[k].time;
        w->type = 2;  /* WorldRay::R_DIFFUSE */
    }
}

/* ---- entry points of the oracle's own checks (tests/test_oracle_pins.py) */
float ora_kat_faster_ln(float x) { return faster_ln(x); }
float ora_kat_fast_logit(float p, float width) { return fast_logit(p, width); }
void ora_kat_square_to_circle(float* x, float* y) { square_to_circle(x, y); }
