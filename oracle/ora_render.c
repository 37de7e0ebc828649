/*
 * ora_render.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * The wavefront path tracer restated from the reference:
 *   PathTraceIntegrator::init_path / next_ray_for_path / update_path /
 *   render_blocks           integrator/path_trace_integrator.cpp:91-384
 *   LightTree::sample, node_prob   accel/light_tree.cpp:6-96
 *   SphereLight::sample            lights/sphere_light.hpp:58-122
 *   RectangleLight::sample         lights/rectangle_light.hpp:40-74
 *   uniform_sample_cone(_pdf), uniform_sample_sphere,
 *   uniform_sample_spherical_triangle, sphere_lambert
 *                                  utils/monte_carlo.hpp:82-233
 *   LambertClosure evaluate / sample_pdf / estimate_eval_over_solid_angle
 *                                  shading/surface_closure.hpp:233-331
 *   power_heuristic                utils/mis.hpp:11-16
 *   SpectralSample, wavelength_n, Color_XYZ, XYZ_NORM_FAC
 *                                  color/color.hpp:15-188
 *   Film::add_sample (sum + count) film/film.hpp:135-153
 *
 * Colour -> spectrum (color/color.hpp:490-511): through the Hanika grid
 * (color/spectrum_grid.h:21-135, ora_spectrum_xyz_to_p) when the caller has
 * set one (ora_spectrum_set_grid: the reference's own tables), else for
 * achromatic colours only, where the grid's spectrum is flat to within
 * 2e-6 relative (pinned against the reference by
 * tests/golden/ref_vectors.npz): (X+Y+Z)/3 at every wavelength.  Without a
 * grid, chromatic colours are reported through ora_render's return value
 * and not rendered.
 */
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include "ora_internal.h"

#ifndef M_PI
#define M_PI 3.1415926535897932384626433
#endif
#define HPI (M_PI / 2)
#define INV_PI (1.0 / M_PI)

#define WL_MIN 380.0f
#define WL_MAX 700.0f
#define NSPEC 4
static const float XYZ_NORM_FAC = 0.009358239977091027f * (700.0f - 380.0f);

static inline v3 va(const float* f) { return v3m(f[0], f[1], f[2]); }
static inline void set3(float* d, v3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

/* ---- spectral ------------------------------------------------------------ */
static inline float wavelength_n(float hero, int n) {
    hero += (float)n * ((WL_MAX - WL_MIN) / NSPEC);
    if (hero > WL_MAX) hero -= (WL_MAX - WL_MIN);
    return hero;
}

static int g_chromatic_seen = 0;

/* Color_to_SpectralSample for one colour: Color_to_XYZ, then the flat
 * white-point spectrum scaled by X+Y+Z (spectrum_xyz_to_p / norm). */
float ora_color_to_spectrum(const float c[3]) {
    const float x = ((c[0] * 0.4339f) + (c[1] * 0.3762f)) + (c[2] * 0.1899f);
    const float y = ((c[0] * 0.2126f) + (c[1] * 0.7152f)) + (c[2] * 0.0722f);
    const float z = ((c[0] * 0.0177f) + (c[1] * 0.1095f)) + (c[2] * 0.8728f);
    if (!(c[0] == c[1] && c[1] == c[2])) g_chromatic_seen = 1;
    const float norm = (float)(1.0 / (double)(x + y + z));
    if (!(norm < FLT_MAX)) return 0.0f;
    return (1.0f / 3.0f) / norm;
}

/* ---- the spectrum grid (color/spectrum_grid.h) ----------------------------
 * The grid of Hanika et al.'s tristimulus upsampling, as data set by the
 * caller (the reference's spectrum_grid / spectrum_data_points tables):
 * cells are 8 ints (inside, num_points, idx[6]), data points are
 * 4 + n_samples floats (xystar[2], uv[2], spectrum[n_samples]). */
static struct {
    int set;
    int32_t* cells;
    float* points;
    int w, h, n_points, n_samples;
    float smin, smax;
    float m[6];
} g_grid;

int ora_spectrum_set_grid(const int32_t* cells, int w, int h, const float* points, int n_points, int n_samples,
                          float smin, float smax, const float* mat_xy_to_uv) {
    free(g_grid.cells);
    free(g_grid.points);
    memset(&g_grid, 0, sizeof g_grid);
    if (!cells) return 0;  /* cleared */
    if (w < 1 || h < 1 || n_points < 1 || n_samples < 2 || !(smin < smax) || !points || !mat_xy_to_uv) return -1;
    g_grid.cells = (int32_t*)malloc(sizeof(int32_t) * 8 * (size_t)(w * h));
    g_grid.points = (float*)malloc(sizeof(float) * (size_t)(4 + n_samples) * (size_t)n_points);
    memcpy(g_grid.cells, cells, sizeof(int32_t) * 8 * (size_t)(w * h));
    memcpy(g_grid.points, points, sizeof(float) * (size_t)(4 + n_samples) * (size_t)n_points);
    g_grid.w = w; g_grid.h = h; g_grid.n_points = n_points; g_grid.n_samples = n_samples;
    g_grid.smin = smin; g_grid.smax = smax;
    memcpy(g_grid.m, mat_xy_to_uv, sizeof g_grid.m);
    g_grid.set = 1;
    return 0;
}

static const float* gpoint(int i) { return g_grid.points + (size_t)(4 + g_grid.n_samples) * (size_t)i; }

/* spectrum_xyz_to_p (color/spectrum_grid.h:21-135): the grid cell of the
 * colour's chromaticity, the cell's data point spectra at lambda (linear
 * between the 5 nm knots), then bilinear weights in an inside cell or the
 * barycentric weights of the fan triangle around idx[0] that holds the
 * point; scaled by X+Y+Z.  Outside the grid, or a NaN chromaticity: 0. */
float ora_spectrum_xyz_to_p(float lambda, const float* xyz) {
    if (!g_grid.set) return 0.0f;
    const float norm = (float)(1.0 / (double)(xyz[0] + xyz[1] + xyz[2]));
    if (!(norm < FLT_MAX)) return 0.0f;
    const float cx = xyz[0] * norm, cy = xyz[1] * norm;
    float uv[2];
    uv[0] = ((g_grid.m[0] * cx) + (g_grid.m[1] * cy)) + g_grid.m[2];
    uv[1] = ((g_grid.m[3] * cx) + (g_grid.m[4] * cy)) + g_grid.m[5];
    if (!(uv[0] >= 0.0f && uv[0] < (float)g_grid.w && uv[1] >= 0.0f && uv[1] < (float)g_grid.h)) return 0.0f;
    const int ui = (int)uv[0], vi = (int)uv[1];
    const int32_t* cell = g_grid.cells + 8 * (ui + g_grid.w * vi);
    const int inside = cell[0], num = cell[1];
    const int32_t* idx = cell + 2;
    const float sb = ((lambda - g_grid.smin) / (g_grid.smax - g_grid.smin)) * (float)(g_grid.n_samples - 1);
    const int sb0 = (int)sb;
    const int sb1 = sb + 1.0f < (float)g_grid.n_samples ? (int)(sb + 1.0f) : g_grid.n_samples - 1;
    const float sbf = sb - (float)sb0;
    float p[6];
    for (int i = 0; i < num; ++i) {
        const float* sp = gpoint(idx[i]) + 4;
        p[i] = (sp[sb0] * (1.0f - sbf)) + (sp[sb1] * sbf);
    }
    float ip = 0.0f;
    if (inside) {
        const float u = uv[0] - (float)ui, v = uv[1] - (float)vi;
        ip = ((((p[0] * (1.0f - u)) * (1.0f - v)) + ((p[2] * (1.0f - u)) * v)) + ((p[3] * u) * v)) +
             ((p[1] * u) * (1.0f - v));
    } else if (num > 0) {
        const float* q0 = gpoint(idx[0]);
        const float ex = uv[0] - q0[2], ey = uv[1] - q0[3];
        float e0x = gpoint(idx[1])[2] - q0[2], e0y = gpoint(idx[1])[3] - q0[3];
        float uu = (e0x * ey) - (ex * e0y);
        for (int i = 0; i < num - 1; ++i) {
            const float* qn = gpoint(i == num - 2 ? idx[1] : idx[i + 2]);
            const float e1x = qn[2] - q0[2], e1y = qn[3] - q0[3];
            const float vv = (ex * e1y) - (e1x * ey);
            const float area = (e0x * e1y) - (e1x * e0y);
            const float u = uu / area, v = vv / area;
            const float w = (1.0f - u) - v;
            if (u < 0.0f || v < 0.0f || w < 0.0f) {
                uu = -vv;
                e0x = e1x;
                e0y = e1y;
                continue;
            }
            ip = ((p[0] * w) + (p[i + 1] * v)) + (p[i == num - 2 ? 1 : i + 2] * u);
            break;
        }
    }
    return ip / norm;
}

/* Color_to_SpectralSample (color/color.hpp:490-511) at the 4 wavelengths of
 * hero: through the grid when one is set, else the flat white-point
 * spectrum for grey colours (chromatic ones are reported). */
static void spec_from_color(const float c[3], float hero, float e[NSPEC]) {
    if (!g_grid.set) {
        const float v = ora_color_to_spectrum(c);
        for (int i = 0; i < NSPEC; ++i) e[i] = v;
        return;
    }
    float xyz[3];
    xyz[0] = ((c[0] * 0.4339f) + (c[1] * 0.3762f)) + (c[2] * 0.1899f);
    xyz[1] = ((c[0] * 0.2126f) + (c[1] * 0.7152f)) + (c[2] * 0.0722f);
    xyz[2] = ((c[0] * 0.0177f) + (c[1] * 0.1095f)) + (c[2] * 0.8728f);
    for (int i = 0; i < NSPEC; ++i)
        e[i] = ora_spectrum_xyz_to_p(wavelength_n(hero, i), xyz) * (1.0f / 0.009358239977091027f);
}

void ora_color_to_spectral_sample(const float c[3], float hero, float e[4]) { spec_from_color(c, hero, e); }

/* Wyman et al. CIE 1931 fits (color/color.hpp:150-165) */
static float X_1931(float w) {
    const float t1 = (w - 442.0f) * ((w < 442.0f) ? 0.0624f : 0.0374f);
    const float t2 = (w - 599.8f) * ((w < 599.8f) ? 0.0264f : 0.0323f);
    const float t3 = (w - 501.1f) * ((w < 501.1f) ? 0.0490f : 0.0382f);
    return (0.362f * expf(-0.5f * t1 * t1)) + (1.056f * expf(-0.5f * t2 * t2)) - (0.065f * expf(-0.5f * t3 * t3));
}
static float Y_1931(float w) {
    const float t1 = (w - 568.8f) * ((w < 568.8f) ? 0.0213f : 0.0247f);
    const float t2 = (w - 530.9f) * ((w < 530.9f) ? 0.0613f : 0.0322f);
    return (0.821f * expf(-0.5f * t1 * t1)) + (0.286f * expf(-0.5f * t2 * t2));
}
static float Z_1931(float w) {
    const float t1 = (w - 437.0f) * ((w < 437.0f) ? 0.0845f : 0.0278f);
    const float t2 = (w - 459.0f) * ((w < 459.0f) ? 0.0385f : 0.0725f);
    return (1.217f * expf(-0.5f * t1 * t1)) + (0.681f * expf(-0.5f * t2 * t2));
}

/* ---- sampling helpers (utils/monte_carlo.hpp) ---------------------------- */
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

/* zup_to_vec (utils/utils.hpp:239-248) */
static v3 zup_to_vec(v3 from, v3 toz) {
    toz = v3norm(toz);
    v3 tox, toy;
    coordinate_system_from_vec3(toz, &tox, &toy);
    return v3add(v3add(v3scale(tox, from.x), v3scale(toy, from.y)), v3scale(toz, from.z));
}

static v3 uniform_sample_sphere(float u, float v) {
    const float z = 1.f - (2.f * u);
    const float r = sqrtf(smax(0.f, 1.f - z * z));
    const float phi = (float)(2.f * M_PI * (double)v);
    return v3m(r * cosf(phi), r * sinf(phi), z);
}

static double dclamp(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }

static v3 uniform_sample_spherical_triangle(v3 A, v3 B, v3 C, float i, float j) {
    const double cos_a = dclamp((double)v3dot(B, C), -1.0, 1.0);
    const double cos_b = dclamp((double)v3dot(C, A), -1.0, 1.0);
    const double cos_c = dclamp((double)v3dot(A, B), -1.0, 1.0);
    const double sin_a = sqrt(1.0 - (cos_a * cos_a));
    const double sin_b = sqrt(1.0 - (cos_b * cos_b));
    const double sin_c = sqrt(1.0 - (cos_c * cos_c));
    if (cos_a == 1.0 || cos_b == 1.0 || cos_c == 1.0) return A;
    const double cos_A = dclamp((cos_a - (cos_b * cos_c)) / (sin_b * sin_c), -1.0, 1.0);
    const double cos_B = dclamp((cos_b - (cos_c * cos_a)) / (sin_c * sin_a), -1.0, 1.0);
    const double cos_C = dclamp((cos_c - (cos_a * cos_b)) / (sin_a * sin_b), -1.0, 1.0);
    const double sin_A = sqrt(1.0 - (cos_A * cos_A));
    const double ang_A = acos(cos_A), ang_B = acos(cos_B), ang_C = acos(cos_C);
    const double area = ang_A + ang_B + ang_C - M_PI;
    const double area_2 = area * i;
    const double s = sin(area_2 - ang_A);
    const double t = cos(area_2 - ang_A);
    const double u = t - cos_A;
    const double v = s + (sin_A * cos_c);
    const double q_top = (((v * t) - (u * s)) * cos_A) - v;
    const double q_bottom = ((v * s) + (u * t)) * sin_A;
    const double q = q_top / q_bottom;
    /* Vec3 * double -> the double is converted to float (Imath Vec3<float>) */
    const v3 C_2 = v3add(v3scale(A, (float)q),
                         v3scale(v3norm(v3sub(C, v3scale(A, v3dot(C, A)))), (float)sqrt(1.0 - (q * q))));
    const double z = 1.0 - (j * (1.0 - (double)v3dot(C_2, B)));
    return v3add(v3scale(B, (float)z), v3scale(v3norm(v3sub(C_2, v3scale(B, v3dot(C_2, B)))), (float)sqrt(1.0 - (z * z))));
}

float ora_rect_light_pdf(float w, float h, v3 arr);
float ora_sphere_light_pdf(v3 pos, double radius, v3 arr);

static float sph_tri_area2(v3 A, v3 B, v3 C) {
    const double cos_a = dclamp((double)v3dot(B, C), -1.0, 1.0);
    const double cos_b = dclamp((double)v3dot(C, A), -1.0, 1.0);
    const double cos_c = dclamp((double)v3dot(A, B), -1.0, 1.0);
    const double sin_a = sqrt(1.0 - (cos_a * cos_a));
    const double sin_b = sqrt(1.0 - (cos_b * cos_b));
    const double sin_c = sqrt(1.0 - (cos_c * cos_c));
    if (cos_a == 1.0 || cos_b == 1.0 || cos_c == 1.0) return 0.0f;
    const double cos_A = dclamp((cos_a - (cos_b * cos_c)) / (sin_b * sin_c), -1.0, 1.0);
    const double cos_B = dclamp((cos_b - (cos_c * cos_a)) / (sin_c * sin_a), -1.0, 1.0);
    const double cos_C = dclamp((cos_c - (cos_a * cos_b)) / (sin_a * sin_b), -1.0, 1.0);
    return (float)(acos(cos_A) + acos(cos_B) + acos(cos_C) - M_PI);
}

/* sphere_lambert (utils/monte_carlo.hpp:207-233) */
static float sphere_lambert(float nlcos, float rcos) {
    const float nlsin = (float)sqrt(1.0 - (double)(nlcos * nlcos));
    const float rsin2 = (float)(1.0 - (double)(rcos * rcos));
    const float rsin = sqrtf(rsin2);
    const float ysin = rcos / nlsin;
    const float ycos2 = (float)(1.0 - (double)(ysin * ysin));
    const float ycos = sqrtf(ycos2);
    const float g = (float)((-2.0 * (double)nlsin * (double)rcos * (double)ycos) + HPI - (double)asinf(ysin) +
                            (double)(ysin * ycos));
    const float h = nlcos * ((ycos * sqrtf(rsin2 - ycos2)) + (rsin2 * asinf(ycos / rsin)));
    const float nl = acosf(nlcos);
    const float r = acosf(rcos);
    if (nl < (float)(HPI - (double)r)) return nlcos * rsin2;
    else if (nl < (float)HPI) return (nlcos * rsin2) + g - h;
    else if (nl < (float)(HPI + (double)r)) return (float)((double)(g + h) * INV_PI);
    return 0.0f;
}

/* LambertClosure::estimate_eval_over_solid_angle */
static float lambert_estimate(v3 in, v3 out, float cos_theta, v3 nor) {
    if (cos_theta < 0.0f) return 1.0f;
    v3 nn = v3norm(nor);
    const v3 v = v3norm(out);
    if (v3dot(nn, in) > 0.0f) nn = v3scale(nn, -1.0f);
    const float cos_nv = v3dot(nn, v);
    return sphere_lambert(cos_nv, cos_theta);
}

/* ---- light tree sampling ------------------------------------------------- */
struct gtr_closure;
typedef struct {
    float n, u, v;
    v3 pos, nor, d;
    float wavelength, time;
    mat4 xform;
    uint64_t id; int id_pos;
    v3 to_light;
    float spec[NSPEC];
    float selection_pdf, light_sample_pdf;
    const struct gtr_closure* gtr;  /* the query's closure: GTR, or Lambert when NULL */
} lquery;
/* LightQuery::bsdf->estimate_eval_over_solid_angle */
static float bsdf_estimate(const lquery* q, v3 d, float cos_theta_max);

static float node_prob(const ora_assembly* a, const lquery* q, size_t index) {
    const ora_ltree_node* nd = &a->lnodes[index];
    const bbox bb = bbox_lerp_seq(q->time, nd->bounds, nd->n_bounds);
    const v3 c = v3scale(v3add(bb.min, bb.max), 0.5f);
    const v3 d = v3sub(c, q->pos);
    const float dist2 = v3len2(d);
    const float r = v3len(v3sub(bb.max, bb.min)) * 0.5f;
    const float r2 = r * r;
    const float inv_surface_area = 1.0f / r2;
    float cos_theta_max;
    if (dist2 <= r2) cos_theta_max = -1.0f;
    else {
        const float s2 = smin(1.0f, r2 / dist2);
        cos_theta_max = sqrtf(1.0f - s2);
    }
    const float approx = smax(0.0f, bsdf_estimate(q, d, cos_theta_max));
    return nd->energy * inv_surface_area * approx;
}

static void sphere_light_sample(const ora_object* ob, lquery* q, v3 arr, float* pdf) {
    float na;
    const int i = lerp_seq_index(q->time, ob->ts, &na);
    v3 pos; float radf; float col[3];
    const float* d0 = ob->data;
    if (i < 0) {
        const float* e = d0 + (i == -1 ? 0 : (ob->ts - 1) * 7);
        pos = va(e); radf = e[3]; col[0] = e[4]; col[1] = e[5]; col[2] = e[6];
    } else {
        const float* a = d0 + i * 7;
        const float* b = d0 + (i + 1) * 7;
        pos = lerpv(na, va(a), va(b));
        radf = lerpf(na, a[3], b[3]);
        for (int k = 0; k < 3; ++k) col[k] = (a[4 + k] * lerp_fac(na)) + (b[4 + k] * na);
    }
    const double radius = radf;
    const double surface_area_inv = 1.0 / (4.0 * M_PI * radius * radius);
    v3 x, y, z = v3sub(pos, arr);
    const double d2 = (double)v3len2(z);
    const double d = sqrt(d2);
    coordinate_system_from_vec3(z, &x, &y);
    x = v3norm(x); y = v3norm(y); z = v3norm(z);
    float ccol[3];
    for (int k = 0; k < 3; ++k) ccol[k] = col[k] * (float)surface_area_inv;
    spec_from_color(ccol, q->wavelength, q->spec);
    for (int k = 0; k < NSPEC; ++k) q->spec[k] *= ob->escale;
    if (d > radius) {
        const double s2 = (radius * radius) / d2 < 1.0 ? (radius * radius) / d2 : 1.0;
        const double c2 = 1.0 - s2;
        const double sin_theta_max = sqrt(s2);
        const double cos_theta_max = sqrt(c2);
        const double du = (double)q->u, dv = (double)q->v;
        const double ct = (1.0 - du) + (du * cos_theta_max);
        const double st = sqrt(1.0 - (ct * ct));
        const double phi0 = dv * 2.0 * M_PI;
        v3 sample = v3norm(v3m((float)(cos(phi0) * st), (float)(sin(phi0) * st), (float)ct));
        const double cos_theta = sample.z;
        const double cos_theta2 = cos_theta * cos_theta;
        const double sin_theta2 = 1.0 - cos_theta2 > 0.0 ? 1.0 - cos_theta2 : 0.0;
        const double sin_theta = sqrt(sin_theta2);
        const double D = 1.0 - (d2 * sin_theta * sin_theta / (radius * radius));
        double cos_a;
        if (D <= 0.0) cos_a = sin_theta_max;
        else cos_a = ((d / radius) * sin_theta2) + (cos_theta * sqrt(D));
        const double sa2 = 1.0 - (cos_a * cos_a);
        const double sin_a = sqrt(sa2 > 0.0 ? sa2 : 0.0);
        const double phi = dv * 2.0 * M_PI;
        sample.x = (float)(cos(phi) * sin_a * radius);
        sample.y = (float)(sin(phi) * sin_a * radius);
        sample.z = (float)(d - (cos_a * radius));
        q->to_light = v3add(v3add(v3scale(x, sample.x), v3scale(y, sample.y)), v3scale(z, sample.z));
        *pdf = (float)(1.0 / (2.0 * M_PI * (1.0 - cos_theta_max)));
    } else {
        q->to_light = uniform_sample_sphere(q->u, q->v);
        *pdf = (float)(1.0f / (4.0f * M_PI));
    }
}

static void rect_light_sample(const ora_object* ob, lquery* q, v3 arr, float* pdf) {
    float na;
    const int i = lerp_seq_index(q->time, ob->ts, &na);
    float w, h, col[3];
    const float* d0 = ob->data;
    if (i < 0) {
        const float* e = d0 + (i == -1 ? 0 : (ob->ts - 1) * 5);
        w = e[0]; h = e[1]; col[0] = e[2]; col[1] = e[3]; col[2] = e[4];
    } else {
        const float f = lerp_fac(na);
        const float* a = d0 + i * 5;
        const float* b = d0 + (i + 1) * 5;
        w = (a[0] * f) + (b[0] * na);
        h = (a[1] * f) + (b[1] * na);
        for (int k = 0; k < 3; ++k) col[k] = (a[2 + k] * f) + (b[2 + k] * na);
    }
    const double inv_area = 1.0 / (double)(w * h);
    const v3 p1 = v3norm(v3sub(v3m(w * 0.5f, h * 0.5f, 0.0f), arr));
    const v3 p2 = v3norm(v3sub(v3m(w * -0.5f, h * 0.5f, 0.0f), arr));
    const v3 p3 = v3norm(v3sub(v3m(w * -0.5f, h * -0.5f, 0.0f), arr));
    const v3 p4 = v3norm(v3sub(v3m(w * 0.5f, h * -0.5f, 0.0f), arr));
    const float area_1 = sph_tri_area2(p2, p1, p3);
    const float area_2 = sph_tri_area2(p4, p1, p3);
    const float prob_1 = area_1 / (area_1 + area_2);
    const float prob_2 = 1.0f - prob_1;
    v3 sv;
    if (q->u < prob_1) sv = uniform_sample_spherical_triangle(p2, p1, p3, q->v, q->u / prob_1);
    else sv = uniform_sample_spherical_triangle(p4, p1, p3, q->v, 1.0f - ((q->u - prob_1) / prob_2));
    sv = v3scale(sv, -arr.z / sv.z);
    q->to_light = sv;
    *pdf = 1.0f / (area_1 + area_2);
    float ccol[3];
    for (int k = 0; k < 3; ++k) ccol[k] = (col[k] * (float)inv_area) * 0.5f;
    spec_from_color(ccol, q->wavelength, q->spec);
    for (int k = 0; k < NSPEC; ++k) q->spec[k] *= ob->escale;
}

/* LightTree::sample (accel/light_tree.cpp:30-96) */
static int light_tree_sample(const ora_assembly* a, lquery* q) {
    if (a->n_lnodes == 0) return 0;
    const ora_ltree_node* nd = &a->lnodes[0];
    float tot_prob = 1.0f;
    while (!nd->is_leaf) {
        float p1 = node_prob(a, q, nd->index1);
        float p2 = node_prob(a, q, nd->index2);
        const float total = p1 + p2;
        if (total <= 0.0f) { p1 = 0.5f; p2 = 0.5f; }
        else { p1 /= total; p2 /= total; }
        if (q->n <= p1) {
            tot_prob *= p1;
            nd = &a->lnodes[nd->index1];
            q->n /= p1;
        } else {
            tot_prob *= p2;
            nd = &a->lnodes[nd->index2];
            q->n = (q->n - p1) / p2;
        }
    }
    const psy_instance_desc* inst = &a->inst[nd->instance_index];
    const int bits = asm_id_bits(a);
    q->id = (q->id << bits) | ((uint64_t)nd->instance_index & (uint64_t)((1 << bits) - 1));
    if (inst->transform_count > 0) {
        const mat4 xf = mat4_lerp_seq(q->time, a->xf + inst->transform_index, inst->transform_count);
        q->pos = mat4_pos(&xf, q->pos);
        const mat4 inv = ora_mat4_inverse(&xf);
        q->nor = v3norm(mat4_dir_transpose(&inv, q->nor));
        q->xform = mat4_mul(&q->xform, &xf);
    }
    if (inst->type == PSY_INSTANCE_OBJECT) {
        const ora_object* ob = &a->obj[inst->data_index];
        float p = 1.0f;
        if (ob->kind == PSY_OBJ_SPHERE_LIGHT) sphere_light_sample(ob, q, q->pos, &p);
        else if (ob->kind == PSY_OBJ_RECTANGLE_LIGHT) rect_light_sample(ob, q, q->pos, &p);
        else return 0;
        const mat4 inv = ora_mat4_inverse(&q->xform);
        q->to_light = mat4_dir(&inv, q->to_light);
        q->selection_pdf *= (tot_prob * (float)a->total_lights);
        q->light_sample_pdf = p;
        return 1;
    }
    const ora_assembly* sub = a->sub[inst->data_index];
    q->selection_pdf *= (tot_prob * (float)a->total_lights) / (float)sub->total_lights;
    return light_tree_sample(sub, q);
}

/* ---- the integrator --------------------------------------------------------- */
typedef struct {
    uint32_t off, dim;
    float time;
    int step;
    int path_length;   /* bounces (the reference integrator's 4 unless set) */
    short px, py;
    ora_inter inter;
    ora_wray prev_ray;
    float wavelength;
    float last_pdf;
    float col[NSPEC], fcol[NSPEC], lcol[NSPEC];
    int done;
} path_t;

/* the Sobol form of the sampler (the device's PSY_RENDER_SOBOL; no
 * counterpart in the reference): Sobol::sample (sampling/sobol.hpp:40-53)
 * of dimension d at the Sampler's index, unscrambled, as a float below 1;
 * set for the length of one ora_render_sobol call */
static const uint64_t* g_sobol_m = NULL;
static uint32_t g_sobol_cols = 0;
static inline float next_dim(path_t* p) {
    if (g_sobol_m) {
        const float v = (float)ora_sobol(g_sobol_m, g_sobol_cols, p->dim++, (uint64_t)p->off, 0);
        return v < 1.0f ? v : 0.99999994f;
    }
    return ora_sampler_get(p->off, p->dim++);
}

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

void ora_camera_ray_from_dims(const ora_scene* s, float rx, float ry, float dx, float dy, float time,
                              float u, float v, ora_wray* w);
void ora_bounce_dir(const ora_inter* h, const ora_wray* prev, float su, float sv, ora_wray* out,
                    float* pdf, v3* world_p, v3* world_n);

/* world-space differential geometry of a stored hit */
static void hit_world(const ora_inter* h, v3* p, v3* n, v3* pos_offset) {
    mat4 sp;
    memcpy(&sp.m[0][0], h->space, sizeof(float) * 16);
    const mat4 inv = ora_mat4_inverse(&sp);
    *p = mat4_pos(&inv, va(h->p));
    v3 nn = mat4_dir_transpose(&sp, va(h->n));
    const float il = 1.0f / v3len(nn);
    *n = v3scale(nn, il);
    v3 po = mat4_dir(&inv, va(h->offset));
    if (h->backfacing) po = v3scale(po, -1.0f);
    *pos_offset = po;
}

static v3 transfer_origin_diff(float t, v3 normal, v3 d, v3 od, v3 dd) {
    const v3 temp = v3add(od, v3scale(dd, t));
    const float td = -v3dot(temp, normal) / v3dot(d, normal);
    const v3 real_projected = v3add(temp, v3scale(d, td));
    return v3scale(v3norm(real_projected), v3len(temp));
}

/* LambertClosure::propagate_differentials */
static void lambert_differentials(float t, const ora_wray* in, v3 gn, ora_wray* out) {
    const float len = v3len(va(out->d));
    const v3 nn = v3norm(gn);
    v3 x, y;
    coordinate_system_from_vec3(va(out->d), &x, &y);
    x = v3norm(x);
    y = v3norm(y);
    set3(out->odx, transfer_origin_diff(t, nn, va(in->d), va(in->odx), va(in->ddx)));
    set3(out->ody, transfer_origin_diff(t, nn, va(in->d), va(in->ody), va(in->ddy)));
    set3(out->ddx, v3sdiv(v3scale(x, 0.15f), len));
    set3(out->ddy, v3sdiv(v3scale(y, 0.15f), len));
}

static void lambert_eval(const float albedo[3], v3 in, v3 out, v3 gn, float hero, float e[NSPEC]) {
    v3 nn = v3norm(gn);
    const v3 v = v3norm(out);
    if (v3dot(nn, in) > 0.0f) nn = v3scale(nn, -1.0f);
    const float fac = smax(v3dot(nn, v), 0.0f) * (float)INV_PI;
    const float c[3] = {albedo[0] * fac, albedo[1] * fac, albedo[2] * fac};
    spec_from_color(c, hero, e);
}

static float lambert_pdf(v3 in, v3 out, v3 gn) {
    v3 nn = v3norm(gn);
    const v3 v = v3norm(out);
    if (v3dot(nn, v3norm(in)) > 0.0f) nn = v3scale(nn, -1.0f);
    return smax(v3dot(nn, v) * (float)INV_PI, 0.0f);
}

/* ---- GTRClosure (shading/surface_closure.hpp:339-608) -------------------- */
typedef struct gtr_closure { float col[3]; float roughness, tail_shape, fresnel, norm; } gtr_closure;

/* clamp (utils/utils.hpp:42-44): std::min(b, std::max(a, v)) */
static float uclamp(float v, float a, float b) { return smin(b, smax(a, v)); }

/* GTRClosure::normalization (:376-382), the pi product in double */
static float gtr_normalization(float r, float t) {
    const float r2 = r * r;
    const float top = (t - 1.0f) * (r2 - 1);
    const double bottom = M_PI * (double)(1.0f - powf(r2, 1.0f - t));
    return (float)((double)top / bottom);
}

/* The GTRClosure(col, roughness, tail_shape, fresnel) constructor with
 * validate (:351-372) */
static gtr_closure gtr_make(const float col[3], const float prm[3]) {
    gtr_closure g;
    memcpy(g.col, col, sizeof g.col);
    g.roughness = smax(0.0f, smin(0.9999f, prm[0]));
    g.tail_shape = smax(0.0001f, prm[1]);
    g.fresnel = prm[2];
    if (g.roughness < 0.000244140625f) g.roughness = 0.0f;
    if (fabsf(g.tail_shape - 1.0f) < 0.0001f) g.tail_shape = 1.0f + 0.0001f;
    g.norm = gtr_normalization(g.roughness, g.tail_shape);
    return g;
}

/* GTRClosure::dist (:398-410) */
static float gtr_dist(const gtr_closure* g, float nh, float rough) {
    const float r2 = rough * rough;
    float D = 0.0f;
    if (nh > 0.0f) {
        const float nh2 = nh * nh;
        D = gtr_normalization(rough, g->tail_shape) / powf(1 + ((r2 - 1) * nh2), g->tail_shape);
    }
    return D;
}

/* schlick_fresnel_from_fac (:173-177) */
static float schlick_from_fac(float f, float c) {
    const float c1 = (1.0f - c);
    const float c2 = c1 * c1;
    return f + ((1.0f - f) * c1 * c2 * c2);
}

/* lerp (utils/utils.hpp:52-55) for floats: the 1.0 - alpha in double */
static float lerp_f(float alpha, float a, float b) {
    return (float)((double)a * (1.0 - (double)alpha) + (double)(b * alpha));
}

/* GTRClosure::evaluate (:452-517) */
static void gtr_eval(const gtr_closure* g, v3 in, v3 out, v3 gn, float hero, float e[NSPEC]) {
    v3 nn = v3norm(gn);
    const v3 aa = v3scale(v3norm(in), -1.0f);
    const v3 bb = v3norm(out);
    const v3 hh = v3norm(v3add(aa, bb));
    if (v3dot(nn, hh) < 0.0f) nn = v3scale(nn, -1.0f);
    const float na = uclamp(v3dot(nn, aa), -1.0f, 1.0f);
    const float nb = uclamp(v3dot(nn, bb), -1.0f, 1.0f);
    const float ha = uclamp(v3dot(hh, aa), -1.0f, 1.0f);
    const float hb = uclamp(v3dot(hh, bb), -1.0f, 1.0f);
    const float nh = uclamp(v3dot(nn, hh), -1.0f, 1.0f);
    const float r2 = g->roughness * g->roughness;
    float D = 1.0f, G1 = 1.0f, G2 = 1.0f;
    spec_from_color(g->col, hero, e);
    for (int i = 0; i < NSPEC; ++i) e[i] = lerp_f(1.0f - g->fresnel, schlick_from_fac(e[i], hb), e[i]);
    if (g->roughness == 0.0f) return;
    if (nh > 0.0f) {
        const float nh2 = nh * nh;
        D = g->norm / powf(1 + ((r2 - 1) * nh2), g->tail_shape);
    }
    {
        const float na2 = na * na;
        const float tan_na = sqrtf((1.0f - na2) / na2);
        const float pos = (ha * na) > 0.0f ? 1.0f : 0.0f;
        const float a = r2 * tan_na;
        const float b = (sqrtf(1.0f + a * a) - 1.0f) * 0.5f;
        G1 = pos / (1.0f + b);
    }
    {
        const float nb2 = nb * nb;
        const float tan_nb = sqrtf((1.0f - nb2) / nb2);
        const float pos = (hb * nb) > 0.0f ? 1.0f : 0.0f;
        const float a = r2 * tan_nb;
        const float b = (sqrtf(1.0f + a * a) - 1.0f) * 0.5f;
        G2 = pos / (1.0f + b);
    }
    const float f = D * G1 * G2;
    for (int i = 0; i < NSPEC; ++i) e[i] = (e[i] * f) * (float)INV_PI;
}

/* GTRClosure::sample_pdf (:556-571) */
static float gtr_pdf(const gtr_closure* g, v3 in, v3 out, v3 gn) {
    v3 nn = v3norm(gn);
    const v3 aa = v3scale(v3norm(in), -1.0f);
    const v3 bb = v3norm(out);
    const v3 hh = v3norm(v3scale(v3add(aa, bb), 0.5f));
    if (v3dot(nn, hh) < 0.0f) nn = v3scale(nn, -1.0f);
    const float nh = uclamp(v3dot(nn, hh), -1.0f, 1.0f);
    return gtr_dist(g, nh, g->roughness) * (float)INV_PI;
}

/* GTRClosure::half_theta_sample (:385-396) */
static float gtr_half_theta_sample(const gtr_closure* g, float u) {
    const float r2 = g->roughness * g->roughness;
    const float top = 1.0f - powf((powf(r2, 1.0f - g->tail_shape) * (1.0f - u)) + u, 1.0f / (1.0f - g->tail_shape));
    const float bottom = (1.0f - r2);
    return sqrtf(top / bottom);
}

/* GTRClosure::sample (:418-437): the direction */
static v3 gtr_sample_dir(const gtr_closure* g, v3 in, v3 gn, float si, float sj) {
    v3 nn = v3norm(gn);
    if (v3dot(nn, in) > 0.0f) nn = v3scale(nn, -1.0f);
    const float theta_cos = gtr_half_theta_sample(g, si);
    const float theta_sin = sqrtf(1.0f - (theta_cos * theta_cos));
    const float angle = (float)((double)sj * M_PI * 2);
    v3 half_dir = v3m(cosf(angle) * theta_sin, sinf(angle) * theta_sin, theta_cos);
    half_dir = v3norm(zup_to_vec(half_dir, nn));
    return v3sub(in, v3scale(v3scale(half_dir, 2.0f), v3dot(in, half_dir)));
}

/* GTRClosure::estimate_eval_over_solid_angle (:520-553) */
static float gtr_estimate(const gtr_closure* g, v3 in, v3 cent, float cos_theta, v3 nor) {
    v3 nn = v3norm(nor);
    const v3 aa = v3scale(v3norm(in), -1.0f);
    const v3 bb = v3norm(cent);
    if (v3dot(nn, in) > 0.0f) nn = v3scale(nn, -1.0f);
    const float theta = acosf(cos_theta);
    const v3 hh = v3norm(v3add(aa, bb));
    const float nh = uclamp(v3dot(nn, hh), -1.0f, 1.0f);
    const float fac = gtr_dist(g, nh, smin(1.0f, sqrtf(g->roughness) + (2.0f * theta / (float)M_PI)));
    return fac * smin(1.0f - cos_theta, 1.0f) * (float)INV_PI;
}

/* calc_uv_differentials (basics/differential_geometry.hpp:145-156) */
static void uv_differentials(v3 dp, v3 dpdu, v3 dpdv, float* du, float* dv) {
    const float dpdu_ilen = 1.0f / v3len(dpdu);
    const v3 dpdu_n = v3scale(dpdu, dpdu_ilen);
    const float dpdv_ilen = 1.0f / v3len(dpdv);
    const v3 dpdv_n = v3scale(dpdv, dpdv_ilen);
    *du = v3dot(dp, dpdu_n) * dpdu_ilen;
    *dv = v3dot(dp, dpdv_n) * dpdv_ilen;
}

/* The rotation about `axis` by `angle`, v' = v M, by Rodrigues' formula as
 * Imath's Matrix44::setAxisAngle lays it out (make_axis_angle_transform,
 * math/transform.hpp:221-226; Imath is not part of the reference tree) */
static mat4 axis_angle(v3 axis, float angle) {
    const v3 u = v3norm(axis);
    const float sn = sinf(angle), cs = cosf(angle), oc = 1 - cs;
    mat4 m = mat4_identity();
    m.m[0][0] = u.x * u.x * oc + cs;
    m.m[0][1] = u.x * u.y * oc + u.z * sn;
    m.m[0][2] = u.x * u.z * oc - u.y * sn;
    m.m[1][0] = u.x * u.y * oc - u.z * sn;
    m.m[1][1] = u.y * u.y * oc + cs;
    m.m[1][2] = u.y * u.z * oc + u.x * sn;
    m.m[2][0] = u.x * u.z * oc + u.y * sn;
    m.m[2][1] = u.y * u.z * oc - u.x * sn;
    m.m[2][2] = u.z * u.z * oc + cs;
    return m;
}

/* reflect_ray_direction_differential (basics/differential_geometry.hpp:113-117) */
static v3 reflect_dd(v3 normal, v3 normal_d, v3 d, v3 dd) {
    const float ddn = v3dot(dd, normal) + v3dot(d, normal_d);
    const v3 tmp = v3add(v3scale(normal_d, v3dot(d, normal)), v3scale(normal, ddn));
    return v3sub(dd, v3scale(tmp, 2.0f));
}

/* GTRClosure::propagate_differentials (:574-606) and clamp_dd
 * (basics/differential_geometry.hpp:125-135); gn, dpdu.. world space */
static void gtr_differentials(float t, const ora_wray* in, v3 gn, v3 dpdu, v3 dpdv, v3 dndu, v3 dndv,
                              ora_wray* out) {
    v3 nn = v3norm(gn);
    const v3 hh = v3norm(v3add(v3norm(v3scale(va(in->d), -1.0f)), v3norm(va(out->d))));
    const v3 odx = transfer_origin_diff(t, nn, va(in->d), va(in->odx), va(in->ddx));
    const v3 ody = transfer_origin_diff(t, nn, va(in->d), va(in->ody), va(in->ddy));
    float dux, dvx, duy, dvy;
    uv_differentials(odx, dpdu, dpdv, &dux, &dvx);
    uv_differentials(ody, dpdu, dpdv, &duy, &dvy);
    v3 dndx = v3add(v3scale(dndu, dux), v3scale(dndv, dvx));
    v3 dndy = v3add(v3scale(dndu, duy), v3scale(dndv, dvy));
    if (v3dot(nn, hh) < 0.0f) {
        nn = v3scale(nn, -1.0f);
        dndx = v3scale(dndx, -1.0f);
        dndy = v3scale(dndy, -1.0f);
    }
    const v3 axis = v3norm(v3cross(nn, hh));
    const float angle = acosf(uclamp(v3dot(nn, hh), 0.0f, 1.0f));
    const mat4 xf = axis_angle(axis, angle);
    const v3 dhdx = mat4_dir(&xf, dndx);
    const v3 dhdy = mat4_dir(&xf, dndy);
    v3 ddx = reflect_dd(hh, dhdx, va(in->d), va(in->ddx));
    v3 ddy = reflect_dd(hh, dhdy, va(in->d), va(in->ddy));
    const float len_d = v3len(va(out->d));
    const float len_dx = v3len(ddx);
    const float len_dy = v3len(ddy);
    if ((len_dx / len_d) > 0.9f) ddx = v3scale(ddx, 0.9f * len_d / len_dx);
    if ((len_dy / len_d) > 0.9f) ddy = v3scale(ddy, 0.9f * len_d / len_dy);
    set3(out->odx, odx);
    set3(out->ody, ody);
    set3(out->ddx, ddx);
    set3(out->ddy, ddy);
}

static float bsdf_estimate(const lquery* q, v3 d, float cos_theta_max) {
    return q->gtr ? gtr_estimate(q->gtr, q->d, d, cos_theta_max, q->nor)
                  : lambert_estimate(q->d, d, cos_theta_max, q->nor);
}

/* world-space dp/du, dp/dv, dn/du, dn/dv of a stored hit
 * (DifferentialGeometry::transformed_from, basics/differential_geometry.hpp:20-40) */
static void hit_world_derivs(const ora_inter* h, v3* dpdu, v3* dpdv, v3* dndu, v3* dndv) {
    mat4 sp;
    memcpy(&sp.m[0][0], h->space, sizeof(float) * 16);
    const mat4 inv = ora_mat4_inverse(&sp);
    *dpdu = mat4_dir(&inv, va(h->dpdu));
    *dpdv = mat4_dir(&inv, va(h->dpdv));
    const float il = 1.0f / v3len(mat4_dir_transpose(&sp, va(h->n)));
    *dndu = v3scale(mat4_dir_transpose(&sp, va(h->dndu)), il);
    *dndv = v3scale(mat4_dir_transpose(&sp, va(h->dndv)), il);
}

static void init_path(path_t* p, uint32_t off, short x, short y) {
    memset(p, 0, sizeof *p);
    p->off = off;
    p->px = x;
    p->py = y;
    p->wavelength = (next_dim(p) * (WL_MAX - WL_MIN)) + WL_MIN;
    for (int i = 0; i < NSPEC; ++i) { p->col[i] = 0.0f; p->fcol[i] = 1.0f; p->lcol[i] = 0.0f; }
}

/* next_ray_for_path (integrator/path_trace_integrator.cpp:106-226) */
static void next_ray(const ora_scene* s, path_t* p, int width, int height, ora_wray* ray) {
    memset(ray, 0, sizeof *ray);
    if (p->step == 0) {
        const float min_x = -1.0f, max_x = 1.0f;
        const float max_y = ((float)height) / (float)width, min_y = -max_y;
        const float samp_x = (fast_logit(next_dim(p), 1.5f) + 0.5f + (float)p->px) / (float)width;
        const float samp_y = (fast_logit(next_dim(p), 1.5f) + 0.5f + (float)p->py) / (float)height;
        const float rx = (float)(((double)samp_x - 0.5) * (double)(max_x - min_x));
        const float ry = (float)((0.5 - (double)samp_y) * (double)(max_y - min_y));
        const float dx = (max_x - min_x) / (float)width;
        const float dy = (max_y - min_y) / (float)height;
        const float sa = next_dim(p), sb = next_dim(p), sc = next_dim(p);
        ora_camera_ray_from_dims(s, rx, ry, dx, dy, sc, sa, sb, ray);
        p->time = ray->time;
    } else if (p->step % 2 && p->inter.closure == PSY_SHADER_GTR &&
               gtr_make(p->inter.closure_col, p->inter.closure_param).roughness == 0.0f) {
        /* shadow ray of a delta closure (a GTR mirror): no light sample and
         * no sampler dimensions; the default WorldRay the reference then
         * traces has no effect (update_path ignores its result), so a ray
         * of zero light and length is traced here */
        for (int i = 0; i < NSPEC; ++i) p->lcol[i] = 0.0f;
        v3 gp, gn, po;
        hit_world(&p->inter, &gp, &gn, &po);
        set3(ray->o, v3add(gp, po));
        set3(ray->d, v3m(1e-20f, 1e-20f, 1e-20f));
        ray->time = p->time;
        ray->type = PSY_RAY_OCCLUSION;
    } else if (p->step % 2) {
        /* shadow ray towards a light sample */
        const int is_gtr = p->inter.closure == PSY_SHADER_GTR;
        const gtr_closure g = is_gtr ? gtr_make(p->inter.closure_col, p->inter.closure_param) : (gtr_closure){{0}};
        v3 gp, gn, po;
        hit_world(&p->inter, &gp, &gn, &po);
        v3 lq_nor = gn;
        if (p->inter.backfacing) lq_nor = v3scale(lq_nor, -1.0f);
        lquery lq;
        memset(&lq, 0, sizeof lq);
        lq.n = next_dim(p); lq.u = next_dim(p); lq.v = next_dim(p);
        lq.pos = gp; lq.nor = lq_nor; lq.d = va(p->prev_ray.d);
        lq.wavelength = p->wavelength; lq.time = p->time;
        lq.xform = mat4_identity();
        lq.selection_pdf = 1.0f;
        lq.light_sample_pdf = 1.0f;
        lq.gtr = is_gtr ? &g : NULL;
        const ora_assembly* root = &s->asms[0];
        if (!light_tree_sample(root, &lq)) {
            for (int i = 0; i < NSPEC; ++i) p->lcol[i] = 0.0f;
            lq.to_light = v3m(0.0f, 0.0f, 1.0f);
        } else {
            const float bsdf_pdf = is_gtr ? gtr_pdf(&g, va(p->prev_ray.d), lq.to_light, gn)
                                          : lambert_pdf(va(p->prev_ray.d), lq.to_light, gn);
            const float a2 = lq.light_sample_pdf * lq.light_sample_pdf;
            const float b2 = bsdf_pdf * bsdf_pdf;
            const float mis_inv_pdf = (a2 / (a2 + b2)) / lq.light_sample_pdf;
            for (int i = 0; i < NSPEC; ++i)
                p->lcol[i] = ((lq.spec[i] * mis_inv_pdf) * (float)root->total_lights) / lq.selection_pdf;
        }
        set3(ray->o, v3add(gp, po));
        set3(ray->d, v3scale(v3sub(lq.to_light, po), 0.9995f));
        ray->time = p->time;
        ray->type = PSY_RAY_OCCLUSION;
        if (is_gtr) {
            v3 dpdu, dpdv, dndu, dndv;
            hit_world_derivs(&p->inter, &dpdu, &dpdv, &dndu, &dndv);
            gtr_differentials(p->inter.t, &p->prev_ray, gn, dpdu, dpdv, dndu, dndv, ray);
        } else {
            lambert_differentials(p->inter.t, &p->prev_ray, gn, ray);
        }
    } else if (p->inter.closure == PSY_SHADER_GTR) {
        /* bounce ray: GTRClosure::sample */
        const gtr_closure g = gtr_make(p->inter.closure_col, p->inter.closure_param);
        const float su = next_dim(p), sv = next_dim(p);
        v3 gp, gn, po, dpdu, dpdv, dndu, dndv;
        hit_world(&p->inter, &gp, &gn, &po);
        hit_world_derivs(&p->inter, &dpdu, &dpdv, &dndu, &dndv);
        const v3 out = gtr_sample_dir(&g, va(p->prev_ray.d), gn, su, sv);
        float pdf = gtr_pdf(&g, va(p->prev_ray.d), out, gn);
        float filt[NSPEC];
        gtr_eval(&g, va(p->prev_ray.d), out, gn, p->wavelength, filt);
        set3(ray->o, v3add(gp, po));
        set3(ray->d, out);
        ray->time = p->time;
        ray->type = 2;
        gtr_differentials(p->inter.t, &p->prev_ray, gn, dpdu, dpdv, dndu, dndv, ray);
        if (g.roughness != 0.0f) {
            if (pdf < 0.00001f) pdf = 0.00001f;
            for (int i = 0; i < NSPEC; ++i) p->fcol[i] *= filt[i] / pdf;
            p->last_pdf = pdf;
        } else {
            for (int i = 0; i < NSPEC; ++i) p->fcol[i] *= filt[i];
            p->last_pdf = 0.0f;
        }
    } else {
        /* bounce ray: LambertClosure::sample */
        const float su = next_dim(p), sv = next_dim(p);
        float pdf;
        v3 gp, gn;
        ora_bounce_dir(&p->inter, &p->prev_ray, su, sv, ray, &pdf, &gp, &gn);
        ray->time = p->time;
        ray->type = 2;
        float filt[NSPEC];
        lambert_eval(p->inter.closure_col, va(p->prev_ray.d), va(ray->d), gn, p->wavelength, filt);
        if (pdf < 0.00001f) pdf = 0.00001f;
        for (int i = 0; i < NSPEC; ++i) p->fcol[i] *= filt[i] / pdf;
        p->last_pdf = pdf;
    }
}

/* update_path (integrator/path_trace_integrator.cpp:232-279) */
static void update_path(const ora_scene* s, path_t* p, const ora_wray* ray, const ora_inter* in) {
    if (p->step % 2) {
        if (!in->hit) {
            v3 gp, gn, po;
            hit_world(&p->inter, &gp, &gn, &po);
            float fac[NSPEC];
            if (p->inter.closure == PSY_SHADER_GTR) {
                const gtr_closure g = gtr_make(p->inter.closure_col, p->inter.closure_param);
                if (g.roughness == 0.0f) goto next;  /* a delta closure adds no light sample */
                gtr_eval(&g, va(p->prev_ray.d), va(ray->d), gn, p->wavelength, fac);
            } else {
                lambert_eval(p->inter.closure_col, va(p->prev_ray.d), va(ray->d), gn, p->wavelength, fac);
            }
            for (int i = 0; i < NSPEC; ++i) p->col[i] += p->fcol[i] * p->lcol[i] * fac[i];
        }
    } else {
        if (in->hit) {
            if (in->closure == PSY_SHADER_EMIT) {
                p->done = 1;
                float em[NSPEC];
                spec_from_color(in->closure_col, p->wavelength, em);
                if (p->last_pdf != 0.0f) {
                    const float a2 = p->last_pdf * p->last_pdf;
                    const float b2 = in->light_pdf * in->light_pdf;
                    const float mis_inv_pdf = (a2 / (a2 + b2)) / p->last_pdf;
                    for (int i = 0; i < NSPEC; ++i) p->col[i] += ((p->fcol[i] * p->last_pdf) * em[i]) * mis_inv_pdf;
                } else {
                    for (int i = 0; i < NSPEC; ++i) p->col[i] += p->fcol[i] * em[i];
                }
            } else {
                p->inter = *in;
                p->prev_ray = *ray;
            }
        } else {
            p->done = 1;
            float bg[NSPEC];
            spec_from_color(s->background, p->wavelength, bg);
            for (int i = 0; i < NSPEC; ++i) p->col[i] += p->fcol[i] * bg[i];
        }
    }
next:
    p->step++;
    if (p->step == p->path_length * 2) p->done = 1;
}

/* Color_XYZ(SpectralSample) * XYZ_NORM_FAC and Film::add_sample */
static void path_xyz(const path_t* p, float out[3]) {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    for (int i = 0; i < NSPEC; ++i) {
        const float w = wavelength_n(p->wavelength, i);
        x += X_1931(w) * p->col[i];
        y += Y_1931(w) * p->col[i];
        z += Z_1931(w) * p->col[i];
    }
    x *= 1.0f / NSPEC; y *= 1.0f / NSPEC; z *= 1.0f / NSPEC;
    out[0] = x * XYZ_NORM_FAC; out[1] = y * XYZ_NORM_FAC; out[2] = z * XYZ_NORM_FAC;
}

typedef struct {
    const ora_scene* s;
    int width, height, spp;
    uint32_t seed;
    int path_length;
    int bucket;
    int nbx, nby;
    int next;
    pthread_mutex_t mu;
    float* path_xyz;    /* width*height*spp*3, ordered (y, x, s) */
    int64_t rays;
} render_job;

static void* render_worker(void* vp) {
    render_job* J = (render_job*)vp;
    const int bs = J->bucket;
    const int64_t cap = (int64_t)bs * bs * J->spp;
    path_t* paths = (path_t*)malloc(sizeof(path_t) * cap);
    ora_wray* rays = (ora_wray*)malloc(sizeof(ora_wray) * cap);
    ora_inter* inters = (ora_inter*)malloc(sizeof(ora_inter) * cap);
    int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * cap);
    int64_t traced = 0;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        const int b = J->next++;
        pthread_mutex_unlock(&J->mu);
        if (b >= J->nbx * J->nby) break;
        const int bx = (b % J->nbx) * bs, by = (b / J->nbx) * bs;
        const int bw = (bx + bs > J->width) ? J->width - bx : bs;
        const int bh = (by + bs > J->height) ? J->height - by : bs;
        int64_t n = 0;
        for (int x = bx; x < bx + bw; ++x)
            for (int y = by; y < by + bh; ++y)
                for (int s = 0; s < J->spp; ++s) {
                    init_path(&paths[n], ora_sampler_offset((uint32_t)x, (uint32_t)y, (uint32_t)s, J->seed), (short)x,
                              (short)y);
                    paths[n].path_length = J->path_length;
                    idx[n] = ((int64_t)y * J->width + x) * J->spp + s;
                    ++n;
                }
        int64_t active = n;
        /* active paths are kept first; order does not change per-path results */
        int64_t* act = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
        for (int64_t i = 0; i < n; ++i) act[i] = i;
        while (active > 0) {
            for (int64_t i = 0; i < active; ++i) next_ray(J->s, &paths[act[i]], J->width, J->height, &rays[i]);
            ora_trace(J->s, rays, active, inters, 1, active);
            traced += active;
            for (int64_t i = 0; i < active; ++i) update_path(J->s, &paths[act[i]], &rays[i], &inters[i]);
            int64_t k = 0;
            for (int64_t i = 0; i < active; ++i)
                if (!paths[act[i]].done) act[k++] = act[i];
            active = k;
        }
        free(act);
        for (int64_t i = 0; i < n; ++i) path_xyz(&paths[i], &J->path_xyz[idx[i] * 3]);
    }
    pthread_mutex_lock(&J->mu);
    J->rays += traced;
    pthread_mutex_unlock(&J->mu);
    free(paths); free(rays); free(inters); free(idx);
    return NULL;
}

/* Renders every pixel sample; path_xyz receives the film contribution of
 * each sample (Color_XYZ(col) * XYZ_NORM_FAC), ordered (y, x, s).  Returns
 * the number of rays traced (camera + bounce + shadow), or -1 when a
 * chromatic colour was encountered (colour upsampling not restated). */
int64_t ora_render(const ora_scene* s, int width, int height, int spp, uint32_t seed, int path_length, int threads,
                   float* path_xyz_out) {
    render_job J;
    memset(&J, 0, sizeof J);
    J.s = s; J.width = width; J.height = height; J.spp = spp; J.seed = seed; J.path_length = path_length;
    int bsz = (int)sqrtf(65536.0f / (float)spp);
    const int maxb = (int)sqrtf((float)(width * height) / ((float)threads * 4.0f));
    if (bsz > maxb) bsz = maxb;
    if (bsz < 1) bsz = 1;
    J.bucket = bsz;
    J.nbx = (width + bsz - 1) / bsz;
    J.nby = (height + bsz - 1) / bsz;
    pthread_mutex_init(&J.mu, NULL);
    J.path_xyz = path_xyz_out;
    g_chromatic_seen = 0;
    if (threads < 1) threads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, render_worker, &J);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&J.mu);
    if (g_chromatic_seen) return -1;
    return J.rays;
}

int64_t ora_render_sobol(const ora_scene* s, int width, int height, int spp, uint32_t seed, int path_length,
                         int threads, float* path_xyz, const uint64_t* m, uint32_t n_cols) {
    g_sobol_m = m;
    g_sobol_cols = n_cols;
    const int64_t r = ora_render(s, width, height, spp, seed, path_length, threads, path_xyz);
    g_sobol_m = NULL;
    g_sobol_cols = 0;
    return r;
}

/* ---- entry points of the oracle's own checks (tests/test_oracle_pins.py) */
void ora_kat_cie_1931(float lambda, float xyz[3]) {
    xyz[0] = X_1931(lambda);
    xyz[1] = Y_1931(lambda);
    xyz[2] = Z_1931(lambda);
}

float ora_kat_sphere_lambert(float nlcos, float rcos) { return sphere_lambert(nlcos, rcos); }

/* LambertClosure eval (flat spectrum of a grey albedo, first wavelength)
 * and sample_pdf for incoming direction `in`, outgoing `out`, normal n */
void ora_kat_lambert(float albedo, const float in[3], const float out[3], const float n[3], float* eval,
                     float* pdf) {
    const float a[3] = {albedo, albedo, albedo};
    float e[NSPEC];
    lambert_eval(a, va(in), va(out), va(n), 500.0f, e);
    *eval = e[0];
    *pdf = lambert_pdf(va(in), va(out), va(n));
}

static ora_object kat_light(int kind, const float* data, int nf) {
    ora_object ob;
    memset(&ob, 0, sizeof ob);
    ob.kind = kind;
    ob.ts = 1;
    ob.escale = 1.0f;
    ob.data = (float*)data;
    (void)nf;
    return ob;
}

/* SphereLight::sample from point arr: the vector to the sampled point and
 * the sample's pdf (grey colour 1) */
void ora_kat_sphere_light_sample(const float c[3], float radius, const float arr[3], float u, float v,
                                 float to_light[3], float* pdf) {
    const float data[7] = {c[0], c[1], c[2], radius, 1.0f, 1.0f, 1.0f};
    ora_object ob = kat_light(PSY_OBJ_SPHERE_LIGHT, data, 7);
    lquery q;
    memset(&q, 0, sizeof q);
    q.u = u; q.v = v; q.wavelength = 500.0f;
    sphere_light_sample(&ob, &q, va(arr), pdf);
    set3(to_light, q.to_light);
}

/* RectangleLight::sample (the light in its own space, centred at the
 * origin in the z = 0 plane) */
void ora_kat_rect_light_sample(float w, float h, const float arr[3], float u, float v, float to_light[3],
                               float* pdf) {
    const float data[5] = {w, h, 1.0f, 1.0f, 1.0f};
    ora_object ob = kat_light(PSY_OBJ_RECTANGLE_LIGHT, data, 5);
    lquery q;
    memset(&q, 0, sizeof q);
    q.u = u; q.v = v; q.wavelength = 500.0f;
    rect_light_sample(&ob, &q, va(arr), pdf);
    set3(to_light, q.to_light);
}

/* LightTree::sample on the scene's root tree from a shading point: the
 * selected light's InstanceID, the selection pdf (times the root's light
 * count, as the reference keeps it), the light-sample pdf and the vector
 * to the light.  Returns 0 when the scene has no light. */
int ora_kat_light_select(const ora_scene* s, const float pos[3], const float nor[3], const float d[3], float time,
                         float n, float u, float v, uint64_t* id, float* selection_pdf, float* light_sample_pdf,
                         float to_light[3]) {
    lquery q;
    memset(&q, 0, sizeof q);
    q.n = n; q.u = u; q.v = v;
    q.pos = va(pos); q.nor = va(nor); q.d = va(d);
    q.wavelength = 500.0f; q.time = time;
    q.xform = mat4_identity();
    q.selection_pdf = 1.0f;
    q.light_sample_pdf = 1.0f;
    const int ok = light_tree_sample(&s->asms[0], &q);
    *id = q.id;
    *selection_pdf = q.selection_pdf;
    *light_sample_pdf = q.light_sample_pdf;
    set3(to_light, q.to_light);
    return ok;
}

int64_t ora_kat_total_lights(const ora_scene* s) { return (int64_t)s->asms[0].total_lights; }

/* ---- GTR closure pieces, for the oracle's own checks --------------------- */
/* GTRClosure(col, prm) validated: roughness, tail shape, normalization */
void ora_kat_gtr_params(const float prm[3], float out[3]) {
    const float c[3] = {1.0f, 1.0f, 1.0f};
    const gtr_closure g = gtr_make(c, prm);
    out[0] = g.roughness; out[1] = g.tail_shape; out[2] = g.norm;
}
/* dist(nh, roughness) */
float ora_kat_gtr_dist(const float prm[3], float nh) {
    const float c[3] = {1.0f, 1.0f, 1.0f};
    const gtr_closure g = gtr_make(c, prm);
    return gtr_dist(&g, nh, g.roughness);
}
float ora_kat_gtr_half_theta(const float prm[3], float u) {
    const float c[3] = {1.0f, 1.0f, 1.0f};
    const gtr_closure g = gtr_make(c, prm);
    return gtr_half_theta_sample(&g, u);
}
/* evaluate (first wavelength of the flat spectrum of a grey colour) and
 * sample_pdf of in -> out about normal n */
void ora_kat_gtr(float grey, const float prm[3], const float in[3], const float out[3], const float n[3],
                 float* eval, float* pdf) {
    const float c[3] = {grey, grey, grey};
    const gtr_closure g = gtr_make(c, prm);
    float e[NSPEC];
    gtr_eval(&g, va(in), va(out), va(n), 500.0f, e);
    *eval = e[0];
    *pdf = gtr_pdf(&g, va(in), va(out), va(n));
}
void ora_kat_gtr_sample(const float prm[3], const float in[3], const float n[3], float si, float sj, float out[3]) {
    const float c[3] = {1.0f, 1.0f, 1.0f};
    const gtr_closure g = gtr_make(c, prm);
    set3(out, gtr_sample_dir(&g, va(in), va(n), si, sj));
}
float ora_kat_gtr_estimate(const float prm[3], const float in[3], const float cent[3], float cos_theta,
                           const float nor[3]) {
    const float c[3] = {1.0f, 1.0f, 1.0f};
    const gtr_closure g = gtr_make(c, prm);
    return gtr_estimate(&g, va(in), va(cent), cos_theta, va(nor));
}
void ora_kat_axis_angle(const float axis[3], float angle, float m[16]) {
    const mat4 x = axis_angle(va(axis), angle);
    memcpy(m, &x.m[0][0], sizeof(float) * 16);
}

/* ---- adaptive sampling film ------------------------------------------------
 * Film::add_sample / variance_estimate (film/film.hpp:135-165) in the
 * adaptive loop of PathTraceIntegrator::render_blocks
 * (integrator/path_trace_integrator.cpp:296-376), over per-sample XYZ values
 * (width*height*spp*3, ordered (y, x, s)) that every pass of a bucket adds
 * again (each pass draws Sampler(x, y, s, seed) for s = 0 .. spp - 1).
 * Buckets are bucket_size squares in row-major order; film = sum / count,
 * var = variance_estimate, passes one int per bucket. */
static float film_hcol(float n) {
    if (n < 0.0f) n = 0.0f;
    return powf(n, 1.0f / 2.2f);
}
static float film_mmax(float a, float b) { return a > b ? a : b; }

void ora_film_adaptive(const float* samples, int W, int H, int spp, int spp_max, float var_max, int bs,
                       float* film, float* var, int32_t* passes) {
    const int nbx = (W + bs - 1) / bs, nby = (H + bs - 1) / bs;
    const size_t P = (size_t)W * H;
    float* sum = (float*)calloc(P * 3, sizeof(float));
    float* vp = (float*)calloc(P * 3, sizeof(float));
    float* vf = (float*)calloc(P * 3, sizeof(float));
    uint16_t* acc = (uint16_t*)calloc(P, sizeof(uint16_t));
    for (int by = 0; by < nby; ++by) {
        for (int bx = 0; bx < nbx; ++bx) {
            const int x0 = bx * bs, y0 = by * bs;
            const int bw = W - x0 < bs ? W - x0 : bs, bh = H - y0 < bs ? H - y0 : bs;
            float max_var = 9999999999.0f, samp_it = 0.0f;
            int np = 0;
            while (max_var > var_max && samp_it < (float)spp_max) {
                /* add_sample in the reference's order: x, then y, then s */
                for (int x = x0; x < x0 + bw; ++x) {
                    for (int y = y0; y < y0 + bh; ++y) {
                        const size_t pix = (size_t)y * W + x;
                        for (int s = 0; s < spp; ++s) {
                            const float* v = samples + (pix * spp + s) * 3;
                            int ok = 1;
                            for (int c = 0; c < 3; ++c)
                                if (isnan(v[c]) || !isfinite(v[c])) ok = 0;
                            if (!ok) continue;
                            for (int c = 0; c < 3; ++c) sum[pix * 3 + c] += v[c];
                            acc[pix]++;
                            const uint16_t k = acc[pix];
                            for (int c = 0; c < 3; ++c) {
                                const float avg = film_hcol(sum[pix * 3 + c] / (float)k);
                                if (k > 1) vf[pix * 3 + c] += fabsf(vp[pix * 3 + c] - avg) * (float)(k - 1);
                                vp[pix * 3 + c] = avg;
                            }
                        }
                    }
                }
                samp_it += (float)spp;
                ++np;
                max_var = 0.0f;
                for (int x = x0; x < x0 + bw; ++x)
                    for (int y = y0; y < y0 + bh; ++y) {
                        const size_t pix = (size_t)y * W + x;
                        const float ve = acc[pix] < 2 ? INFINITY
                                                      : (vf[pix * 3] / (float)(acc[pix] - 1)) / (float)sqrt((double)acc[pix]);
                        max_var = film_mmax(max_var, ve);
                    }
            }
            if (passes) passes[by * nbx + bx] = np;
        }
    }
    for (size_t pix = 0; pix < P; ++pix)
        for (int c = 0; c < 3; ++c) {
            if (film) film[pix * 3 + c] = acc[pix] ? sum[pix * 3 + c] / (float)acc[pix] : 0.0f;
            if (var)
                var[pix * 3 + c] = acc[pix] < 2 ? INFINITY
                                                : (vf[pix * 3 + c] / (float)(acc[pix] - 1)) / (float)sqrt((double)acc[pix]);
        }
    free(sum); free(vp); free(vf); free(acc);
}
