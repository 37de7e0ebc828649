// psy_render_dev.hpp -- per-path device functions of the wavefront path
// tracer; psy_kernels.hip wraps them in kernels.  Included after
// psy_trace.hpp, inside namespace psy.
//
// Restates, per path and per wavefront step:
//   init_path / next_ray_for_path / update_path
//                       integrator/path_trace_integrator.cpp:91-279
//   LightTree::sample, node_prob          accel/light_tree.cpp:6-96
//   SphereLight::sample / sample_pdf      lights/sphere_light.hpp:58-141
//   RectangleLight::sample / sample_pdf   lights/rectangle_light.hpp:40-93
//   uniform_sample_spherical_triangle, uniform_sample_sphere,
//   spherical_triangle_solid_angle, sphere_lambert
//                                         utils/monte_carlo.hpp:82-233
//   LambertClosure evaluate / sample_pdf / estimate_eval_over_solid_angle
//                                         shading/surface_closure.hpp:233-331
//   power_heuristic                       utils/mis.hpp:11-16
//   SpectralSample, Color_XYZ, XYZ_NORM_FAC
//                                         color/color.hpp:15-188
//   Color_to_SpectralSample, spectrum_xyz_to_p
//                                         color/color.hpp:490-511,
//                                         color/spectrum_grid.h:21-135
// Colours are upsampled through the spectrum grid the caller set on the
// context; without one, grey (r == g == b) colours take the grid's flat
// white-point spectrum, (X+Y+Z)/3 at every wavelength, and the host refuses
// scenes with other colours.
#pragma once

#include "psy_trace.hpp"

namespace psy {

constexpr float WL_MIN = 380.0f;
constexpr float WL_MAX = 700.0f;
constexpr int NSPEC = 4;  // hero wavelength + 3 (color/color.hpp:15)
constexpr double PSY_PI_D = 3.1415926535897932384626433;

// ---- path state planes ------------------------------------------------------
// One float plane per field, `cap` entries each; u32 fields are bit patterns.
enum : int {
    P_OFF = 0,       // sampler offset (hash of pixel and seed + sample index)
    P_DIM = 1,       // next sampler dimension
    P_PIX = 2,       // x | y << 16
    P_STEP = 3,      // step (bits 0-7), bit 8: stored hit is backfacing
    P_TIME = 4,
    P_WL = 5,        // hero wavelength
    P_LASTPDF = 6,
    P_COL = 7,       // 4 spectral samples each
    P_FCOL = 11,
    P_LCOL = 15,
    P_GP = 19,       // stored hit: world position, normal, origin offset
    P_GN = 22,
    P_POFF = 25,
    P_T = 28,
    P_SHADER = 29,   // u32 shader index of the stored hit
    P_PD = 30,       // the ray that found it: direction and differentials
    P_PODX = 33,
    P_PODY = 36,
    P_PDDX = 39,
    P_PDDY = 42,
    P_SIDX = 45,     // sample index within the pixel
    P_DPDU = 46,     // stored hit's world-space dp/du, dp/dv, dn/du, dn/dv (GTR shaders only)
    P_DPDV = 49,
    P_DNDU = 52,
    P_DNDV = 55,
    P_COUNT = 58
};
static_assert(P_COUNT == PATH_PLANES, "path state layout");
constexpr uint32_t STEP_BACKFACING = 1u << 8;

struct PathBuf {
    float* p;     // P_COUNT planes
    int64_t cap;  // entries per plane
    PSY_DEV float& f(int plane, int64_t i) const { return p[plane * cap + i]; }
    PSY_DEV uint32_t u(int plane, int64_t i) const { return f2u(p[plane * cap + i]); }
    PSY_DEV void setu(int plane, int64_t i, uint32_t v) const { p[plane * cap + i] = u2f(v); }
    PSY_DEV V3 v3(int plane, int64_t i) const { return mk(f(plane, i), f(plane + 1, i), f(plane + 2, i)); }
    PSY_DEV void set3(int plane, int64_t i, V3 v) const {
        f(plane, i) = v.x; f(plane + 1, i) = v.y; f(plane + 2, i) = v.z;
    }
};

// ---- spectral ---------------------------------------------------------------
// Color_to_SpectralSample for a grey colour: Color -> XYZ (E white), the
// flat spectrum of the white point scaled by X+Y+Z.
PSY_DEV float color_spec(float r, float g, float b) {
    const float x = ((r * 0.4339f) + (g * 0.3762f)) + (b * 0.1899f);
    const float y = ((r * 0.2126f) + (g * 0.7152f)) + (b * 0.0722f);
    const float z = ((r * 0.0177f) + (g * 0.1095f)) + (b * 0.8728f);
    const float norm = (float)(1.0 / (double)(x + y + z));
    if (!(norm < 3.402823466e+38f)) return 0.0f;
    return (1.0f / 3.0f) / norm;
}

PSY_DEV float wavelength_n(float hero, int n) {
    hero += (float)n * ((WL_MAX - WL_MIN) / NSPEC);
    if (hero > WL_MAX) hero -= (WL_MAX - WL_MIN);
    return hero;
}

// Color_to_SpectralSample (color/color.hpp:490-511) at the 4 wavelengths of
// `hero`.  With a grid, spectrum_xyz_to_p (color/spectrum_grid.h:21-135):
// the cell of the colour's chromaticity (in the grid's uv space), then at
// each wavelength the cell's data point spectra, linear between knots,
// weighted bilinearly in an inside cell or by the barycentric weights of
// the fan triangle around idx[0] that holds the point; times X+Y+Z and
// 1 / equal_energy_reflectance.  The cell and its weights do not depend on
// the wavelength and are found once.  Outside the grid (or a NaN
// chromaticity) the spectrum is 0.
PSY_DEV void color_spec4(const DevSpectrum& S, float r, float g, float b, float hero, float out[NSPEC]) {
    if (!S.cells) {
        const float v = color_spec(r, g, b);
        for (int k = 0; k < NSPEC; ++k) out[k] = v;
        return;
    }
    for (int k = 0; k < NSPEC; ++k) out[k] = 0.0f;
    const float x = ((r * 0.4339f) + (g * 0.3762f)) + (b * 0.1899f);
    const float y = ((r * 0.2126f) + (g * 0.7152f)) + (b * 0.0722f);
    const float z = ((r * 0.0177f) + (g * 0.1095f)) + (b * 0.8728f);
    const float norm = (float)(1.0 / (double)(x + y + z));
    if (!(norm < 3.402823466e+38f)) return;
    const float cx = x * norm, cy = y * norm;
    const float u = ((S.m[0] * cx) + (S.m[1] * cy)) + S.m[2];
    const float v = ((S.m[3] * cx) + (S.m[4] * cy)) + S.m[5];
    if (!(u >= 0.0f && u < (float)S.w && v >= 0.0f && v < (float)S.h)) return;
    const int ui = (int)u, vi = (int)v;
    const int32_t* cell = S.cells + 8 * (ui + S.w * vi);
    const int inside = cell[0], num = cell[1];
    const int stride = 4 + S.n_samples;
    // corner/fan weights: inside -> bilinear in (fu, fv); fan -> triangle
    // (idx[0], idx[t1], idx[t2]) with weights (bw, bv, bu)
    float fu = 0.0f, fv = 0.0f, bw = 0.0f, bv = 0.0f, bu = 0.0f;
    int t1 = -1, t2 = -1;
    if (inside) {
        fu = u - (float)ui;
        fv = v - (float)vi;
    } else if (num > 0) {
        const float* q0 = S.points + (int64_t)stride * cell[2];
        const float* q1 = S.points + (int64_t)stride * cell[3];
        const float ex = u - q0[2], ey = v - q0[3];
        float e0x = q1[2] - q0[2], e0y = q1[3] - q0[3];
        float uu = (e0x * ey) - (ex * e0y);
        for (int i = 0; i < num - 1; ++i) {
            const int j = i == num - 2 ? 1 : i + 2;
            const float* qn = S.points + (int64_t)stride * cell[2 + j];
            const float e1x = qn[2] - q0[2], e1y = qn[3] - q0[3];
            const float vv = (ex * e1y) - (e1x * ey);
            const float area = (e0x * e1y) - (e1x * e0y);
            const float cu = uu / area, cv = vv / area;
            const float cw = (1.0f - cu) - cv;
            if (cu < 0.0f || cv < 0.0f || cw < 0.0f) {
                uu = -vv;
                e0x = e1x;
                e0y = e1y;
                continue;
            }
            t1 = i + 1;
            t2 = j;
            bw = cw;
            bv = cv;
            bu = cu;
            break;
        }
        if (t1 < 0) return;
    } else {
        return;
    }
    for (int k = 0; k < NSPEC; ++k) {
        const float lam = wavelength_n(hero, k);
        const float sb = ((lam - S.smin) / (S.smax - S.smin)) * (float)(S.n_samples - 1);
        const int sb0 = (int)sb;
        const int sb1 = sb + 1.0f < (float)S.n_samples ? (int)(sb + 1.0f) : S.n_samples - 1;
        const float sbf = sb - (float)sb0;
        auto pv = [&](int c) {
            const float* sp = S.points + (int64_t)stride * cell[2 + c] + 4;
            return (sp[sb0] * (1.0f - sbf)) + (sp[sb1] * sbf);
        };
        float ip;
        if (inside) {
            ip = ((((pv(0) * (1.0f - fu)) * (1.0f - fv)) + ((pv(2) * (1.0f - fu)) * fv)) + ((pv(3) * fu) * fv)) +
                 ((pv(1) * fu) * (1.0f - fv));
        } else {
            ip = ((pv(0) * bw) + (pv(t1) * bv)) + (pv(t2) * bu);
        }
        out[k] = (ip / norm) * (1.0f / 0.009358239977091027f);
    }
}

// Wyman et al. CIE 1931 fits (color/color.hpp:150-165)
PSY_DEV float x_1931(float w) {
    const float t1 = (w - 442.0f) * ((w < 442.0f) ? 0.0624f : 0.0374f);
    const float t2 = (w - 599.8f) * ((w < 599.8f) ? 0.0264f : 0.0323f);
    const float t3 = (w - 501.1f) * ((w < 501.1f) ? 0.0490f : 0.0382f);
    return (0.362f * expf(-0.5f * t1 * t1)) + (1.056f * expf(-0.5f * t2 * t2)) - (0.065f * expf(-0.5f * t3 * t3));
}
PSY_DEV float y_1931(float w) {
    const float t1 = (w - 568.8f) * ((w < 568.8f) ? 0.0213f : 0.0247f);
    const float t2 = (w - 530.9f) * ((w < 530.9f) ? 0.0613f : 0.0322f);
    return (0.821f * expf(-0.5f * t1 * t1)) + (0.286f * expf(-0.5f * t2 * t2));
}
PSY_DEV float z_1931(float w) {
    const float t1 = (w - 437.0f) * ((w < 437.0f) ? 0.0845f : 0.0278f);
    const float t2 = (w - 459.0f) * ((w < 459.0f) ? 0.0385f : 0.0725f);
    return (1.217f * expf(-0.5f * t1 * t1)) + (0.681f * expf(-0.5f * t2 * t2));
}

// Color_XYZ(SpectralSample) * XYZ_NORM_FAC: a path's film contribution
PSY_DEV void path_film_xyz(float hero, const float col[NSPEC], float out[3]) {
    const float norm_fac = 0.009358239977091027f * (700.0f - 380.0f);
    float x = 0.0f, y = 0.0f, z = 0.0f;
    for (int i = 0; i < NSPEC; ++i) {
        const float w = wavelength_n(hero, i);
        x += x_1931(w) * col[i];
        y += y_1931(w) * col[i];
        z += z_1931(w) * col[i];
    }
    x *= 1.0f / NSPEC;
    y *= 1.0f / NSPEC;
    z *= 1.0f / NSPEC;
    out[0] = x * norm_fac;
    out[1] = y * norm_fac;
    out[2] = z * norm_fac;
}

// ---- sampling helpers -------------------------------------------------------
PSY_DEV double dclamp(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }

PSY_DEV V3 uniform_sample_sphere(float u, float v) {
    const float z = 1.f - (2.f * u);
    const float r = sqrtf(smax(0.f, 1.f - z * z));
    const float phi = (float)(2.f * PSY_PI_D * (double)v);
    return mk(r * cosf(phi), r * sinf(phi), z);
}

// Solid angle of a spherical triangle (unit vectors A, B, C).
PSY_DEV_FN float sph_tri_area(V3 A, V3 B, V3 C) {
    const double cos_a = dclamp((double)dot(B, C), -1.0, 1.0);
    const double cos_b = dclamp((double)dot(C, A), -1.0, 1.0);
    const double cos_c = dclamp((double)dot(A, B), -1.0, 1.0);
    const double sin_a = sqrt(1.0 - (cos_a * cos_a));
    const double sin_b = sqrt(1.0 - (cos_b * cos_b));
    const double sin_c = sqrt(1.0 - (cos_c * cos_c));
    if (cos_a == 1.0 || cos_b == 1.0 || cos_c == 1.0) return 0.0f;
    const double cos_A = dclamp((cos_a - (cos_b * cos_c)) / (sin_b * sin_c), -1.0, 1.0);
    const double cos_B = dclamp((cos_b - (cos_c * cos_a)) / (sin_c * sin_a), -1.0, 1.0);
    const double cos_C = dclamp((cos_c - (cos_a * cos_b)) / (sin_a * sin_b), -1.0, 1.0);
    return (float)(acos(cos_A) + acos(cos_B) + acos(cos_C) - PSY_PI_D);
}

PSY_DEV_FN V3 uniform_sample_spherical_triangle(V3 A, V3 B, V3 C, float i, float j) {
    const double cos_a = dclamp((double)dot(B, C), -1.0, 1.0);
    const double cos_b = dclamp((double)dot(C, A), -1.0, 1.0);
    const double cos_c = dclamp((double)dot(A, B), -1.0, 1.0);
    const double sin_a = sqrt(1.0 - (cos_a * cos_a));
    const double sin_b = sqrt(1.0 - (cos_b * cos_b));
    const double sin_c = sqrt(1.0 - (cos_c * cos_c));
    if (cos_a == 1.0 || cos_b == 1.0 || cos_c == 1.0) return A;
    const double cos_A = dclamp((cos_a - (cos_b * cos_c)) / (sin_b * sin_c), -1.0, 1.0);
    const double cos_B = dclamp((cos_b - (cos_c * cos_a)) / (sin_c * sin_a), -1.0, 1.0);
    const double cos_C = dclamp((cos_c - (cos_a * cos_b)) / (sin_a * sin_b), -1.0, 1.0);
    const double sin_A = sqrt(1.0 - (cos_A * cos_A));
    const double ang_A = acos(cos_A), ang_B = acos(cos_B), ang_C = acos(cos_C);
    const double area = ang_A + ang_B + ang_C - PSY_PI_D;
    const double area_2 = area * i;
    const double s = sin(area_2 - ang_A);
    const double t = cos(area_2 - ang_A);
    const double u = t - cos_A;
    const double v = s + (sin_A * cos_c);
    const double q_top = (((v * t) - (u * s)) * cos_A) - v;
    const double q_bottom = ((v * s) + (u * t)) * sin_A;
    const double q = q_top / q_bottom;
    // Vec3<float> * double: the factor is rounded to float first
    const V3 C_2 = add(scale(A, (float)q), scale(normalized(sub(C, scale(A, dot(C, A)))), (float)sqrt(1.0 - (q * q))));
    const double z = 1.0 - (j * (1.0 - (double)dot(C_2, B)));
    return add(scale(B, (float)z), scale(normalized(sub(C_2, scale(B, dot(C_2, B)))), (float)sqrt(1.0 - (z * z))));
}

PSY_DEV_FN float sphere_lambert(float nlcos, float rcos) {
    const double HPI = PSY_PI_D / 2;
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
    else if (nl < (float)(HPI + (double)r)) return (float)((double)(g + h) * (1.0 / PSY_PI_D));
    return 0.0f;
}

// LambertClosure::estimate_eval_over_solid_angle
PSY_DEV_FN float lambert_estimate(V3 in, V3 out, float cos_theta, V3 nor) {
    if (cos_theta < 0.0f) return 1.0f;
    V3 nn = normalized(nor);
    const V3 v = normalized(out);
    if (dot(nn, in) > 0.0f) nn = scale(nn, -1.0f);
    const float cos_nv = dot(nn, v);
    return sphere_lambert(cos_nv, cos_theta);
}

// LambertClosure::evaluate: the albedo * max(0, n.v)/pi, upsampled at the
// 4 wavelengths of `hero`
PSY_DEV void lambert_eval(const DevSpectrum& S, const float albedo[3], V3 in, V3 out, V3 gn, float hero,
                          float e[NSPEC]) {
    V3 nn = normalized(gn);
    const V3 v = normalized(out);
    if (dot(nn, in) > 0.0f) nn = scale(nn, -1.0f);
    const float fac = smax(dot(nn, v), 0.0f) * (float)(1.0 / PSY_PI_D);
    color_spec4(S, albedo[0] * fac, albedo[1] * fac, albedo[2] * fac, hero, e);
}
// The same for the grey albedo 1 on the flat spectrum (the adjoint's
// parameter-free factor)
PSY_DEV float lambert_eval1(V3 in, V3 out, V3 gn) {
    V3 nn = normalized(gn);
    const V3 v = normalized(out);
    if (dot(nn, in) > 0.0f) nn = scale(nn, -1.0f);
    const float fac = smax(dot(nn, v), 0.0f) * (float)(1.0 / PSY_PI_D);
    return color_spec(fac, fac, fac);
}

// LambertClosure::sample_pdf
PSY_DEV float lambert_pdf(V3 in, V3 out, V3 gn) {
    V3 nn = normalized(gn);
    const V3 v = normalized(out);
    if (dot(nn, normalized(in)) > 0.0f) nn = scale(nn, -1.0f);
    return smax(dot(nn, v) * (float)(1.0 / PSY_PI_D), 0.0f);
}

// ---- GTR microfacet closure (shading/surface_closure.hpp:339-608) -----------
// The shader's roughness, tail shape and normalization come validated from
// the host (GTRClosure::validate).  A roughness of 0 is a perfect mirror:
// a delta closure, whose paths take no light samples.

// clamp(v, a, b) of utils/utils.hpp:42-44: std::min(b, std::max(a, v))
PSY_DEV float uclamp(float v, float a, float b) { return smin(b, smax(a, v)); }

// GTRClosure::normalization (:376-382), its pi product in double
PSY_DEV_FN float gtr_normalization(float r, float t) {
    const float r2 = r * r;
    const float top = (t - 1.0f) * (r2 - 1);
    const double bottom = PSY_PI_D * (double)(1.0f - powf(r2, 1.0f - t));
    return (float)((double)top / bottom);
}

// GTRClosure::dist (:398-410) for roughness `rough`, whose normalization
// is `norm`
PSY_DEV_FN float gtr_dist(float nh, float rough, float tail, float norm) {
    const float r2 = rough * rough;
    float D = 0.0f;
    if (nh > 0.0f) {
        const float nh2 = nh * nh;
        D = norm / powf(1 + ((r2 - 1) * nh2), tail);
    }
    return D;
}

// schlick_fresnel_from_fac (:173-177)
PSY_DEV float schlick_from_fac(float f, float c) {
    const float c1 = (1.0f - c);
    const float c2 = c1 * c1;
    return f + ((1.0f - f) * c1 * c2 * c2);
}

// GTRClosure::evaluate (:452-517) at the 4 wavelengths of `hero`
PSY_DEV_FN void gtr_eval(const DevSpectrum& S, const DevShader& g, V3 in, V3 out, V3 gn, float hero,
                         float e[NSPEC]) {
    V3 nn = normalized(gn);
    const V3 aa = scale(normalized(in), -1.0f);
    const V3 bb = normalized(out);
    const V3 hh = normalized(add(aa, bb));
    if (dot(nn, hh) < 0.0f) nn = scale(nn, -1.0f);
    const float na = uclamp(dot(nn, aa), -1.0f, 1.0f);
    const float nb = uclamp(dot(nn, bb), -1.0f, 1.0f);
    const float ha = uclamp(dot(hh, aa), -1.0f, 1.0f);
    const float hb = uclamp(dot(hh, bb), -1.0f, 1.0f);
    const float nh = uclamp(dot(nn, hh), -1.0f, 1.0f);
    const float r2 = g.roughness * g.roughness;
    color_spec4(S, g.color[0], g.color[1], g.color[2], hero, e);
    for (int k = 0; k < NSPEC; ++k) e[k] = lerpf(1.0f - g.fresnel, schlick_from_fac(e[k], hb), e[k]);
    if (g.roughness == 0.0f) return;  // a mirror: the Fresnel-weighted colour
    float D = 1.0f;
    if (nh > 0.0f) {
        const float nh2 = nh * nh;
        D = g.norm / powf(1 + ((r2 - 1) * nh2), g.tail_shape);
    }
    const float na2 = na * na;
    const float tan_na = sqrtf((1.0f - na2) / na2);
    const float g1_pos = (ha * na) > 0.0f ? 1.0f : 0.0f;
    const float g1_a = r2 * tan_na;
    const float g1_b = (sqrtf(1.0f + g1_a * g1_a) - 1.0f) * 0.5f;
    const float G1 = g1_pos / (1.0f + g1_b);
    const float nb2 = nb * nb;
    const float tan_nb = sqrtf((1.0f - nb2) / nb2);
    const float g2_pos = (hb * nb) > 0.0f ? 1.0f : 0.0f;
    const float g2_a = r2 * tan_nb;
    const float g2_b = (sqrtf(1.0f + g2_a * g2_a) - 1.0f) * 0.5f;
    const float G2 = g2_pos / (1.0f + g2_b);
    const float f = D * G1 * G2;
    for (int k = 0; k < NSPEC; ++k) e[k] = (e[k] * f) * (float)(1.0 / PSY_PI_D);
}

// GTRClosure::sample_pdf (:556-571)
PSY_DEV_FN float gtr_pdf(const DevShader& g, V3 in, V3 out, V3 gn) {
    V3 nn = normalized(gn);
    const V3 aa = scale(normalized(in), -1.0f);
    const V3 bb = normalized(out);
    const V3 hh = normalized(scale(add(aa, bb), 0.5f));
    if (dot(nn, hh) < 0.0f) nn = scale(nn, -1.0f);
    const float nh = uclamp(dot(nn, hh), -1.0f, 1.0f);
    return gtr_dist(nh, g.roughness, g.tail_shape, g.norm) * (float)(1.0 / PSY_PI_D);
}

// GTRClosure::sample (:418-437): a half vector from the distribution
// (half_theta_sample, :385-396), the incoming direction mirrored about it
PSY_DEV_FN V3 gtr_sample_dir(const DevShader& g, V3 in, V3 gn, float si, float sj) {
    V3 nn = normalized(gn);
    if (dot(nn, in) > 0.0f) nn = scale(nn, -1.0f);
    const float r2 = g.roughness * g.roughness;
    const float top =
        1.0f - powf((powf(r2, 1.0f - g.tail_shape) * (1.0f - si)) + si, 1.0f / (1.0f - g.tail_shape));
    const float bottom = (1.0f - r2);
    const float theta_cos = sqrtf(top / bottom);
    const float theta_sin = sqrtf(1.0f - (theta_cos * theta_cos));
    const float angle = (float)((double)sj * PSY_PI_D * 2);
    V3 half_dir = mk(cosf(angle) * theta_sin, sinf(angle) * theta_sin, theta_cos);
    half_dir = normalized(zup_to_vec(half_dir, nn));
    return sub(in, scale(scale(half_dir, 2.0f), dot(in, half_dir)));
}

// GTRClosure::estimate_eval_over_solid_angle (:520-553): the distribution
// at the half vector towards the centre, widened by the cone's angle
PSY_DEV_FN float gtr_estimate(float rough, float tail, V3 in, V3 cent, float cos_theta, V3 nor) {
    V3 nn = normalized(nor);
    const V3 aa = scale(normalized(in), -1.0f);
    const V3 bb = normalized(cent);
    if (dot(nn, in) > 0.0f) nn = scale(nn, -1.0f);
    const float theta = acosf(cos_theta);
    const V3 hh = normalized(add(aa, bb));
    const float nh = uclamp(dot(nn, hh), -1.0f, 1.0f);
    const float r = smin(1.0f, sqrtf(rough) + (2.0f * theta / (float)PSY_PI_D));
    const float fac = gtr_dist(nh, r, tail, gtr_normalization(r, tail));
    return fac * smin(1.0f - cos_theta, 1.0f) * (float)(1.0 / PSY_PI_D);
}

// calc_uv_differentials (basics/differential_geometry.hpp:145-156)
PSY_DEV void uv_differentials(V3 dp, V3 dpdu, V3 dpdv, float* du, float* dv) {
    const float dpdu_ilen = 1.0f / length(dpdu);
    const V3 dpdu_n = scale(dpdu, dpdu_ilen);
    const float dpdv_ilen = 1.0f / length(dpdv);
    const V3 dpdv_n = scale(dpdv, dpdv_ilen);
    *du = dot(dp, dpdu_n) * dpdu_ilen;
    *dv = dot(dp, dpdv_n) * dpdv_ilen;
}

// The rotation by `angle` about `axis` as a row-vector matrix (v' = v M),
// Rodrigues' formula in the element order of Imath's Matrix44::setAxisAngle,
// which make_axis_angle_transform (math/transform.hpp:221-226) calls; the
// axis is normalized once more there.
PSY_DEV M44 axis_angle_matrix(V3 axis, float angle) {
    const V3 u = normalized(axis);
    const float sn = sinf(angle), cs = cosf(angle);
    const float oc = 1 - cs;
    M44 m = m_identity();
    m.m[0] = u.x * u.x * oc + cs;
    m.m[1] = u.x * u.y * oc + u.z * sn;
    m.m[2] = u.x * u.z * oc - u.y * sn;
    m.m[4] = u.x * u.y * oc - u.z * sn;
    m.m[5] = u.y * u.y * oc + cs;
    m.m[6] = u.y * u.z * oc + u.x * sn;
    m.m[8] = u.x * u.z * oc + u.y * sn;
    m.m[9] = u.y * u.z * oc - u.x * sn;
    m.m[10] = u.z * u.z * oc + cs;
    return m;
}

// reflect_ray_direction_differential (basics/differential_geometry.hpp:113-117)
PSY_DEV V3 reflect_dd(V3 normal, V3 normal_d, V3 d, V3 dd) {
    const float ddn = dot(dd, normal) + dot(d, normal_d);
    const V3 tmp = add(scale(normal_d, dot(d, normal)), scale(normal, ddn));
    return sub(dd, scale(tmp, 2.0f));
}

// GTRClosure::propagate_differentials (:574-606) with clamp_dd
// (basics/differential_geometry.hpp:125-135); geo is the hit's world-space
// normal and derivatives.
PSY_DEV_FN void gtr_differentials(float t, V3 in_d, V3 in_odx, V3 in_ody, V3 in_ddx, V3 in_ddy, V3 gn, V3 dpdu,
                                  V3 dpdv, V3 dndu, V3 dndv, V3 out_d, V3* odx, V3* ody, V3* ddx, V3* ddy) {
    V3 nn = normalized(gn);
    const V3 hh = normalized(add(normalized(scale(in_d, -1.0f)), normalized(out_d)));
    *odx = transfer_origin_diff(t, nn, in_d, in_odx, in_ddx);
    *ody = transfer_origin_diff(t, nn, in_d, in_ody, in_ddy);
    float dux, dvx, duy, dvy;
    uv_differentials(*odx, dpdu, dpdv, &dux, &dvx);
    uv_differentials(*ody, dpdu, dpdv, &duy, &dvy);
    V3 dndx = add(scale(dndu, dux), scale(dndv, dvx));
    V3 dndy = add(scale(dndu, duy), scale(dndv, dvy));
    if (dot(nn, hh) < 0.0f) {
        nn = scale(nn, -1.0f);
        dndx = scale(dndx, -1.0f);
        dndy = scale(dndy, -1.0f);
    }
    const V3 axis = normalized(cross(nn, hh));
    const float angle = acosf(uclamp(dot(nn, hh), 0.0f, 1.0f));
    const M44 xf = axis_angle_matrix(axis, angle);
    const V3 dhdx = m_dir(xf, dndx);
    const V3 dhdy = m_dir(xf, dndy);
    *ddx = reflect_dd(hh, dhdx, in_d, in_ddx);
    *ddy = reflect_dd(hh, dhdy, in_d, in_ddy);
    const float len_d = length(out_d);
    const float len_dx = length(*ddx);
    const float len_dy = length(*ddy);
    if ((len_dx / len_d) > 0.9f) *ddx = scale(*ddx, 0.9f * len_d / len_dx);
    if ((len_dy / len_d) > 0.9f) *ddy = scale(*ddy, 0.9f * len_d / len_dy);
}

// ---- lights -----------------------------------------------------------------
PSY_DEV void rect_light_at(const float* d, int ts, float time, float* w, float* h, float col[3]) {
    float na = 0.0f;
    const int i = lerp_seq_index(time, ts, &na);
    if (i == -1 || i == -2) {
        const float* e = d + (i == -1 ? 0 : (ts - 1) * 5);
        *w = e[0]; *h = e[1];
        col[0] = e[2]; col[1] = e[3]; col[2] = e[4];
        return;
    }
    const float f = lerp_fac(na);
    const float* a = d + i * 5;
    const float* b = d + (i + 1) * 5;
    *w = (a[0] * f) + (b[0] * na);
    *h = (a[1] * f) + (b[1] * na);
    for (int k = 0; k < 3; ++k) col[k] = (a[2 + k] * f) + (b[2 + k] * na);
}

PSY_DEV void sphere_light_at(const float* d, int ts, float time, V3* pos, float* rad, float col[3]) {
    float na = 0.0f;
    const int i = lerp_seq_index(time, ts, &na);
    if (i == -1 || i == -2) {
        const float* e = d + (i == -1 ? 0 : (ts - 1) * 7);
        *pos = mk(e[0], e[1], e[2]); *rad = e[3];
        col[0] = e[4]; col[1] = e[5]; col[2] = e[6];
        return;
    }
    const float* a = d + i * 7;
    const float* b = d + (i + 1) * 7;
    *pos = lerpv(na, mk(a[0], a[1], a[2]), mk(b[0], b[1], b[2]));
    *rad = lerpf(na, a[3], b[3]);
    for (int k = 0; k < 3; ++k) col[k] = (a[4 + k] * lerp_fac(na)) + (b[4 + k] * na);
}

// RectangleLight::sample_pdf
PSY_DEV_FN float rect_light_pdf(float w, float h, V3 arr) {
    const V3 p1 = normalized(sub(mk(w * 0.5f, h * 0.5f, 0.0f), arr));
    const V3 p2 = normalized(sub(mk(w * -0.5f, h * 0.5f, 0.0f), arr));
    const V3 p3 = normalized(sub(mk(w * -0.5f, h * -0.5f, 0.0f), arr));
    const V3 p4 = normalized(sub(mk(w * 0.5f, h * -0.5f, 0.0f), arr));
    return 1.0f / (sph_tri_area(p2, p1, p3) + sph_tri_area(p4, p1, p3));
}

// SphereLight::sample_pdf
PSY_DEV_FN float sphere_light_pdf(V3 pos, double radius, V3 arr) {
    const double d2 = (double)length2(sub(pos, arr));
    const double d = sqrt(d2);
    if (d > radius) {
        const double s2 = (radius * radius) / d2 < 1.0 ? (radius * radius) / d2 : 1.0;
        const double cos_theta_max = sqrt(1.0 - s2);
        return (float)(1.0 / (2.0 * PSY_PI_D * (1.0 - cos_theta_max)));
    }
    return (float)(1.0f / (4.0f * PSY_PI_D));
}

struct LQuery {
    float n, u, v;
    V3 pos, nor, d;
    float time;
    float wl;         // hero wavelength
    M44 xform;
    V3 to_light;
    float spec[NSPEC];       // the light's emission at the sample, times its emission scale
    float spec_base[NSPEC];  // the emission without the scale
    float scale;      // that emission scale
    int32_t obj;      // the sampled light object
    float selection_pdf, light_sample_pdf;
    // the closure whose estimate weights the tree's nodes: Lambert, or GTR
    // with this validated roughness and tail shape
    int32_t bsdf;
    float g_rough, g_tail;
};

PSY_DEV_FN void sphere_light_sample(const DevSpectrum& SP, const float* od, int ts, LQuery& q, float* pdf) {
    V3 pos; float radf; float col[3];
    sphere_light_at(od, ts, q.time, &pos, &radf, col);
    const V3 arr = q.pos;
    const double radius = radf;
    const double surface_area_inv = 1.0 / (4.0 * PSY_PI_D * radius * radius);
    V3 x, y, z = sub(pos, arr);
    const double d2 = (double)length2(z);
    const double d = sqrt(d2);
    coordinate_system_from_vec3(z, &x, &y);
    x = normalized(x); y = normalized(y); z = normalized(z);
    color_spec4(SP, col[0] * (float)surface_area_inv, col[1] * (float)surface_area_inv,
                col[2] * (float)surface_area_inv, q.wl, q.spec);
    if (d > radius) {
        const double s2 = (radius * radius) / d2 < 1.0 ? (radius * radius) / d2 : 1.0;
        const double c2 = 1.0 - s2;
        const double sin_theta_max = sqrt(s2);
        const double cos_theta_max = sqrt(c2);
        const double du = (double)q.u, dv = (double)q.v;
        // uniform_sample_cone
        const double ct = (1.0 - du) + (du * cos_theta_max);
        const double st = sqrt(1.0 - (ct * ct));
        const double phi0 = dv * 2.0 * PSY_PI_D;
        V3 sample = normalized(mk((float)(cos(phi0) * st), (float)(sin(phi0) * st), (float)ct));
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
        const double phi = dv * 2.0 * PSY_PI_D;
        sample.x = (float)(cos(phi) * sin_a * radius);
        sample.y = (float)(sin(phi) * sin_a * radius);
        sample.z = (float)(d - (cos_a * radius));
        q.to_light = add(add(scale(x, sample.x), scale(y, sample.y)), scale(z, sample.z));
        *pdf = (float)(1.0 / (2.0 * PSY_PI_D * (1.0 - cos_theta_max)));
    } else {
        q.to_light = uniform_sample_sphere(q.u, q.v);
        *pdf = (float)(1.0f / (4.0f * PSY_PI_D));
    }
}

PSY_DEV_FN void rect_light_sample(const DevSpectrum& SP, const float* od, int ts, LQuery& q, float* pdf) {
    float w, h, col[3];
    rect_light_at(od, ts, q.time, &w, &h, col);
    const V3 arr = q.pos;
    const double inv_area = 1.0 / (double)(w * h);
    const V3 p1 = normalized(sub(mk(w * 0.5f, h * 0.5f, 0.0f), arr));
    const V3 p2 = normalized(sub(mk(w * -0.5f, h * 0.5f, 0.0f), arr));
    const V3 p3 = normalized(sub(mk(w * -0.5f, h * -0.5f, 0.0f), arr));
    const V3 p4 = normalized(sub(mk(w * 0.5f, h * -0.5f, 0.0f), arr));
    const float area_1 = sph_tri_area(p2, p1, p3);
    const float area_2 = sph_tri_area(p4, p1, p3);
    const float prob_1 = area_1 / (area_1 + area_2);
    const float prob_2 = 1.0f - prob_1;
    V3 sv;
    if (q.u < prob_1) sv = uniform_sample_spherical_triangle(p2, p1, p3, q.v, q.u / prob_1);
    else sv = uniform_sample_spherical_triangle(p4, p1, p3, q.v, 1.0f - ((q.u - prob_1) / prob_2));
    sv = scale(sv, -arr.z / sv.z);
    q.to_light = sv;
    *pdf = 1.0f / (area_1 + area_2);
    color_spec4(SP, (col[0] * (float)inv_area) * 0.5f, (col[1] * (float)inv_area) * 0.5f,
                (col[2] * (float)inv_area) * 0.5f, q.wl, q.spec);
}

PSY_DEV void lbox(const DevLights& L, int b, V3* lo, V3* hi) {
    const float4 a = L.bounds[2 * b], c = L.bounds[2 * b + 1];
    *lo = mk(a.x, a.y, a.z);
    *hi = mk(c.x, c.y, c.z);
}

PSY_DEV_FN float node_prob(const DevLights& L, const LQuery& q, int index) {
    const DevLNode nd = L.nodes[index];
    V3 lo, hi;
    float na = 0.0f;
    const int i = lerp_seq_index(q.time, nd.n_bounds, &na);
    if (i == -1) lbox(L, nd.bounds, &lo, &hi);
    else if (i == -2) lbox(L, nd.bounds + nd.n_bounds - 1, &lo, &hi);
    else {
        V3 alo, ahi, blo, bhi;
        lbox(L, nd.bounds + i, &alo, &ahi);
        lbox(L, nd.bounds + i + 1, &blo, &bhi);
        lo = lerpv(na, alo, blo);
        hi = lerpv(na, ahi, bhi);
    }
    const V3 c = scale(add(lo, hi), 0.5f);
    const V3 d = sub(c, q.pos);
    const float dist2 = length2(d);
    const float r = length(sub(hi, lo)) * 0.5f;
    const float r2 = r * r;
    const float inv_surface_area = 1.0f / r2;
    float cos_theta_max;
    if (dist2 <= r2) {
        cos_theta_max = -1.0f;
    } else {
        const float s2 = smin(1.0f, r2 / dist2);
        cos_theta_max = sqrtf(1.0f - s2);
    }
    const float est = q.bsdf == PSY_SHADER_GTR ? gtr_estimate(q.g_rough, q.g_tail, q.d, d, cos_theta_max, q.nor)
                                               : lambert_estimate(q.d, d, cos_theta_max, q.nor);
    const float approx = smax(0.0f, est);
    return nd.energy * inv_surface_area * approx;
}

// LightTree::sample, descending through sub-assembly trees.  False when the
// scene has no light.
PSY_DEV_FN bool light_tree_sample(const DevScene& s, const DevLights& L, const DevSpectrum& SP, LQuery& q) {
    int a = 0;
    for (int guard = 0; guard < 64; ++guard) {
        const DevAsm A = L.asms[a];
        if (A.n_nodes == 0) return false;
        int ni = A.node0;
        float tot_prob = 1.0f;
        for (;;) {
            const DevLNode nd = L.nodes[ni];
            if (nd.leaf) break;
            float p1 = node_prob(L, q, A.node0 + nd.index1);
            float p2 = node_prob(L, q, A.node0 + nd.index2);
            const float total = p1 + p2;
            if (total <= 0.0f) {
                p1 = 0.5f;
                p2 = 0.5f;
            } else {
                p1 /= total;
                p2 /= total;
            }
            if (q.n <= p1) {
                tot_prob *= p1;
                ni = A.node0 + nd.index1;
                q.n /= p1;
            } else {
                tot_prob *= p2;
                ni = A.node0 + nd.index2;
                q.n = (q.n - p1) / p2;
            }
        }
        const DevAsmInst in = L.inst[A.inst0 + L.nodes[ni].instance];
        if (in.xf_count > 0) {
            const M44 xf = m_lerp_seq(q.time, L.xf + A.xf0 + in.xf_index, in.xf_count);
            q.pos = m_pos(xf, q.pos);
            const M44 inv = m_inverse(xf);
            q.nor = normalized(m_dir_transpose(inv, q.nor));
            q.xform = m_mul(q.xform, xf);
        }
        if (in.type == PSY_INSTANCE_OBJECT) {
            const DevObject ob = s.obj[A.obj0 + in.data_index];
            const float* od = s.odata + ob.data;
            float p = 1.0f;
            if (ob.kind == K_SPHERE_LIGHT) sphere_light_sample(SP, od, ob.ts, q, &p);
            else if (ob.kind == K_RECT_LIGHT) rect_light_sample(SP, od, ob.ts, q, &p);
            else return false;
            q.to_light = m_dir(m_inverse(q.xform), q.to_light);
            q.obj = A.obj0 + in.data_index;
            q.scale = s.escale[q.obj];
            for (int k = 0; k < NSPEC; ++k) {
                q.spec_base[k] = q.spec[k];
                q.spec[k] *= q.scale;
            }
            q.selection_pdf *= (tot_prob * (float)A.total_lights);
            q.light_sample_pdf = p;
            return true;
        }
        const int sub_a = L.sub[A.sub0 + in.data_index];
        q.selection_pdf *= (tot_prob * (float)A.total_lights) / (float)L.asms[sub_a].total_lights;
        a = sub_a;
    }
    return false;
}

// ---- path kernels -------------------------------------------------------------

PSY_DEV float next_dim(const RenderLaunch& A, uint32_t off, uint32_t& dim) {
    if (A.sobol_m) return sobol_sampler_get(A.sobol_m, A.sobol_cols, off, dim++);
    return sampler_get(A.dims, A.perm, off, dim++);
}

// init_path + the camera ray: path slot i of the chunk, sample `first + i`
// of the image in (y, x, s) order.
// A world ray in registers (WorldRay, basics/ray.hpp:162-185).
struct WRay {
    V3 o, d, odx, ody, ddx, ddy;
    float time;
    uint32_t type;
};

// The closest-hit part of an Intersection, in registers.
struct HitRec {
    float t, u, v, aux;
    uint32_t inst;
};

PSY_DEV void load_wray(const float* w, int64_t rcap, int64_t j, WRay& r) {
    r.o = mk(w[0 * rcap + j], w[1 * rcap + j], w[2 * rcap + j]);
    r.d = mk(w[3 * rcap + j], w[4 * rcap + j], w[5 * rcap + j]);
    r.odx = mk(w[6 * rcap + j], w[7 * rcap + j], w[8 * rcap + j]);
    r.ody = mk(w[9 * rcap + j], w[10 * rcap + j], w[11 * rcap + j]);
    r.ddx = mk(w[12 * rcap + j], w[13 * rcap + j], w[14 * rcap + j]);
    r.ddy = mk(w[15 * rcap + j], w[16 * rcap + j], w[17 * rcap + j]);
    r.time = w[18 * rcap + j];
    r.type = f2u(w[19 * rcap + j]);
}

PSY_DEV void store_wray(float* w, int64_t rcap, int64_t j, const WRay& r) {
    store_wray(w, rcap, j, r.o, r.d, r.odx, r.ody, r.ddx, r.ddy, r.time, r.type);
}

// init_path + the camera ray (next_ray_for_path step 0) of path slot i of
// the chunk, sample `first + i` of the image in (y, x, s) order.
PSY_DEV void render_init_path(const RenderLaunch& A, int64_t gid, int64_t i, const PathBuf& P, WRay& ray) {
    const int s = (int)(gid % A.spp);
    const int64_t pix = gid / A.spp;
    const int px = (int)(pix % A.width);
    const int py = (int)(pix / A.width);
    const uint32_t off = sampler_offset((uint32_t)px, (uint32_t)py, (uint32_t)s, A.seed);
    uint32_t dim = 0;
    const float wl = (next_dim(A, off, dim) * (WL_MAX - WL_MIN)) + WL_MIN;
    const float sx = next_dim(A, off, dim);
    const float sy = next_dim(A, off, dim);
    const float sa = next_dim(A, off, dim);
    const float sb = next_dim(A, off, dim);
    const float time = next_dim(A, off, dim);
    V3 z;
    camera_ray(A.cam, A.width, A.height, px, py, sx, sy, sa, sb, time, &ray.o, &ray.d, &ray.ddx, &ray.ddy, &z);
    ray.odx = z;
    ray.ody = z;
    ray.time = time;
    ray.type = 1u;  // WorldRay::CAMERA
    P.setu(P_OFF, i, off);
    P.setu(P_DIM, i, dim);
    P.setu(P_PIX, i, (uint32_t)px | ((uint32_t)py << 16));
    P.setu(P_SIDX, i, (uint32_t)s);
    P.setu(P_STEP, i, 0u);
    P.f(P_TIME, i) = time;
    P.f(P_WL, i) = wl;
    P.f(P_LASTPDF, i) = 0.0f;
    for (int k = 0; k < NSPEC; ++k) {
        P.f(P_COL + k, i) = 0.0f;
        P.f(P_FCOL + k, i) = 1.0f;
        P.f(P_LCOL + k, i) = 0.0f;
    }
}

PSY_DEV void render_init_one(const RenderLaunch& A, int64_t first, int64_t i, const PathBuf& P, float* rays,
                             int64_t rcap, uint32_t* act) {
    WRay r;
    render_init_path(A, first + i, i, P, r);
    store_wray(rays, rcap, i, r);
    act[i] = (uint32_t)i;
}

// Emission and light pdf of a camera / bounce ray that ended on a light
// object (SphereLight / RectangleLight::intersect_ray closures).
// *spec is the emission times the object's emission scale, *base without it.
PSY_DEV_FN void light_hit(const DevScene& s, const DevSpectrum& SP, uint32_t li, V3 wo, float time, float wl,
                          float spec[NSPEC], float base[NSPEC], float* light_pdf) {
    const DevInstance I = s.inst[li];
    const DevObject ob = s.obj[I.object];
    V3 o = wo;
    if (I.xf_count > 0) o = m_pos(m_lerp_seq(time, s.xf + I.xf_offset, I.xf_count), wo);
    const float* od = s.odata + ob.data;
    if (ob.kind == K_RECT_LIGHT) {
        float w, h, col[3];
        rect_light_at(od, ob.ts, time, &w, &h, col);
        *light_pdf = rect_light_pdf(w, h, o);
        const double area = (double)(w * h);
        color_spec4(SP, (col[0] * 0.5f) / (float)area, (col[1] * 0.5f) / (float)area,
                    (col[2] * 0.5f) / (float)area, wl, base);
    } else {
        V3 c; float r, col[3];
        sphere_light_at(od, ob.ts, time, &c, &r, col);
        const double area = 4.0 * PSY_PI_D * r * r;
        *light_pdf = sphere_light_pdf(c, (double)r, o);
        color_spec4(SP, col[0] / (float)area, col[1] / (float)area, col[2] / (float)area, wl, base);
    }
    for (int k = 0; k < NSPEC; ++k) spec[k] = base[k] * s.escale[I.object];
}


// ---- Intersection records ------------------------------------------------------
// The fields of the reference's Intersection (basics/intersection.hpp:22-49)
// for one closest hit, as its tracer fills them: differential geometry in the
// hit object's space (Sphere::intersect_ray, object/sphere.cpp:120-189;
// intersect_rays_with_patch, object/patch_utils.hpp:94-131; the lights'
// intersect_ray), the merged instance transform at the ray's time (`space`,
// world -> object), the light pdf of a light hit, and the closure the
// shader binds (shading/surface_shader.hpp:28-53; magenta emission without a
// shader, tracer/tracer.cpp:248-250) as a kind and an RGB colour.
struct IsectRec {
    V3 p, n, dpdu, dpdv, dndu, dndv, offset;
    M44 space;
    float light_pdf;
    uint32_t backfacing;
    uint32_t closure;  // PSY_SHADER_LAMBERT / PSY_SHADER_EMIT
    float col[3];
    int32_t shader;    // the instance's shader, -1 for lights and none
};

// Sphere differential geometry (object/sphere.cpp:131-189): latitude /
// longitude parameterization scaled to [0,1]^2.
PSY_DEV_FN void sphere_dg(V3 n, float radi, V3* dpdu, V3* dpdv, V3* dndu, V3* dndv) {
    const V3 pp = scale(n, radi);
    const float lat_sin = sqrtf((n.x * n.x) + (n.y * n.y));
    const float long_cos = n.x / lat_sin;
    const float long_sin = n.y / lat_sin;
    const float pi2 = (float)(PSY_PI_D * 2);
    const float fpi = (float)PSY_PI_D;
    *dpdu = scale(mk(pp.y * -1.0f, pp.x, 0.0f), pi2);
    *dpdv = scale(mk(pp.z * long_cos, pp.z * long_sin, -radi * lat_sin), fpi);
    const V3 d2pduu = scale(mk(pp.x, pp.y, 0.0f), (-pi2 * pi2));
    const V3 d2pduv = scale(scale(scale(mk(-long_sin, long_cos, 0.0f), fpi), pp.z), pi2);
    const V3 d2pdvv = scale(mk(pp.x, pp.y, pp.z), (float)(-PSY_PI_D * PSY_PI_D));
    const float E = dot(*dpdu, *dpdu), F = dot(*dpdu, *dpdv), G = dot(*dpdv, *dpdv);
    const float e = dot(n, d2pduu), f = dot(n, d2pduv), g = dot(n, d2pdvv);
    const float invEGF2 = 1.0f / ((E * G) - (F * F));
    *dndu = add(scale(*dpdu, ((f * F) - (e * G)) * invEGF2), scale(*dpdv, ((e * F) - (f * E)) * invEGF2));
    *dndv = add(scale(*dpdu, ((g * F) - (f * G)) * invEGF2), scale(*dpdv, ((f * F) - (g * E)) * invEGF2));
}

PSY_DEV_FN void intersection_record(const DevScene& s, uint32_t li, float t, float u, float v, float aux, V3 wo,
                                    V3 wd, float time, IsectRec& R) {
    const DevInstance I = s.inst[li];
    const DevObject ob = s.obj[I.object];
    R.space = m_identity();
    V3 o = wo, d = wd;
    if (I.xf_count > 0) {
        R.space = m_lerp_seq(time, s.xf + I.xf_offset, I.xf_count);
        o = m_pos(R.space, wo);
        d = m_dir(R.space, wd);
    }
    const float* od = s.odata + ob.data;
    R.p = add(o, scale(d, t));
    R.light_pdf = 9999.0f;  // Intersection's default where no light sets it
    R.shader = -1;
    R.dpdu = R.dpdv = R.dndu = R.dndv = splat(0.0f);
    if (ob.kind == K_RECT_LIGHT) {
        float w, h, col[3];
        rect_light_at(od, ob.ts, time, &w, &h, col);
        R.p = mk(o.x + (d.x * t), o.y + (d.y * t), 0.0f);
        R.n = mk(0.0f, 0.0f, 1.0f);
        R.backfacing = d.z > 0.0f ? 1u : 0u;
        R.light_pdf = rect_light_pdf(w, h, o);
        R.offset = scale(R.n, 0.000001f);
        const double area = (double)(w * h);
        R.closure = PSY_SHADER_EMIT;
        for (int k = 0; k < 3; ++k) R.col[k] = ((col[k] * 0.5f) / (float)area) * s.escale[I.object];
        return;
    }
    if (ob.kind == K_SPHERE_LIGHT) {
        V3 c; float r, col[3];
        sphere_light_at(od, ob.ts, time, &c, &r, col);
        const double area = 4.0 * PSY_PI_D * r * r;
        R.n = normalized(sub(R.p, c));
        R.backfacing = dot(R.n, normalized(d)) > 0.0f ? 1u : 0u;
        R.light_pdf = sphere_light_pdf(c, (double)r, o);
        R.offset = scale(R.n, 0.000001f);
        R.closure = PSY_SHADER_EMIT;
        for (int k = 0; k < 3; ++k) R.col[k] = (col[k] / (float)area) * s.escale[I.object];
        return;
    }
    if (ob.kind == K_SPHERE) {
        V3 c; float r;
        sphere_at(od, ob.ts, 4, time, &c, &r);
        R.n = normalized(sub(R.p, c));
        sphere_dg(R.n, r, &R.dpdu, &R.dpdv, &R.dndu, &R.dndv);
        R.offset = scale(R.n, 0.000001f);
    } else {
        if (ob.kind == K_BILINEAR) {
            Patch<4> pp;
            load_patch<4>(s.odata, ob.data, pp);
            patch_dg<4>(pp, u, v, &R.n, &R.dpdu, &R.dpdv, &R.dndu, &R.dndv);
        } else {
            Patch<16> pp;
            load_patch<16>(s.odata, ob.data, pp);
            patch_dg<16>(pp, u, v, &R.n, &R.dpdu, &R.dpdv, &R.dndu, &R.dndv);
        }
        R.offset = scale(R.n, aux * 1.74f);
    }
    R.backfacing = dot(R.n, normalized(d)) > 0.0f ? 1u : 0u;
    R.shader = I.shader;
    if (I.shader >= 0) {
        const DevShader sh = s.shaders[I.shader];
        R.closure = (uint32_t)sh.kind;
        R.col[0] = sh.color[0]; R.col[1] = sh.color[1]; R.col[2] = sh.color[2];
    } else {
        R.closure = PSY_SHADER_EMIT;
        R.col[0] = 1.0f; R.col[1] = 0.0f; R.col[2] = 1.0f;
    }
}

// ---- adjoint (path replay) ---------------------------------------------------
// The adjoint pass re-runs every path from its sample index with the same
// sampler, scene and parameters, so it reproduces the forward paths exactly
// without storing them.  Each radiance contribution C of a path is the
// product of a factor B that no parameter enters, the grey albedos of the
// path's scattering vertices a_v and, for light emission, the light's
// emission scale e:  C = B * prod(a_v) * e.  With w = dLoss/dC of the path
// (its pixel's film gradient through Color_XYZ), the pass adds
//   dLoss/de   += w * B * prod(a_v)
//   dLoss/da_v += w * B * e * prod(a_u, u != v)   for every vertex v.
// Light selection does not depend on the parameters (the light trees keep
// the energies of the scene as created), so every C is a polynomial in them.
enum : int {
    Q_W = 0,       // w: dLoss/dC for this path
    Q_B = 1,       // B of the throughput (albedos taken as 1)
    Q_LB = 2,      // B of the pending shadow ray's light sample
    Q_LOBJ = 3,    // its light object (i32), -1 for none
    Q_CHAIN = 4,   // shader of scattering vertex 0, 1, ... (u32)
    Q_COUNT = Q_CHAIN + ADJ_MAX_VERTS
};
static_assert(Q_COUNT == ADJ_STATE_PLANES, "adjoint state layout");

// The forward integrator: no adjoint state.
struct NoAdjoint {
    static constexpr bool on = false;
    PathBuf Q{nullptr, 0};
    int32_t n_albedo = 0;
    PSY_DEV int light_param(int) const { return -1; }
    PSY_DEV void add(int, float) const {}
};

// Gradients of one contribution w*B (`wb`) with `nverts` scattering vertices
// and light object `light_obj` (-1: no scaled emitter).
template <class Adj>
PSY_DEV void adj_contrib(const DevScene& s, const Adj& adj, uint32_t i, int nverts, float wb, int32_t light_obj) {
    if (wb == 0.0f) return;
    float a[ADJ_MAX_VERTS];
    uint32_t sh[ADJ_MAX_VERTS];
    float all = 1.0f;
#pragma unroll
    for (int v = 0; v < ADJ_MAX_VERTS; ++v) {
        sh[v] = 0u;
        a[v] = 1.0f;
        if (v < nverts) {
            sh[v] = adj.Q.u(Q_CHAIN + v, i);
            a[v] = s.shaders[sh[v]].color[0];
            all *= a[v];
        }
    }
    float e = 1.0f;
    const int lp = light_obj >= 0 ? adj.light_param(light_obj) : -1;
    if (lp >= 0) {
        e = s.escale[light_obj];
        adj.add(adj.n_albedo + lp, wb * all);
    }
#pragma unroll
    for (int v = 0; v < ADJ_MAX_VERTS; ++v) {
        if (v < nverts) {
            float ex = e;
#pragma unroll
            for (int u = 0; u < ADJ_MAX_VERTS; ++u)
                if (u != v && u < nverts) ex *= a[u];
            adj.add((int)sh[v], wb * ex);
        }
    }
}

// The image sample a path slot holds: (y * width + x) * spp + s.
PSY_DEV int64_t path_sample(const RenderLaunch& A, const PathBuf& P, uint32_t i) {
    const uint32_t pix = P.u(P_PIX, i);
    return ((int64_t)(pix >> 16) * A.width + (int64_t)(pix & 0xffffu)) * A.spp + (int64_t)P.u(P_SIDX, i);
}

// update_path for the ray just traced (slot j of the stream) of path i, then
// next_ray_for_path if the path goes on (returns true, the ray in `nr`);
// a finished path writes its film contribution to sample_xyz.
// PAR: the parity of every path's step when the launch fixes it (0 or 1: a
// wavefront step, where all paths of the stream are at the same step, so
// each launch compiles only its half), else -1 (read per path).
template <class Adj, int PAR = -1>
PSY_DEV bool render_update_path(const RenderLaunch& A, int64_t first, const PathBuf& P, uint32_t i, const WRay& ray,
                                const HitRec& hr, float* sample_xyz, int64_t sbase, WRay& nr, const Adj& adj) {
    bool alive = false;
    V3 no = splat(0.0f), nd = splat(0.0f), nodx = splat(0.0f), nody = splat(0.0f), nddx = splat(0.0f),
       nddy = splat(0.0f);
    float ntime = 0.0f;
    uint32_t ntype = 0;
    {
        const DevScene& s = A.s;
        uint32_t stepw = P.u(P_STEP, i);
        const int step = (int)(stepw & 0xffu);
        const bool odd = PAR < 0 ? (step & 1) != 0 : PAR == 1;
        bool done = false;
        const uint32_t hinst = hr.inst;
        const bool hit = hinst != 0xffffffffu;
        const V3 rd = ray.d;
        float col[NSPEC], fcol[NSPEC];
        for (int k = 0; k < NSPEC; ++k) { col[k] = P.f(P_COL + k, i); fcol[k] = P.f(P_FCOL + k, i); }
        if (odd) {
            // result of a shadow ray: unoccluded -> add the light sample
            // (the adjoint takes Lambert and Emit shaders only: its scene has no GTR)
            const DevShader SH = s.shaders[P.u(P_SHADER, i)];
            const bool sgtr = !Adj::on && SH.kind == PSY_SHADER_GTR;
            if (!hit && !(sgtr && SH.roughness == 0.0f)) {
                float fac[NSPEC];
                if (sgtr) gtr_eval(A.spec, SH, P.v3(P_PD, i), rd, P.v3(P_GN, i), P.f(P_WL, i), fac);
                else lambert_eval(A.spec, SH.color, P.v3(P_PD, i), rd, P.v3(P_GN, i), P.f(P_WL, i), fac);
                for (int k = 0; k < NSPEC; ++k) col[k] += fcol[k] * P.f(P_LCOL + k, i) * fac[k];
                if constexpr (Adj::on) {
                    const Adj& Z = adj;
                    const float fac1 = lambert_eval1(P.v3(P_PD, i), rd, P.v3(P_GN, i));
                    const float wb = Z.Q.f(Q_W, i) * ((Z.Q.f(Q_B, i) * Z.Q.f(Q_LB, i)) * fac1);
                    adj_contrib(s, Z, i, (step >> 1) + 1, wb, (int32_t)Z.Q.u(Q_LOBJ, i));
                }
            }
        } else if (hit) {
            const DevInstance I = s.inst[hinst];
            const V3 ro = ray.o;
            const float time = ray.time;
            bool emit = false;
            float em[NSPEC], em_base[NSPEC], light_pdf = 9999.0f;
            int32_t em_obj = -1;
            if (I.kind == K_SPHERE_LIGHT || I.kind == K_RECT_LIGHT) {
                emit = true;
                light_hit(s, A.spec, hinst, ro, time, P.f(P_WL, i), em, em_base, &light_pdf);
                em_obj = I.object;
            } else if (I.shader < 0 || s.shaders[I.shader].kind == PSY_SHADER_EMIT) {
                // an EmitShader, or no shader at all: the reference's
                // magenta emitter (Color(1, 0, 1), no MIS light pdf)
                emit = true;
                const float magenta[3] = {1.0f, 0.0f, 1.0f};
                const float* c = I.shader < 0 ? magenta : s.shaders[I.shader].color;
                color_spec4(A.spec, c[0], c[1], c[2], P.f(P_WL, i), em);
                for (int k = 0; k < NSPEC; ++k) em_base[k] = em[k];
            }
            if (emit) {
                done = true;
                const float last_pdf = P.f(P_LASTPDF, i);
                float wb = 0.0f;
                if (last_pdf != 0.0f) {
                    const float a2 = last_pdf * last_pdf;
                    const float b2 = light_pdf * light_pdf;
                    const float mis_inv_pdf = (a2 / (a2 + b2)) / last_pdf;
                    for (int k = 0; k < NSPEC; ++k) col[k] += ((fcol[k] * last_pdf) * em[k]) * mis_inv_pdf;
                    if constexpr (Adj::on) wb = ((adj.Q.f(Q_B, i) * last_pdf) * em_base[0]) * mis_inv_pdf;
                } else {
                    for (int k = 0; k < NSPEC; ++k) col[k] += fcol[k] * em[k];
                    if constexpr (Adj::on) wb = adj.Q.f(Q_B, i) * em_base[0];
                }
                if constexpr (Adj::on) adj_contrib(s, adj, i, step >> 1, adj.Q.f(Q_W, i) * wb, em_obj);
            } else {
                // store the hit in world space (DifferentialGeometry::transformed_from)
                M44 space;
                V3 pl, nl, offl;
                bool backfacing = false;
                const float t = hr.t;
                hit_geometry(s, hinst, t, hr.u, hr.v, hr.aux, ro, rd, time, &space, &pl, &nl, &offl, &backfacing);
                const M44 inv = m_inverse(space);
                const V3 gp = m_pos(inv, pl);
                V3 gn = m_dir_transpose(space, nl);
                const float il = 1.0f / length(gn);
                gn = scale(gn, il);
                V3 po = m_dir(inv, offl);
                if (backfacing) po = scale(po, -1.0f);
                P.set3(P_GP, i, gp);
                P.set3(P_GN, i, gn);
                P.set3(P_POFF, i, po);
                P.f(P_T, i) = t;
                P.setu(P_SHADER, i, (uint32_t)I.shader);
                if (!Adj::on && s.shaders[I.shader].kind == PSY_SHADER_GTR) {
                    // the rest of the world-space DifferentialGeometry
                    // (basics/differential_geometry.hpp:20-40), which the
                    // GTR closure's differentials need
                    IsectRec R;
                    intersection_record(s, hinst, t, hr.u, hr.v, hr.aux, ro, rd, time, R);
                    P.set3(P_DPDU, i, m_dir(inv, R.dpdu));
                    P.set3(P_DPDV, i, m_dir(inv, R.dpdv));
                    P.set3(P_DNDU, i, scale(m_dir_transpose(space, R.dndu), il));
                    P.set3(P_DNDV, i, scale(m_dir_transpose(space, R.dndv), il));
                }
                if constexpr (Adj::on) adj.Q.setu(Q_CHAIN + (step >> 1), i, (uint32_t)I.shader);
                P.set3(P_PD, i, rd);
                P.set3(P_PODX, i, ray.odx);
                P.set3(P_PODY, i, ray.ody);
                P.set3(P_PDDX, i, ray.ddx);
                P.set3(P_PDDY, i, ray.ddy);
                stepw = backfacing ? (stepw | STEP_BACKFACING) : (stepw & ~STEP_BACKFACING);
            }
        } else {
            done = true;
            float bg[NSPEC];
            color_spec4(A.spec, A.bg[0], A.bg[1], A.bg[2], P.f(P_WL, i), bg);
            for (int k = 0; k < NSPEC; ++k) col[k] += fcol[k] * bg[k];
            if constexpr (Adj::on)
                adj_contrib(s, adj, i, step >> 1, adj.Q.f(Q_W, i) * (adj.Q.f(Q_B, i) * bg[0]), -1);
        }
        const int nstep = step + 1;
        const bool next_odd = !odd;  // nstep & 1
        if (nstep == A.path_length * 2) done = true;

        if (!done) {
            // next_ray_for_path for step nstep
            const uint32_t off = P.u(P_OFF, i);
            uint32_t dim = P.u(P_DIM, i);
            const V3 gp = P.v3(P_GP, i), gn = P.v3(P_GN, i), po = P.v3(P_POFF, i), pd = P.v3(P_PD, i);
            const float ht = P.f(P_T, i);
            const float time = P.f(P_TIME, i);
            const DevShader SH = s.shaders[P.u(P_SHADER, i)];
            const float* alb = SH.color;
            const bool gtr = !Adj::on && SH.kind == PSY_SHADER_GTR;
            const bool delta = gtr && SH.roughness == 0.0f;  // GTRClosure::is_delta
            if (next_odd && delta) {
                // a delta closure takes no light sample (and no sampler
                // dimensions): a shadow ray of zero light, whose result is
                // not used
                for (int k = 0; k < NSPEC; ++k) P.f(P_LCOL + k, i) = 0.0f;
                if constexpr (Adj::on) {
                    adj.Q.f(Q_LB, i) = 0.0f;
                    adj.Q.setu(Q_LOBJ, i, 0xffffffffu);
                }
                no = add(gp, po);
                nd = mk(1e-20f, 1e-20f, 1e-20f);
                ntype = PSY_RAY_OCCLUSION;
            } else if (next_odd) {
                // shadow ray towards a light sample
                V3 lq_nor = gn;
                if (stepw & STEP_BACKFACING) lq_nor = scale(lq_nor, -1.0f);
                LQuery q;
                q.n = next_dim(A, off, dim);
                q.u = next_dim(A, off, dim);
                q.v = next_dim(A, off, dim);
                q.pos = gp;
                q.nor = lq_nor;
                q.d = pd;
                q.time = time;
                q.wl = P.f(P_WL, i);
                q.xform = m_identity();
                q.to_light = splat(0.0f);
                for (int k = 0; k < NSPEC; ++k) {
                    q.spec[k] = 0.0f;
                    q.spec_base[k] = 0.0f;
                }
                q.scale = 1.0f;
                q.obj = -1;
                q.selection_pdf = 1.0f;
                q.light_sample_pdf = 1.0f;
                q.bsdf = gtr ? PSY_SHADER_GTR : PSY_SHADER_LAMBERT;
                q.g_rough = SH.roughness;
                q.g_tail = SH.tail_shape;
                float lcol[NSPEC] = {0.0f, 0.0f, 0.0f, 0.0f}, lbase = 0.0f;
                if (!light_tree_sample(s, A.L, A.spec, q)) {
                    q.to_light = mk(0.0f, 0.0f, 1.0f);
                    q.obj = -1;
                } else {
                    const float bsdf_pdf = gtr ? gtr_pdf(SH, pd, q.to_light, gn) : lambert_pdf(pd, q.to_light, gn);
                    const float a2 = q.light_sample_pdf * q.light_sample_pdf;
                    const float b2 = bsdf_pdf * bsdf_pdf;
                    const float mis_inv_pdf = (a2 / (a2 + b2)) / q.light_sample_pdf;
                    for (int k = 0; k < NSPEC; ++k)
                        lcol[k] = ((q.spec[k] * mis_inv_pdf) * (float)A.L.asms[0].total_lights) / q.selection_pdf;
                    if constexpr (Adj::on)
                        lbase = ((q.spec_base[0] * mis_inv_pdf) * (float)A.L.asms[0].total_lights) / q.selection_pdf;
                }
                for (int k = 0; k < NSPEC; ++k) P.f(P_LCOL + k, i) = lcol[k];
                if constexpr (Adj::on) {
                    adj.Q.f(Q_LB, i) = lbase;
                    adj.Q.setu(Q_LOBJ, i, (uint32_t)q.obj);
                }
                no = add(gp, po);
                nd = scale(sub(q.to_light, po), 0.9995f);
                ntype = PSY_RAY_OCCLUSION;
            } else if (gtr) {
                // bounce ray: GTRClosure::sample
                const float su = next_dim(A, off, dim);
                const float sv = next_dim(A, off, dim);
                nd = gtr_sample_dir(SH, pd, gn, su, sv);
                float pdf = gtr_pdf(SH, pd, nd, gn);
                float filt[NSPEC];
                gtr_eval(A.spec, SH, pd, nd, gn, P.f(P_WL, i), filt);
                no = add(gp, po);
                ntype = 2u;  // WorldRay::R_DIFFUSE
                if (!delta) {
                    if (pdf < 0.00001f) pdf = 0.00001f;
                    for (int k = 0; k < NSPEC; ++k) fcol[k] *= filt[k] / pdf;
                    P.f(P_LASTPDF, i) = pdf;
                } else {
                    for (int k = 0; k < NSPEC; ++k) fcol[k] *= filt[k];
                    P.f(P_LASTPDF, i) = 0.0f;
                }
            } else {
                // bounce ray: LambertClosure::sample
                float su = next_dim(A, off, dim);
                float sv = next_dim(A, off, dim);
                V3 nn = normalized(gn);
                if (dot(nn, pd) > 0.0f) nn = scale(nn, -1.0f);
                su = (su * 2) - 1;
                sv = (sv * 2) - 1;
                square_to_circle(&su, &sv);
                const double zz = 1.0 - (double)((su * su) + (sv * sv));
                const float zc = (float)sqrt(zz > 0.0 ? zz : 0.0);
                const V3 dir = mk(su, sv, zc);
                float pdf = dir.z * (float)(1.0 / PSY_PI_D);
                nd = zup_to_vec(dir, nn);
                no = add(gp, po);
                ntype = 2u;  // WorldRay::R_DIFFUSE
                float filt[NSPEC];
                lambert_eval(A.spec, alb, pd, nd, gn, P.f(P_WL, i), filt);
                if (pdf < 0.00001f) pdf = 0.00001f;
                for (int k = 0; k < NSPEC; ++k) fcol[k] *= filt[k] / pdf;
                if constexpr (Adj::on) adj.Q.f(Q_B, i) *= lambert_eval1(pd, nd, gn) / pdf;
                P.f(P_LASTPDF, i) = pdf;
            }
            if (gtr) {
                // GTRClosure::propagate_differentials (a delta closure's
                // zero-light shadow ray keeps zero differentials)
                if (!(next_odd && delta))
                    gtr_differentials(ht, pd, P.v3(P_PODX, i), P.v3(P_PODY, i), P.v3(P_PDDX, i), P.v3(P_PDDY, i), gn,
                                      P.v3(P_DPDU, i), P.v3(P_DPDV, i), P.v3(P_DNDU, i), P.v3(P_DNDV, i), nd, &nodx,
                                      &nody, &nddx, &nddy);
            } else {
                // LambertClosure::propagate_differentials
                const float len = length(nd);
                const V3 gnn = normalized(gn);
                V3 x, y;
                coordinate_system_from_vec3(nd, &x, &y);
                x = normalized(x);
                y = normalized(y);
                nodx = transfer_origin_diff(ht, gnn, pd, P.v3(P_PODX, i), P.v3(P_PDDX, i));
                nody = transfer_origin_diff(ht, gnn, pd, P.v3(P_PODY, i), P.v3(P_PDDY, i));
                nddx = sdiv(scale(x, 0.15f), len);
                nddy = sdiv(scale(y, 0.15f), len);
            }
            ntime = time;
            P.setu(P_DIM, i, dim);
            alive = true;
        }
        for (int k = 0; k < NSPEC; ++k) { P.f(P_COL + k, i) = col[k]; P.f(P_FCOL + k, i) = fcol[k]; }
        P.setu(P_STEP, i, (stepw & ~0xffu) | (uint32_t)nstep);
        if (!alive && sample_xyz) {
            float xyz[3];
            path_film_xyz(P.f(P_WL, i), col, xyz);
            const int64_t g = path_sample(A, P, i) - sbase;
            sample_xyz[3 * g + 0] = xyz[0];
            sample_xyz[3 * g + 1] = xyz[1];
            sample_xyz[3 * g + 2] = xyz[2];
        }
    }
    nr.o = no; nr.d = nd; nr.odx = nodx; nr.ody = nody; nr.ddx = nddx; nr.ddy = nddy;
    nr.time = ntime;
    nr.type = ntype;
    return alive;
}

// A lane of the fused path kernel starts tracing world ray w (WorldRay::
// to_ray, basics/ray.hpp:310-323, then the traversal state).
PSY_DEV void lane_init_wray(const DevScene& s, const WRay& w, uint32_t i, Lane& L) {
    L.r.o = w.o;
    L.r.d = w.d;
    L.r.odx = w.odx;
    L.r.ody = w.ody;
    L.r.ddx = w.ddx;
    L.r.ddy = w.ddy;
    L.r.time = w.time;
    L.r.occl = w.type == PSY_RAY_OCCLUSION;
    L.r.w = ray_widths(L.r.o, L.r.d, L.r.odx, L.r.ody, L.r.ddx, L.r.ddy);
    L.r.dinv = divv(splat(1.0f), L.r.d);
    lane_start(s, L, L.r.occl ? 1.0f : INFINITY, i);
}

// The ray a lane just traced, as a world ray.
PSY_DEV void lane_wray(const Lane& L, uint32_t type, WRay& w) {
    w.o = L.r.o;
    w.d = L.r.d;
    w.odx = L.r.odx;
    w.ody = L.r.ody;
    w.ddx = L.r.ddx;
    w.ddy = L.r.ddy;
    w.time = L.r.time;
    w.type = type;
}

// render_update_path for slot j of a ray / hit stream.
template <class Adj = NoAdjoint, int PAR = -1>
PSY_DEV bool render_update_one(const RenderLaunch& A, int64_t first, const PathBuf& P, uint32_t i, int64_t j,
                               const float* rays_in, const DevHits& hits, int64_t rcap, float* sample_xyz,
                               int64_t sbase, WRay& nr, const Adj& adj = Adj()) {
    WRay ray;
    load_wray(rays_in, rcap, j, ray);
    HitRec hr;
    hr.t = hits.t[j];
    hr.inst = hits.inst[j];
    hr.u = hits.u[j];
    hr.v = hits.v[j];
    hr.aux = hits.aux ? hits.aux[j] : 0.0f;
    return render_update_path<Adj, PAR>(A, first, P, i, ray, hr, sample_xyz, sbase, nr, adj);
}

// The adjoint state of a new path: w = dLoss/dC from its pixel's film
// gradient (film = mean over the pixel's spp samples of Color_XYZ(col) *
// XYZ_NORM_FAC, and a contribution C adds to every spectral sample of col),
// B = 1, no pending light sample.
PSY_DEV void adjoint_init_one(const RenderLaunch& A, const PathBuf& P, const PathBuf& Q, int64_t i,
                              const float* d_film) {
    const uint32_t pix = P.u(P_PIX, i);
    const int64_t p = (int64_t)(pix >> 16) * A.width + (int64_t)(pix & 0xffffu);
    const float ones[NSPEC] = {1.0f, 1.0f, 1.0f, 1.0f};
    float xyz[3];
    path_film_xyz(P.f(P_WL, i), ones, xyz);
    const float w = ((d_film[3 * p] * xyz[0]) + (d_film[3 * p + 1] * xyz[1]) + (d_film[3 * p + 2] * xyz[2])) /
                    (float)A.spp;
    Q.f(Q_W, i) = w;
    Q.f(Q_B, i) = 1.0f;
    Q.f(Q_LB, i) = 0.0f;
    Q.setu(Q_LOBJ, i, 0xffffffffu);
}

// Film::add_sample over a pixel's samples, in sample order: the mean of the
// finite samples (NaN / inf samples are skipped as the reference does).
PSY_DEV void render_resolve_one(const float* sample_xyz, int64_t sbase, int64_t first_pixel, int64_t p, int spp,
                                float* film) {
    const int64_t g = first_pixel + p;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    int cnt = 0;
    for (int s = 0; s < spp; ++s) {
        const float* v = sample_xyz + 3 * (g * spp + s - sbase);
        const float a = v[0], b = v[1], c = v[2];
        if (!(fabsf(a) < INFINITY) || !(fabsf(b) < INFINITY) || !(fabsf(c) < INFINITY)) continue;
        sx += a; sy += b; sz += c;
        ++cnt;
    }
    const float inv = cnt > 0 ? 1.0f / (float)cnt : 0.0f;
    film[3 * g + 0] = sx * inv;
    film[3 * g + 1] = sy * inv;
    film[3 * g + 2] = sz * inv;
}

}  // namespace psy
