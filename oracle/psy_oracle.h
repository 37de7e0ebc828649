/*
 * psy_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * The oracle: a plain-C CPU restatement of the reference's breadth-first
 * ray-stream hot path (psychopath_cpp), used only by tests/, the smoke
 * check and the CPU-baseline leg of bench.py, as the checker.  The product
 * library (libpsy_gpu.so) never links or calls it.
 *
 * Each function cites the reference file:line it restates in its
 * definition.  Scene input is the product's own psy_scene_desc, so both
 * sides consume identical inputs.
 */
#ifndef PSY_ORACLE_H
#define PSY_ORACLE_H

#include <stdint.h>
#include "../include/psy_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* WorldRay (basics/ray.hpp:162-185), AoS */
typedef struct {
    float o[3], d[3], odx[3], ody[3], ddx[3], ddy[3];
    float time;
    uint32_t type;   /* PSY_RAY_OCCLUSION = occlusion ray */
} ora_wray;

/* Intersection (basics/intersection.hpp:22-49), flattened to POD */
typedef struct {
    int32_t hit;
    int32_t backfacing;
    uint64_t id;          /* InstanceID::id */
    float t;
    float light_pdf;
    float space[16];      /* Transform::to */
    float u, v;
    float p[3], dpdu[3], dpdv[3], n[3], dndu[3], dndv[3];
    float offset[3];
    int32_t closure;      /* 0 none, PSY_SHADER_LAMBERT, PSY_SHADER_EMIT */
    float closure_col[3];
    float closure_prob;
    float patch_dim;      /* max_dim of the accepted patch leaf (0 if none) */
    float closure_param[3];  /* GTR: roughness, tail shape, Fresnel as the shader gives them */
} ora_inter;

/* Ray with derived footprint (the product's PSY_RAY_PLANES fields) */
typedef struct {
    float o[3], d[3], max_t, time, owx, owy, dwx, dwy, fwx, fwy;
    uint32_t flags;
} ora_ray_out;

/* sampling */
float ora_halton(uint32_t dim, uint32_t index);
void ora_halton_batch(uint32_t dim, const uint32_t* index, int64_t n, float* out);
double ora_sobol(const uint64_t* m, uint32_t n_cols, uint32_t dim, uint64_t index, uint64_t scramble);
void ora_sobol_batch(const uint64_t* m, uint32_t n_cols, uint32_t dim, const uint64_t* index, int64_t n,
                     uint64_t scramble, double* out);
uint32_t ora_hash_u32(uint32_t n, uint32_t seed);
uint32_t ora_sampler_offset(uint32_t x, uint32_t y, uint32_t n, uint32_t seed);
float ora_sampler_get(uint32_t offset, uint32_t dim);

/* math KAT helpers */
void ora_matrix_inverse(const float in[16], float out[16]);
int ora_bbox_ray(const float bmin[3], const float bmax[3], const float o[3],
                 const float d[3], float max_t, float* h0, float* h1);
unsigned ora_bbox4_ray(const float bounds[24], const float o[3], const float d[3],
                       float max_t, float near_out[4]);
void ora_world_to_rays(const ora_wray* w, int64_t n, ora_ray_out* out);
float ora_vec3_length(const float v[3]);
void ora_vec3_normalize(const float v[3], float out[3]);
float ora_vec3_dot(const float a[3], const float b[3]);
void ora_vec3_cross(const float a[3], const float b[3], float out[3]);
int ora_partition_ints(int* v, int n, int key, int* calls);

/* scene */
typedef struct ora_scene ora_scene;
/* Grey albedo per shader (Lambert shaders) and emission scale per light
 * object in scene order; the light trees keep the scene's energies. */
int ora_scene_set_params(ora_scene* s, const float* albedo, int n_albedo, const float* escale, int n_escale);
ora_scene* ora_scene_create(const psy_scene_desc* desc);
void ora_scene_destroy(ora_scene* s);
int64_t ora_scene_bvh4_nodes(const ora_scene* s, int assembly);

/* Tracer::trace (tracer/tracer.cpp:38-121) over n world rays.  Rays are
 * processed in chunks of `chunk` rays (one Tracer per chunk, as one
 * Tracer per worker bucket in the reference), spread over `threads`
 * threads.  Returns the number of rays traced. */
int64_t ora_trace(const ora_scene* s, const ora_wray* rays, int64_t n,
                  ora_inter* out, int threads, int64_t chunk);
/* The same with the BVH replaced by an exhaustive instance loop. */
int64_t ora_trace_brute(const ora_scene* s, const ora_wray* rays, int64_t n,
                        ora_inter* out, int threads, int64_t chunk);

/* Camera rays for every pixel sample, ordered (y, x, s):
 * init_path + next_ray_for_path step 0
 * (integrator/path_trace_integrator.cpp:91-125). */
void ora_camera_rays(const ora_scene* s, int width, int height, int spp,
                     uint32_t seed, int64_t first, int64_t count,
                     ora_wray* out, float* wavelength);

/* Lambert bounce rays from hits (integrator/path_trace_integrator.cpp:
 * 183-209).  valid[i] = 0 where the hit is missing or not Lambert. */
void ora_bounce_rays(const ora_scene* s, const ora_wray* in, const ora_inter* hits,
                     int64_t first, int64_t count, int width, int spp,
                     uint32_t seed, uint32_t dim_u, uint32_t dim_v,
                     ora_wray* out, uint32_t* valid);

/* PathTraceIntegrator::integrate (integrator/path_trace_integrator.cpp:28-384)
 * for every pixel sample; path_xyz (width*height*spp*3) receives each
 * sample's film contribution, ordered (y, x, s).  Returns rays traced, or
 * -1 if a chromatic colour was met (colour upsampling is restated for
 * achromatic colours only). */
int64_t ora_render(const ora_scene* s, int width, int height, int spp, uint32_t seed,
                   int path_length, int threads, float* path_xyz);
/* The same with every path dimension from Sobol::sample over matrices m
 * (n_cols columns) at the Sampler's index (the device's PSY_RENDER_SOBOL) */
int64_t ora_render_sobol(const ora_scene* s, int width, int height, int spp, uint32_t seed, int path_length,
                         int threads, float* path_xyz, const uint64_t* m, uint32_t n_cols);
/* Color_to_SpectralSample value for an achromatic colour */
float ora_color_to_spectrum(const float c[3]);
/* The spectrum grid (color/spectrum_grid.h) as data: cells of 8 ints
 * (inside, num_points, idx[6]), w x h row-major; data points of
 * 4 + n_samples floats (xystar[2], uv[2], spectrum[n_samples]); the xy -> uv
 * 3x2 matrix.  NULL cells clears it.  Returns 0, or -1 on bad sizes. */
int ora_spectrum_set_grid(const int32_t* cells, int w, int h, const float* points, int n_points, int n_samples,
                          float smin, float smax, const float* mat_xy_to_uv);
/* spectrum_xyz_to_p (color/spectrum_grid.h:21-135) over the set grid */
float ora_spectrum_xyz_to_p(float lambda, const float* xyz);
/* Color_to_SpectralSample (color/color.hpp:509-511) at hero's 4 wavelengths */
void ora_color_to_spectral_sample(const float c[3], float hero, float e[4]);

/* ---- single functions, for the oracle's own checks ------------------- */
/* Sphere::intersect_ray's quadratic (object/sphere.cpp:67-140) */
int ora_kat_sphere_t(const float o[3], const float d[3], const float c[3], float r, float max_t, float* t);
/* closest_ray_t2 (utils/utils.hpp:443-456) */
void ora_kat_closest_ray_t2(const float o1[3], const float d1[3], const float o2[3], const float d2[3],
                            float* t, float* dist);
/* Ray::min_width (basics/ray.hpp:132-153); w6 = owx owy dwx dwy fwx fwy */
float ora_kat_min_width(const float w6[6], float tnear, float tfar);
/* Bilinear / Bicubic split_u (dir 0) / split_v (object/bilinear.hpp:63-73,
 * object/bicubic.hpp:74-110); kind PSY_OBJ_BILINEAR or PSY_OBJ_BICUBIC */
void ora_kat_patch_split(int kind, const float* p, int dir, float* p1, float* p2);
/* differential_geometry (object/bilinear.hpp:100-141, bicubic.hpp:140-189) */
void ora_kat_patch_dg(int kind, const float* p, float u, float v, float n[3], float dpdu[3], float dpdv[3],
                      float dndu[3], float dndv[3]);
/* InstanceID::push_back sequence (basics/instance_id.hpp:18-23) */
uint64_t ora_kat_instance_id(const uint64_t* sub, const int* bits, int n);
/* SphereLight / RectangleLight sample_pdf and sample (lights/) */
float ora_kat_sphere_light_pdf(const float pos[3], float radius, const float arr[3]);
float ora_kat_rect_light_pdf(float w, float h, const float arr[3]);
void ora_kat_sphere_light_sample(const float c[3], float radius, const float arr[3], float u, float v,
                                 float to_light[3], float* pdf);
void ora_kat_rect_light_sample(float w, float h, const float arr[3], float u, float v, float to_light[3],
                               float* pdf);
/* faster_ln, fast_logit (utils/utils.hpp:335-343, utils/monte_carlo.hpp:17-21),
 * square_to_circle (utils/monte_carlo.hpp:29-53) */
float ora_kat_faster_ln(float x);
float ora_kat_fast_logit(float p, float width);
void ora_kat_square_to_circle(float* x, float* y);
/* X_1931 / Y_1931 / Z_1931 (color/color.hpp:150-165) */
void ora_kat_cie_1931(float lambda, float xyz[3]);
/* sphere_lambert (utils/monte_carlo.hpp:207-233) */
float ora_kat_sphere_lambert(float nlcos, float rcos);
/* LambertClosure evaluate / sample_pdf (shading/surface_closure.hpp:233-331) */
void ora_kat_lambert(float albedo, const float in[3], const float out[3], const float n[3], float* eval,
                     float* pdf);
/* LightTree::sample on the root tree (accel/light_tree.cpp:30-96) */
int ora_kat_light_select(const ora_scene* s, const float pos[3], const float nor[3], const float d[3], float time,
                         float n, float u, float v, uint64_t* id, float* selection_pdf, float* light_sample_pdf,
                         float to_light[3]);
int64_t ora_kat_total_lights(const ora_scene* s);
/* GTRClosure (shading/surface_closure.hpp:339-608); prm = roughness, tail
 * shape, Fresnel as a GTRShader gives them: validated roughness, tail shape
 * and normalization; dist; half_theta_sample; evaluate (grey colour, first
 * wavelength) with sample_pdf; sample's direction; the solid-angle estimate;
 * and the axis-angle rotation its differentials use (row-vector matrix) */
void ora_kat_gtr_params(const float prm[3], float out[3]);
float ora_kat_gtr_dist(const float prm[3], float nh);
float ora_kat_gtr_half_theta(const float prm[3], float u);
void ora_kat_gtr(float grey, const float prm[3], const float in[3], const float out[3], const float n[3],
                 float* eval, float* pdf);
void ora_kat_gtr_sample(const float prm[3], const float in[3], const float n[3], float si, float sj, float out[3]);
float ora_kat_gtr_estimate(const float prm[3], const float in[3], const float cent[3], float cos_theta,
                           const float nor[3]);
void ora_kat_axis_angle(const float axis[3], float angle, float m[16]);
/* SubdivisionSurface::finalize's patch extraction (object/subdivision_surface.cpp:
 * 85-233): Bezier patches, 48 floats each, into *out (free with ora_free);
 * returns the count, or -1 for a mesh it refuses.  The _ex form takes the
 * deepest level (5 in the reference), the level below which every active face
 * is split, whether to refine only the faces within two rings of the split
 * ones (else every face), and whether every active face of the deepest level
 * becomes an end cap. */
int64_t ora_subdiv_patches(const float* verts, int n_verts, const int32_t* counts, int n_faces,
                           const int32_t* indices, float** out);
int64_t ora_subdiv_patches_ex(const float* verts, int n_verts, const int32_t* counts, int n_faces,
                              const int32_t* indices, int max_level, int min_level, int prune, int cap_all,
                              float** out);
void ora_free(void* p);
/* The adaptive sampling loop's film (film/film.hpp:135-165 in
 * integrator/path_trace_integrator.cpp:296-376) over per-sample XYZ values
 * that every pass of a bucket adds again */
void ora_film_adaptive(const float* samples, int W, int H, int spp, int spp_max, float var_max, int bs,
                       float* film, float* var, int32_t* passes);

#ifdef __cplusplus
}
#endif

#endif
