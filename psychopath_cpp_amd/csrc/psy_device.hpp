// psy_device.hpp -- HBM layout of an uploaded scene and the launcher
// interface between the host library and the HIP kernels.
//
// Layout (all arrays in device memory, read-only after psy_scene_create):
//   bvh    : BVH4 over flattened leaf instances, 128-byte nodes; child
//            boxes stored per axis as 4-wide lanes (the reference BBox4
//            layout, basics/bbox.hpp:461-476), 16 bytes of child refs,
//            16 bytes of padding (one node = one 128-byte L2 line).
//   inst   : 32-byte leaf-instance records (object, merged transforms,
//            effective shader, and for patches the dice-cache slot, so a
//            traversal reaching a patch leaf reads one record) -- the
//            assembly hierarchy flattened.
//   obj    : 32-byte object records; control points / centres in `odata`.
//   xf     : merged per-leaf transform lists (world -> object), 64 B each.
//   dice   : the patch dice cache: for every patch a complete 1-based
//            heap-ordered split tree of `dice_levels`+1 levels in a slot of
//            2^(levels+1) entries (entry 0 unused), 32 bytes per node:
//            float4(min.xyz, u_center), float4(max.xyz, v_center).
#pragma once

#include <stdint.h>
#include "psy_math.hpp"

namespace psy {

enum : int32_t {
    K_SPHERE = 1, K_BILINEAR = 2, K_BICUBIC = 3, K_SPHERE_LIGHT = 4, K_RECT_LIGHT = 5
};

constexpr int32_t CHILD_EMPTY = INT32_MIN;  // unused BVH4 lane
// child ref: >= 0 inner node index; < 0 and != CHILD_EMPTY: leaf ~ref

struct alignas(16) BvhNode {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    int32_t child[4];
    int32_t pad[4];  // keeps nodes 128-byte aligned
};
static_assert(sizeof(BvhNode) == 128, "BVH node must be 128 bytes");

struct alignas(16) DevInstance {
    int32_t object;
    int32_t kind;
    int32_t xf_offset;
    int32_t xf_count;
    int32_t shader;       // effective shader index, -1 = none
    int32_t dice_levels;  // patches: the object's cached split depth
    int64_t dice;         // patches: the object's first dice-cache node
};
static_assert(sizeof(DevInstance) == 32, "instance record must be 32 bytes");

struct alignas(16) DevObject {
    int32_t kind;
    int32_t ts;
    int32_t data;         // float offset into odata
    int32_t dice_levels;  // cached split depth (patches)
    int64_t dice;         // first node (in 32-byte units) in the dice pool
    int32_t pad[2];
};
static_assert(sizeof(DevObject) == 32, "object record must be 32 bytes");

// A surface shader; GTR ones carry GTRClosure's validated roughness and
// tail shape, its Fresnel blend and its normalization factor
// (shading/surface_closure.hpp:351-382), computed once on the host.
struct DevShader { int32_t kind; float color[3]; float roughness, tail_shape, fresnel, norm; };

constexpr int MAX_CAM_SAMPLES = 8;
struct DevCamera {
    M44 xf[MAX_CAM_SAMPLES]; int32_t n_xf;
    float tfov[MAX_CAM_SAMPLES]; int32_t n_tfov;
    float aperture[MAX_CAM_SAMPLES]; int32_t n_ap;
    float focus[MAX_CAM_SAMPLES]; int32_t n_fd;
};

// Everything a kernel needs to trace against a scene (passed by value).
struct DevScene {
    const BvhNode* bvh;
    const BvhNode* bvh1;      // moving scenes: each node's child box change
                              // from time 0 (bvh) to time 1, which a ray
                              // scales by its time; null for static scenes
    int32_t n_bvh;
    int32_t root;             // child ref of the root (node 0, or a leaf)
    const DevInstance* inst;
    const DevObject* obj;
    const float* odata;
    const M44* xf;
    const float4* dice;
    const DevShader* shaders;
    const float* escale;      // per object: emission scale of a light object (1 by default)
    float dice_rate;
    float min_upoly;
    int32_t n_inst;
    uint32_t flags;           // PSY_SCENE_*
};

// Halton tables (dimension d >= 1): perm table offset, pow_base, terms.
struct HaltonDim {
    uint32_t perm_offset, pow_base, n_terms, pad;
    uint32_t div[8];
    uint32_t power[8];
    double scale;
    double pad2;
};

struct DevHits { float* t; uint32_t* inst; float* u; float* v; float* aux; };

// ---- light trees and the assembly hierarchy (for light sampling) -----------
// One binary light tree per assembly (accel/light_tree.hpp:7-66); node
// indices are local to the assembly's node range.
struct alignas(16) DevLNode {
    int32_t leaf;
    int32_t index1, index2;  // children (inner nodes)
    int32_t instance;        // the assembly's instance (leaves)
    float energy;
    int32_t bounds, n_bounds;  // time-sampled bounds in DevLights::bounds
    int32_t pad;
};
static_assert(sizeof(DevLNode) == 32, "light tree node must be 32 bytes");
struct alignas(16) DevAsm {
    int32_t node0, n_nodes;  // light tree nodes
    int32_t inst0;           // first DevAsmInst
    int32_t xf0;             // first transform of the assembly's list
    int32_t sub0;            // first child-assembly scene index in DevLights::sub
    int32_t obj0;            // first global object
    int32_t total_lights;
    int32_t pad;
};
struct alignas(16) DevAsmInst { int32_t type, data_index, xf_index, xf_count; };
struct DevLights {
    const DevLNode* nodes;
    const float4* bounds;  // 2 per box: (min.xyz, 0), (max.xyz, 0)
    const DevAsm* asms;
    const DevAsmInst* inst;
    const M44* xf;
    const int32_t* sub;
    int32_t n_asms;
};

// ---- launchers (psy_kernels.hip) -------------------------------------------
int launch_build_dice(const float* odata, const DevObject* objs, const int32_t* patch_ids,
                      int n_patches, int levels, float4* dice, hipStream_t st);
// Second-pass defaults: the queue length from which its wide build walks
// the queue, and the split-tree node expansions per ray before the third
// pass takes the ray.
constexpr uint32_t DEEP_WIDE_MIN = 4096;
constexpr int32_t DEEP_BUDGET = 512;
// How the main trace pass is launched: persistent (persist_blocks blocks,
// lanes refilled when refill_min of a wave's 64 are idle) or one thread per
// ray.
struct TraceTuning {
    bool persistent;
    bool coherent = false;  // persistent: step one kind of work per wave step
    bool plain_coherent = false;  // one thread per ray, stepping one kind of work per wave step
    int persist_blocks;    // persistent grid with the 16-entry stack
    int persist_blocks32;  // and with the 32-entry stack
    int refill_min;
    int stack = 16;        // main-pass LDS stack entries: 16 or 32
    const uint32_t* n_dev = nullptr;  // device-side ray count (<= n), or NULL
    // second pass: queue length from which its wide build walks the queue
    // (0: always the wide build, UINT32_MAX: always the narrow one), and
    // the split-tree node expansions a ray may take there before it is
    // handed to the third pass
    uint32_t wide_min = DEEP_WIDE_MIN;
    int32_t deep_budget = DEEP_BUDGET;
    // the main pass's continuation buffer (cont_words words; NULL: none)
    // and the live-lane count below which a wave continues its rays on it
    uint32_t* cont = nullptr;
    int64_t cont_words = 0;
    int cont_below = 0;
    // hipEvent_t handles recorded around the passes when the context times
    // them: before the main pass, after it, after the second pass.
    void* ev[3] = {nullptr, nullptr, nullptr};
};
// Three stream-ordered passes; queue (6 n words: second- and third-pass
// queues, second-pass deep-node lists and pool) and counters (8 x u32) are
// scratch.
// Ray plane p of ray i is at rays[p * rs + i] (rs >= n).
int launch_trace(const DevScene& s, bool world, bool occl_only, int stack_need, const float* rays,
                 int64_t n, int64_t rs, const DevHits& h, uint32_t* queue, uint32_t* counters,
                 const TraceTuning& tune, hipStream_t st);
int trace_persist_occupancy(bool world, int stack, int* blocks_per_cu);
// Device address of the allocation-free trace's scratch array on `device`,
// which must be the current device (looked up on the first call per device).
// Layout: 6 * PSY_STATIC_MAX_RAYS words of queues, deep-node lists and
// pool, then the counters.
int static_scratch_address(int device, uint32_t** out);
int launch_world_to_rays(const float* wrays, int64_t n, float* rays, hipStream_t st);
// Intersection records (PSY_ISECT_PLANES planes) of the hits of world rays.
// Film::add_sample / variance_estimate passes per bucket (state: 10 W*H floats)
int launch_film_adaptive(const float* samples, int W, int H, int spp, int bs, int spp_max, float var_max, float* state,
                         float* film, float* var_out, int32_t* passes, hipStream_t st);
int launch_intersections(const DevScene& s, const float* wrays, int64_t n, const DevHits& h, float* out,
                         hipStream_t st);

// ---- integrator (psy_render_dev.hpp) ----------------------------------------
// The spectrum grid of colour upsampling (color/spectrum_grid.h) as the
// caller handed it over: cells of 8 ints (inside, num_points, idx[6]),
// w x h row-major; data points of 4 + n_samples floats (xystar[2], uv[2],
// spectrum).  cells == NULL: no grid (grey colours only, flat spectrum).
struct DevSpectrum {
    const int32_t* cells;
    const float* points;
    int32_t w, h, n_points, n_samples;
    float smin, smax;
    float m[6];  // xy -> uv
};
struct RenderLaunch {
    DevScene s;
    DevLights L;
    DevCamera cam;
    const HaltonDim* dims;
    const uint16_t* perm;
    const uint64_t* sobol_m;  // PSY_RENDER_SOBOL: the Sobol matrices (sobol_cols columns), else null
    uint32_t sobol_cols;
    int32_t width, height, spp, path_length;
    uint32_t seed;
    float bg[3];
    DevSpectrum spec;
};
constexpr int PATH_PLANES = 58;  // float planes of a path state
// Path buffers of one chunk of `cap` paths: states (PATH_PLANES), two world
// ray streams, hits, two active lists, counters (u32 x 4).
struct RenderBuffers {
    float* paths;
    float* rays[2];
    DevHits hits;
    uint32_t* act[2];
    uint32_t* count;  // [0], [1]: next active counts (ping-pong)
    int64_t cap;
};
int launch_render_init(const RenderLaunch& R, int64_t first, int64_t n, const RenderBuffers& B, hipStream_t st);
// Finished paths write their film XYZ to sample_xyz[3 * (sample - sbase)].
// Every path of the stream is at step `step` (the wavefront loop's).
// shade_waves 4: an even step's kernel is the one held to 4 waves per SIMD.
int launch_render_update(const RenderLaunch& R, int64_t first, const RenderBuffers& B, int cur, int64_t n_in,
                         int step, int shade_waves, float* sample_xyz, int64_t sbase, hipStream_t st);
// Stream form: device-side state of a pool of path slots.
struct StreamState {
    uint32_t count[2];         // ping-pong stream lengths
    uint32_t n_wait;           // slots parked after an even step
    uint32_t pad;
    unsigned long long next;   // next image sample to start
    unsigned long long end;    // end of the sample window
    unsigned long long rays;   // rays traced
};
int launch_stream_init(const RenderLaunch& R, const RenderBuffers& B, int64_t first, int64_t n, int64_t end,
                       StreamState* S, bool reset_rays, hipStream_t st);
int launch_stream_shade(const RenderLaunch& R, const RenderBuffers& B, int cur, StreamState* S, uint32_t* wait,
                        int blocks, float* sample_xyz, int64_t sbase, hipStream_t st);

// Adjoint pass buffers: Q planes (adjoint path state, `cap` entries each),
// the gradient vector (n_params: albedos by shader, then emission scales by
// light object), per-block partial rows when the sums fit in LDS.
constexpr int ADJ_MAX_VERTS = 8;  // scattering vertices kept per path (path_length limit)
constexpr int ADJ_STATE_PLANES = 4 + ADJ_MAX_VERTS;
constexpr int ADJ_LDS_MAX_PARAMS = 16384;
struct AdjointBuffers {
    float* q;
    const float* d_film;
    int32_t n_albedo, n_params;
    const int32_t* lparam;
    float* grad;
    float* partials;
    int rows;
    bool lds;
};
int launch_fill_f32(float* p, int64_t n, float v, hipStream_t st);
int launch_adjoint_init(const RenderLaunch& R, const RenderBuffers& B, const AdjointBuffers& J, int64_t n,
                        hipStream_t st);
// Every path of the stream is at step `step`.
int launch_render_adjoint(const RenderLaunch& R, int64_t first, const RenderBuffers& B, const AdjointBuffers& J,
                          int cur, int64_t n_in, int step, hipStream_t st);
int launch_adjoint_reduce(const AdjointBuffers& J, float* d_albedo, float* d_emission, hipStream_t st);
int launch_render_resolve(const float* sample_xyz, int64_t sbase, int64_t first_pixel, int64_t n_pixels, int spp,
                          float* film, hipStream_t st);
int launch_zero_u32(uint32_t* p, int n, hipStream_t st);
// Buffers of the fused path kernel for one chunk of `cap` paths.
struct PersistBuffers {
    float* paths;
    uint32_t* counters;     // u32 x 4 (see k_render_persist)
    uint32_t* resume_ids;   // paths handed back to the kernel, and their rays
    float* resume_rays;
    uint32_t* defer_ids;    // paths whose ray needs the second trace pass
    float* defer_rays;
    float* sample_xyz;
    int64_t sbase;
    int64_t cap;
};
int launch_render_persist(const RenderLaunch& R, int64_t first, int64_t n_new, const PersistBuffers& B,
                          int blocks, int refill_min, hipStream_t st);
int render_persist_occupancy(int* blocks_per_cu);
int launch_render_update_io(const RenderLaunch& R, int64_t first, float* paths, int64_t cap, const uint32_t* act_in,
                            const float* rays_in, const DevHits& hits, int64_t n_in, uint32_t* act_out,
                            float* rays_out, uint32_t* count_out, float* sample_xyz, int64_t sbase, hipStream_t st);
int launch_render_count_steps(float* paths, int64_t cap, int64_t n, unsigned long long* out, hipStream_t st);
// counters[0..7] = 0 except counters[3] = n_resume
int launch_set_counters(uint32_t* p, uint32_t n_resume, hipStream_t st);
int launch_sobol(const uint64_t* m, uint32_t n_cols, uint32_t dim, const uint64_t* idx, int64_t n,
                 uint64_t scramble, double* out, hipStream_t st);
// out[4 i + k] = Color_to_SpectralSample(rgb[3 i ..], hero[i]).e[k]
int launch_color_spectrum(const DevSpectrum& S, const float* rgb, const float* hero, int64_t n, float* out,
                          hipStream_t st);
int launch_halton(const HaltonDim* dims, const uint16_t* perm, uint32_t dim, const uint32_t* idx,
                  int64_t n, float* out, hipStream_t st);
int launch_pixel_samples(const HaltonDim* dims, const uint16_t* perm, int x0, int y0, int w, int h,
                         int spp, uint32_t seed, uint32_t dim0, int n_dims, float* out, hipStream_t st);
int launch_camera_rays(const DevCamera& cam, const HaltonDim* dims, const uint16_t* perm, int width,
                       int height, int spp, uint32_t seed, float* wrays, float* wavelength,
                       hipStream_t st);
int launch_bounce_rays(const DevScene& s, const HaltonDim* dims, const uint16_t* perm,
                       const float* in_wrays, const DevHits& hits, int64_t n, int width, int spp,
                       uint32_t seed, uint32_t dim_u, uint32_t dim_v, float* out_wrays,
                       uint32_t* valid, hipStream_t st);

}  // namespace psy
