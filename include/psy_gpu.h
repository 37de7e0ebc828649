/*
 * psy_gpu.h -- C ABI of the MI355X ray-stream library (libpsy_gpu.so).
 *
 * Plain C, POD structs, device pointers as void/float/uint32_t pointers,
 * int status codes with psy_last_error().  No torch or HIP types appear in
 * the signatures; a `stream` argument is a hipStream_t passed as void*
 * (NULL = the null stream).
 *
 * Every entry point below names the reference interface it stands in for
 * (paths relative to the psychopath_cpp source tree).
 *
 * Conventions carried over from the reference:
 *   - matrices are row-major float[16] in row-vector convention (p' = p*M),
 *     the layout of Matrix44::data (math/matrix.hpp:18-22, :99-106);
 *   - an instance's transforms map its parent space to its local space;
 *     n motion samples are evenly spaced over time in [0,1]
 *     (scene/assembly.hpp:26-65, utils/utils.hpp:76-91);
 *   - ray SoA "planes": plane p of an n-ray buffer starts at base + p*n.
 */
#ifndef PSY_GPU_H
#define PSY_GPU_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSY_ABI_VERSION 3

/* ---- status codes ---------------------------------------------------- */
enum psy_status {
    PSY_OK = 0,
    PSY_ERR_INVALID = 1,      /* bad argument / inconsistent description */
    PSY_ERR_HIP = 2,          /* a HIP runtime call failed                */
    PSY_ERR_UNSUPPORTED = 3,  /* valid input the device path cannot take  */
    PSY_ERR_NO_DEVICE = 4     /* no GPU visible                            */
};

/* ---- scene description (host memory, read during psy_scene_create) --- */

/* Object kinds (object/object.hpp:21-138, lights/sphere_light.hpp,
 * lights/rectangle_light.hpp). */
enum psy_object_kind {
    PSY_OBJ_SPHERE = 1,          /* per time sample: cx cy cz r                  (4)  */
    PSY_OBJ_BILINEAR = 2,        /* per time sample: v1 v2 v3 v4, xyz each      (12) */
    PSY_OBJ_BICUBIC = 3,         /* per time sample: 16 control points, xyz     (48) */
    PSY_OBJ_SPHERE_LIGHT = 4,    /* per time sample: cx cy cz r  R G B            (7) */
    PSY_OBJ_RECTANGLE_LIGHT = 5, /* per time sample: width height  R G B          (5) */
    PSY_OBJ_SUBDIVISION = 6      /* per time sample: n_vertices control points, xyz   */
};

/* A SubdivisionSurface (object/subdivision_surface.hpp:15-71) is its
 * control mesh: n_vertices points per time sample in `data`, and n_faces
 * faces of face_vert_counts[f] (>= 3) vertices each, listed in
 * face_vert_indices.  psy_scene_create turns it into the bicubic patches of
 * psy_subdivision_patches; every hit on them carries the instance's id.
 * The four mesh fields are read for this kind only. */
typedef struct psy_object_desc {
    int32_t kind;          /* enum psy_object_kind                      */
    int32_t time_samples;  /* >= 1                                      */
    const float* data;     /* time_samples * floats-per-kind            */
    int32_t n_vertices;    /* PSY_OBJ_SUBDIVISION: points per sample    */
    int32_t n_faces;       /* PSY_OBJ_SUBDIVISION                       */
    const int32_t* face_vert_counts;   /* PSY_OBJ_SUBDIVISION: n_faces  */
    const int32_t* face_vert_indices;  /* PSY_OBJ_SUBDIVISION           */
} psy_object_desc;

enum psy_instance_type { PSY_INSTANCE_OBJECT = 0, PSY_INSTANCE_ASSEMBLY = 1 };

/* One Instance (scene/assembly.hpp:26-65).  shader_index indexes the
 * scene-wide shader table (-1 = none, i.e. inherit from the parent). */
typedef struct psy_instance_desc {
    int32_t type;             /* enum psy_instance_type                      */
    int32_t data_index;       /* object index or child-assembly index        */
    int32_t transform_index;  /* first transform in the assembly's list     */
    int32_t transform_count;  /* 0 = untransformed                           */
    int32_t shader_index;     /* -1 = none                                   */
} psy_instance_desc;

/* Surface shaders (shading/surface_shader.hpp:28-71): EmitShader,
 * LambertShader, and GTRShader, the Disney GTR microfacet closure
 * (shading/surface_closure.hpp:339-608) with its roughness (0 gives a
 * perfect mirror, a delta closure), tail shape and Fresnel blend; the
 * three GTR fields are ignored for the other kinds. */
enum psy_shader_kind { PSY_SHADER_LAMBERT = 1, PSY_SHADER_EMIT = 2, PSY_SHADER_GTR = 3 };
typedef struct psy_shader_desc {
    int32_t kind;
    float color[3];  /* linear RGB with white point rgb<1,1,1> */
    float roughness, tail_shape, fresnel;  /* PSY_SHADER_GTR */
} psy_shader_desc;

/* One Assembly (scene/assembly.hpp:74-330).  Child assemblies are indices
 * into psy_scene_desc.assemblies; entry 0 of that array is the root. */
typedef struct psy_assembly_desc {
    const psy_object_desc* objects;     int32_t n_objects;
    const psy_instance_desc* instances; int32_t n_instances;
    const float* transforms;            int32_t n_transforms;  /* 16 floats each */
    const int32_t* assemblies;          int32_t n_assemblies;  /* scene indices  */
} psy_assembly_desc;

/* Camera (basics/camera.hpp:14-105): per-sample lists, fov in radians. */
typedef struct psy_camera_desc {
    const float* transforms; int32_t n_transforms;  /* camera-to-world, 16 floats */
    const float* fovs;       int32_t n_fovs;
    const float* aperture_radii;  int32_t n_aperture_radii;
    const float* focus_distances; int32_t n_focus_distances;
} psy_camera_desc;

typedef struct psy_scene_desc {
    const psy_assembly_desc* assemblies; int32_t n_assemblies;  /* [0] = root */
    const psy_shader_desc* shaders;      int32_t n_shaders;
    psy_camera_desc camera;
    float background[3];
    float dice_rate;       /* Config::dice_rate (config.cpp:7), reference default 0.25 */
    float min_upoly_size;  /* Config::min_upoly_size (config.cpp:8), default 1e-5       */
    uint32_t flags;        /* PSY_SCENE_* */
} psy_scene_desc;
/* RectangleLight::intersect_ray tests the hit's y against half the width
 * (lights/rectangle_light.hpp:127), so a light taller than wide loses the
 * part of its top beyond half its width (the light samples still cover it),
 * and one wider than tall is hit beyond half its height wherever the
 * light's own bounds (which stop there) are not tested first, as when it is
 * a scene's only instance.  Scenes keep that behaviour unless they set this
 * flag, which tests y against half the height. */
#define PSY_SCENE_RECT_LIGHT_FIX 1u

/* ---- scene files (.psy) ------------------------------------------------ */
/* Host-only parsing of the reference's scene format into the description
 * above.  Replaces: DataTree::build_from_file (parser/data_tree.cpp:
 * 333-380, here over text already in memory) and Parser::parse_next_frame
 * (parser/parser.cpp:37-183) with the node parsers it calls
 * (parser/parser.cpp:186-662).  What the reference's parser reads is read
 * the same way (Fov in degrees; numbers as -?[0-9]+[.]?[0-9]*; names with
 * their '$'; object and assembly names resolved in their own assembly,
 * shader names up the parent chain); what it skips with a message (unknown
 * node types, instances of unknown data, unknown shader types) is skipped
 * and reported by psy_parsed_scene_warnings.  Scenes the device path cannot
 * render fail with PSY_ERR_UNSUPPORTED: lights whose motion sample lists
 * differ in length (other than 1). */
typedef struct psy_parsed_scene psy_parsed_scene;
typedef struct psy_render_settings {
    int32_t width, height;   /* RenderSettings Resolution (default 1 1)      */
    int32_t spp;             /* SamplesPerPixel (default 1)                  */
    int32_t seed;            /* Seed (default 1)                             */
    float dice_rate;         /* DicingRate (default 0.25; also in the desc)  */
    char name[256];          /* the Scene's name without its '$'             */
    char output_path[1024];  /* Output Path, between its quotes              */
} psy_render_settings;
/* Number of Scene nodes (frames) in the text. */
int psy_parse_scene_count(const char* text, int64_t len, int32_t* n_scenes);
/* Parses Scene node `scene_index` (0-based) of the text. */
int psy_parse_scene(const char* text, int64_t len, int32_t scene_index, psy_parsed_scene** out);
/* The scene description (valid until psy_parsed_scene_destroy). */
const psy_scene_desc* psy_parsed_scene_desc(const psy_parsed_scene* scene);
int psy_parsed_scene_settings(const psy_parsed_scene* scene, psy_render_settings* out);
/* What the parser skipped, one line each ("" when nothing). */
const char* psy_parsed_scene_warnings(const psy_parsed_scene* scene);
void psy_parsed_scene_destroy(psy_parsed_scene* scene);
/* DataTree::print_tree (parser/data_tree.cpp:383-398) of every top-level
 * node: writes at most cap bytes (NUL-terminated) and the full size
 * (with the NUL) to *needed. */
int psy_parse_dump_tree(const char* text, int64_t len, char* out, int64_t cap, int64_t* needed);

/* ---- film output (host memory) ------------------------------------------ */
/* Replaces: Film::scanline_image_8bbc (film/film.hpp:173-225): the per-pixel
 * mean XYZ film (as psy_render writes it, copied to the host) as 8-bit
 * sRGB with white point E (XYZ_to_sRGB_E, color/color.hpp:425-439), times
 * 255, dithered, clamped and truncated; scanline order, 3 bytes a pixel.
 * sample_counts (or NULL: every pixel sampled): pixels with 0 samples show
 * the reference's grey checkerboard.  The dither of each channel is a hash
 * of (dither_seed, pixel, channel) in [-0.5, 0.5); dither_seed 0: none. */
int psy_film_to_srgb8(const float* film_xyz, const int32_t* sample_counts, int32_t width, int32_t height,
                      uint32_t dither_seed, uint8_t* rgb8);
/* write_png_from_film's file write (renderer/renderer.cpp:24-42): an 8-bit
 * RGB PNG (uncompressed deflate). */
int psy_write_png(const char* path, const uint8_t* rgb8, int32_t width, int32_t height);
/* A float dump the reference does not have: the film as linear sRGB with
 * white point E (unclamped), PFM, little-endian. */
int psy_write_pfm(const char* path, const float* film_xyz, int32_t width, int32_t height);

/* ---- ray / hit streams (device memory) -------------------------------- */

/* World rays, WorldRay (basics/ray.hpp:162-185): 20 planes of n entries.
 * Planes 0-2 o, 3-5 d, 6-8 odx, 9-11 ody, 12-14 ddx, 15-17 ddy, 18 time,
 * 19 type (uint32 bit pattern: PSY_RAY_OCCLUSION = WorldRay::OCCLUSION). */
#define PSY_WRAY_PLANES 20
#define PSY_RAY_OCCLUSION 32u   /* WorldRay::OCCLUSION (basics/ray.hpp:177) */

/* Object-space rays with their footprint already derived, i.e. the fields
 * of Ray that WorldRay::to_ray() fills (basics/ray.hpp:24-36, :207-252):
 * 15 planes.  0-2 o, 3-5 d, 6 max_t, 7 time, 8 owx, 9 owy, 10 dwx, 11 dwy,
 * 12 fwx, 13 fwy, 14 flags (uint32 bit pattern, bit 30 = occlusion as in
 * Ray::id_and_flags).  Valid only for scenes whose patch instances are
 * untransformed; otherwise pass world rays. */
#define PSY_RAY_PLANES 15

/* Hit record, the closest-hit part of Intersection
 * (basics/intersection.hpp:22-49).  All device pointers; `aux` may be NULL.
 *   t     : hit distance (+inf on a miss)
 *   inst  : flattened leaf-instance index, 0xffffffff on a miss
 *   u, v  : surface parameters (0 for lights)
 *   aux   : patch leaf extent max_dim (offset = n*max_dim*1.74), 0 else   */
typedef struct psy_hits {
    float* t;
    uint32_t* inst;
    float* u;
    float* v;
    float* aux;
} psy_hits;

#define PSY_TRACE_OCCLUSION_ONLY 1u  /* any-hit: stop at the first hit */

/* ---- opaque handles ----------------------------------------------------- */
typedef struct psy_ctx psy_ctx;
typedef struct psy_scene psy_scene;

typedef struct psy_scene_info {
    int64_t n_leaf_instances;    /* flattened (assembly hierarchy expanded)  */
    int64_t n_objects;
    int64_t n_bvh_nodes;
    int64_t n_patches;
    int32_t dice_depth;          /* cached split depth of the dice cache     */
    int32_t bvh_stack_need;      /* worst-case traversal stack entries       */
    int64_t device_bytes;        /* scene bytes resident in HBM              */
    int64_t dice_cache_bytes;
} psy_scene_info;

/* ---- context ------------------------------------------------------------ */
/* Replaces: the per-worker Tracer construction (tracer/tracer.hpp:57-63);
 * the GPU context owns device memory, scene uploads and kernels. */
int psy_ctx_create(int device, psy_ctx** out);
int psy_ctx_destroy(psy_ctx* ctx);
/* Launch form of the trace's main pass (no counterpart in the reference,
 * whose tracer is a per-thread CPU loop).  PSY_TRACE_MODE_PLAIN (default):
 * one thread per ray.  PSY_TRACE_MODE_PERSISTENT: one grid that fills the
 * GPU; each wave refills its finished lanes from a global ray counter once
 * refill_min (1..64) of its 64 lanes are idle; blocks_per_cu <= 0 takes the
 * kernel's occupancy.  PSY_TRACE_MODE_COHERENT: persistent, and each step
 * of a wave advances only the lanes of its majority kind of work
 * (split-tree walk, or BVH node / leaf), so the wave issues one kind of
 * step at a time.  PSY_TRACE_MODE_PLAIN_COHERENT: one thread per ray, with
 * the same majority stepping (faster on instanced, moving and bicubic
 * scenes; the autotuner tries it).  Results are identical. */
#define PSY_TRACE_MODE_PLAIN 0
#define PSY_TRACE_MODE_PERSISTENT 1
#define PSY_TRACE_MODE_COHERENT 2
#define PSY_TRACE_MODE_PLAIN_COHERENT 3
int psy_ctx_set_trace_mode(psy_ctx* ctx, int mode, int refill_min, int blocks_per_cu);
/* Main-pass traversal stack in LDS (no counterpart in the reference, whose
 * BitStack and node stack live in CPU memory): 16 entries (default, 5 waves
 * per SIMD) or 32 (2 waves per SIMD; fewer rays handed to the second pass
 * on deep hierarchies).  Results are identical. */
int psy_ctx_set_trace_stack(psy_ctx* ctx, int entries);
/* The trace's second pass (no counterpart in the reference, which splits
 * patches per ray batch without a cache): rays that the main pass hands on
 * are retraced one lane each, by a narrow build (1 wave per SIMD) for short
 * queues or a wide build (2 waves per SIMD) for queues of wide_min rays or
 * more (default 4096; 0: always the wide build; UINT32_MAX: always the
 * narrow one); a ray that takes more than `budget` (>= 1, default 512)
 * split-tree node expansions there goes on to the third, wave-cooperative
 * pass.  Results do not depend on either setting. */
int psy_ctx_set_second_pass(psy_ctx* ctx, uint32_t wide_min, int32_t budget);
/* The main pass's continuations (no counterpart in the reference, whose
 * lanes are CPU threads): the rays of one 64-lane wave end after very
 * different numbers of node steps, so once fewer than `below` of a wave's
 * rays are still tracing (1..64, default 24), their traversal state goes
 * to a continuation buffer in the context's scratch and a follow-on launch
 * finishes them 64 to a wave; 0 turns it off.  Results are identical.
 * psy_trace_rays_static never continues. */
int psy_ctx_set_continuation(psy_ctx* ctx, int below);
/* Paths per wavefront chunk of psy_render (no counterpart in the reference,
 * whose buckets are sqrt(65536/spp) pixels square): whole pixels, at least
 * one; the default 2^21 paths.  Results do not depend on it. */
int psy_ctx_set_render_chunk(psy_ctx* ctx, int64_t paths);
/* Chunks of PSY_RENDER_WAVEFRONT in flight at once (no counterpart in the
 * reference): 1..4, default 2.  With k > 1, a render of more than one chunk
 * runs up to k chunks at once, each beyond the first on a stream of the
 * context's own with scratch of its own (a set of path, ray and trace
 * buffers per chunk: k times the render scratch), and the host queues each
 * chunk's next step as soon as that chunk's previous one ends, so each
 * chunk's short late bounces run beside the others' work.  The caller's
 * stream waits for all of them at the end.  While such a render runs, the
 * calling thread polls the chunks' events (one host core stays busy).
 * Same samples as with 1. */
int psy_ctx_set_render_streams(psy_ctx* ctx, int streams);
/* Register budget of PSY_RENDER_WAVEFRONT's shading kernel for the steps
 * that take a hit and sample a light (no counterpart in the reference):
 * 0 (default) lets the compiler use what the light sampling needs (2 waves
 * per SIMD); 4 holds it to 128 VGPRs for 4 waves per SIMD, spilling the
 * rest to scratch, which pays where the light tree is deep (many lights)
 * and costs elsewhere.  Same samples either way. */
int psy_ctx_set_render_shade_waves(psy_ctx* ctx, int waves);
/* Form of psy_render (no counterpart in the reference).
 * PSY_RENDER_WAVEFRONT (default): one path per thread per step; each step
 * traces the whole active stream, then shades it and compacts the paths
 * that go on (the reference's loop, integrator/path_trace_integrator.cpp:
 * 321-349).  PSY_RENDER_FUSED: persistent waves in which each lane runs one
 * path through trace, shading and its next ray, refilled from the chunk's
 * paths when refill_min (1..64) of a wave's lanes are idle; blocks_per_cu
 * <= 0 takes the kernel's occupancy.  Both give the same samples. */
#define PSY_RENDER_WAVEFRONT 0
#define PSY_RENDER_FUSED 1
/* PSY_RENDER_STREAM: the batched form.  A pool of (render chunk) path
 * slots works through the image's samples; the shading step gives a
 * finished path's slot to the next sample, so one compacted stream mixes
 * paths of many image tiles at every bounce depth and stays full until the
 * samples run out.  Stream lengths stay on the device (the trace reads
 * them there) and psy_ctx_set_stream_steps steps are captured as one HIP
 * graph, with one host synchronization per graph launch.  refill_min and
 * blocks_per_cu are ignored.  Same samples as the other forms. */
#define PSY_RENDER_STREAM 2
int psy_ctx_set_render_mode(psy_ctx* ctx, int mode, int refill_min, int blocks_per_cu);
/* Steps per captured graph of PSY_RENDER_STREAM: even, 2..64 (default 4). */
int psy_ctx_set_stream_steps(psy_ctx* ctx, int steps_per_graph);
/* Rays of the latest trace call on this context that the main pass handed
 * to the second pass (split below the dice cache, or a BVH stack deeper
 * than the main pass keeps).  Synchronizes the device; for diagnostics. */
int psy_ctx_retrace_count(psy_ctx* ctx, int64_t* out);
/* Rays of the latest trace call (psy_trace_*, or the last step of a
 * psy_render on the caller's stream) that took the second pass, and those of them that ran out of
 * its split budget and took the third, wave-cooperative pass.  Synchronizes
 * the device. */
int psy_ctx_trace_pass_counts(psy_ctx* ctx, int64_t* second, int64_t* third);
/* Pass timing (no counterpart in the reference): while enabled, each trace
 * call records HIP events on its stream around the main pass and the
 * second pass.  psy_ctx_trace_timing waits for the recorded calls and
 * returns the summed milliseconds of each pass and the number of calls
 * since the previous query, then forgets them. */
int psy_ctx_set_trace_timing(psy_ctx* ctx, int enable);
int psy_ctx_trace_timing(psy_ctx* ctx, double* main_ms, double* second_ms, int64_t* calls);
/* Device memory for callers that do not link HIP themselves (the C++ host
 * layer, include/psychopath_gpu.hpp): plain allocation on the context's
 * device and synchronous copies.  No counterpart in the reference. */
int psy_malloc(psy_ctx* ctx, size_t bytes, void** out);
int psy_free(psy_ctx* ctx, void* p);
int psy_copy_to_device(psy_ctx* ctx, void* dst, const void* src, size_t bytes);
int psy_copy_to_host(psy_ctx* ctx, void* dst, const void* src, size_t bytes);
const char* psy_last_error(void);
int psy_abi_version(void);

/* ---- scene -------------------------------------------------------------- */
/* Replaces: Scene::finalize -> Assembly::finalize -> BVH4::build /
 * LightTree::build (scene/scene.hpp:50-52, scene/assembly.hpp:206-238,
 * accel/bvh4.cpp:17-159).  Flattens the assembly hierarchy, builds the
 * instance BVH and the patch dice cache in HBM.  dice_depth (0..24) is the
 * cached split depth; < 0 picks it from the patch count: up to 24 levels
 * within 2 GiB of cache, else up to 18 within 16 GiB. */
int psy_scene_create(psy_ctx* ctx, const psy_scene_desc* desc, int dice_depth,
                     psy_scene** out);
/* A scene may be destroyed before or after the context it was created on. */
int psy_scene_destroy(psy_scene* scene);
int psy_scene_get_info(const psy_scene* scene, psy_scene_info* info);
/* InstanceID (basics/instance_id.hpp:9-40) of every flattened leaf, as the
 * reference's Tracer would report it in Intersection::id (host memory). */
int psy_scene_instance_ids(const psy_scene* scene, uint64_t* out, int64_t n);
/* Re-dices the scene: rebuilds its dice cache at dice_depth (as in
 * psy_scene_create; < 0 the memory budget's depth), beside the old one
 * (which is freed after), and waits for the device; on failure the scene is
 * unchanged.  A split deeper than the cache is diced by the second pass, so
 * every hit keeps its distance and instance; only where two micro-patches
 * give exactly the same distance can the depth change which one supplies
 * (u, v) (the first found is kept, and the second pass finds a deferred
 * subtree after the rest of the walk).  No counterpart in the reference,
 * which splits per ray batch; psy_render_autotune uses it to pick the
 * depth. */
int psy_scene_set_dice_depth(psy_scene* scene, int dice_depth);

/* Replaces: SubdivisionSurface::finalize (object/subdivision_surface.cpp:
 * 85-233), the Catmull-Clark patch extraction it leaves to OpenSubdiv.
 * The limit surface of a control mesh (boundary edges creased, boundary
 * vertices smooth) as bicubic Bezier patches: 48 floats each, 16 control
 * points in rows along v, in PSY_OBJ_BICUBIC's layout.  Faces are refined
 * adaptively around extraordinary vertices down to level 5; each face left
 * irregular there gets an end cap matching the limit position and tangents
 * at its corners.  The mesh must be manifold with consistently oriented
 * faces of 3 or more distinct vertices.  Host only, no device is used.
 * patches == NULL: only *n_patches is set; else max_patches must be at
 * least the count. */
int psy_subdivision_patches(const float* verts, int32_t n_verts, const int32_t* face_vert_counts,
                            int32_t n_faces, const int32_t* face_vert_indices, float* patches,
                            int64_t max_patches, int64_t* n_patches);

/* ---- tracing ------------------------------------------------------------ */
/* Replaces: uint32_t Tracer::trace(const WorldRay*, const WorldRay*,
 *           Intersection*, Intersection*)  (tracer/tracer.hpp:76,
 *           tracer/tracer.cpp:38-121).  Closest hit per ray (or any hit
 *           with PSY_TRACE_OCCLUSION_ONLY / per-ray occlusion type).
 * Work is queued on `stream` and the call returns without waiting.  The
 * trace, render and adjoint calls of one context share its scratch
 * memory (the second- and third-pass queues): a call on a stream other
 * than the previous call's is made to wait on the device for that call,
 * so calls on different streams of one context run one after the other.
 * (Calls queued on a stream under graph capture are not ordered this
 * way.)  The caller orders its own ray and hit buffers.  One context is
 * not to be used from several host threads at once. */
int psy_trace_world(psy_ctx* ctx, const psy_scene* scene,
                    const float* world_rays, int64_t n,
                    const psy_hits* hits, uint32_t flags, void* stream);
/* Same, for rays whose footprint was already derived (WorldRay::to_ray,
 * basics/ray.hpp:310-323); the 56-byte-per-ray form used for streams. */
int psy_trace_rays(psy_ctx* ctx, const psy_scene* scene,
                   const float* rays, int64_t n,
                   const psy_hits* hits, uint32_t flags, void* stream);
/* No counterpart in the reference: psy_trace_rays for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * results as psy_trace_rays, but its second- and third-pass queues and
 * counters live in a fixed array of the library's own on each GPU (48 MiB
 * + 256 B, held until the process ends) instead of context memory, and it
 * calls no hipMalloc, hipFree or other allocating routine.  At most
 * PSY_STATIC_MAX_RAYS rays per call (PSY_ERR_INVALID above).  Every call on
 * one GPU, from any context, shares that array: calls that overlap in time
 * on one GPU race, and the caller must order them (one stream, or events).
 * It does not record pass timing, and psy_ctx_trace_pass_counts does not
 * report its calls. */
#define PSY_STATIC_MAX_RAYS 2097152
int psy_trace_rays_static(psy_ctx* ctx, const psy_scene* scene,
                          const float* rays, int64_t n,
                          const psy_hits* hits, uint32_t flags, void* stream);
/* The rest of the reference's Intersection (basics/intersection.hpp:22-49)
 * for the closest hits of a psy_trace_world call, as its tracer fills it
 * at the accepted hit (tracer/tracer.cpp:241-250 with
 * object/sphere.cpp:120-189, object/patch_utils.hpp:94-131,
 * lights/sphere_light.hpp:159-246, lights/rectangle_light.hpp:110-147 and
 * SurfaceShader::shade, shading/surface_shader.hpp:28-53), for a caller that
 * shades the hits: `out` gets PSY_ISECT_PLANES planes of n floats (device):
 *   0-2 p, 3-5 n, 6-8 dpdu, 9-11 dpdv, 12-14 dndu, 15-17 dndv  (geo, in
 *        the hit object's space; 0 for lights' derivatives),
 *   18-20 offset, 21-36 space (Transform: world -> object, row-major
 *        matrix at the ray's time; identity when untransformed),
 *   37 light_pdf (light hits; 9999, the reference's default, elsewhere),
 *   38 backfacing (uint32 0/1),
 *   39 closure kind (uint32: PSY_SHADER_LAMBERT or PSY_SHADER_EMIT, 0 on
 *        a miss, whose other planes hold Intersection's defaults), 40-42 the closure's colour (LambertClosure / EmitClosure
 *        colour, linear RGB; a light's is its emitted colour per area
 *        times its emission scale), 43 shader index (int32, -1 for lights
 *        and for the magenta emitter of a surface without a shader).
 * world_rays and hits are those of the trace (hits.aux must be set). */
#define PSY_ISECT_PLANES 44
int psy_intersections(psy_ctx* ctx, const psy_scene* scene, const float* world_rays, int64_t n,
                      const psy_hits* hits, float* out, void* stream);
/* WorldRay::to_ray over a stream: world rays -> object rays (planes as
 * PSY_RAY_PLANES).  Replaces tracer/tracer.cpp:47-51. */
int psy_world_to_rays(psy_ctx* ctx, const float* world_rays, int64_t n,
                      float* rays, void* stream);

/* ---- rendering ---------------------------------------------------------- */
typedef struct psy_render_params {
    int32_t width, height;  /* image size (Film, film/film.hpp:40-60)            */
    int32_t spp;            /* samples per pixel                                */
    uint32_t seed;          /* ImageSampler seed (parser default 1)             */
    int32_t path_length;    /* bounces; the reference integrator uses 4         */
    uint32_t flags;         /* PSY_RENDER_* below                               */
} psy_render_params;
/* Sample every path's dimensions from Sobol::sample (sampling/sobol.hpp:
 * 40-53) over the matrices set by psy_sobol_set_matrices, instead of the
 * reference's Halton Sampler (sampling/image_sampler.hpp:16-38); no
 * counterpart in the reference, whose integrator samples with Halton only.
 * A path keeps the Sampler's index (hash of pixel and seed, plus the
 * sample number) and dimension order; dimension d takes Sobol dimension
 * d at that index, unscrambled, rounded to float and kept below 1. */
#define PSY_RENDER_SOBOL 1u

/* Replaces: PathTraceIntegrator(scene, film, spp, spp, 0, seed, threads)
 * + integrate() (integrator/path_trace_integrator.hpp:103-120,
 * integrator/path_trace_integrator.cpp:282-384): every pixel sample of the
 * image is one path (init_path / next_ray_for_path / update_path with light
 * tree + MIS next-event estimation, Lambert and emit closures, hero-
 * wavelength spectral sampling) traced as a wavefront on the device.
 *   film_xyz   : device, width*height*3 floats, or NULL: per pixel the mean
 *                of its finite sample values (Film::add_sample's sum over
 *                count, film/film.hpp:135-153), CIE XYZ.
 *   sample_xyz : device, width*height*spp*3 floats, or NULL: every sample's
 *                film contribution Color_XYZ(col) * XYZ_NORM_FAC, ordered
 *                (y, x, sample).
 *   rays_traced: host, or NULL: camera + bounce + shadow rays traced
 *                (Global::Stats::rays_shot).
 * Colours other than grey (and surfaces without a shader, which emit
 * magenta) need the spectrum grid (psy_spectrum_set_grid;
 * PSY_ERR_UNSUPPORTED otherwise).  Synchronizes `stream` once per wavefront
 * step. */
int psy_render(psy_ctx* ctx, const psy_scene* scene, const psy_render_params* params,
               float* film_xyz, float* sample_xyz, int64_t* rays_traced, void* stream);

/* Adaptive sampling: PathTraceIntegrator(scene, film, spp, spp_max,
 * variance_max, seed, thread_count) + integrate() (integrator/
 * path_trace_integrator.cpp:28-90, 282-384) with its Film statistics
 * (film/film.hpp:135-165).  The image is cut into the reference's square
 * buckets (psy_bucket_size: sqrt(65536 / spp), at most
 * sqrt(width * height / (4 thread_count)), at least 1; row-major bucket
 * order); a bucket takes passes of spp samples per pixel while its largest
 * X variance estimate exceeds variance_max and fewer than spp_max (at least
 * spp) samples were taken.  Each pass of the reference draws the same
 * samples (Sampler(x, y, s, seed), s = 0 .. spp - 1), so they are rendered
 * once and each pass adds them to the film statistics again: the film and
 * pass counts are the reference's, rays_traced is one pass's.
 *   film_xyz     : device, width*height*3, or NULL: the film's per-pixel
 *                  sum / count (CIE XYZ, as psy_render's film)
 *   variance_xyz : device, width*height*3, or NULL: variance_estimate
 *                  (+inf below 2 samples)
 *   passes       : device, one int32 per bucket, or NULL */
int32_t psy_bucket_size(int32_t width, int32_t height, int32_t spp, int32_t thread_count);
int psy_render_adaptive(psy_ctx* ctx, const psy_scene* scene, const psy_render_params* params,
                        int32_t spp_max, float variance_max, int32_t thread_count,
                        float* film_xyz, float* variance_xyz, int32_t* passes,
                        int64_t* rays_traced, void* stream);
/* The film part alone, over per-sample values (sample_xyz: device, the
 * layout psy_render writes) and a given bucket size. */
int psy_film_adaptive(psy_ctx* ctx, const float* sample_xyz, int32_t width, int32_t height, int32_t spp,
                      int32_t spp_max, float variance_max, int32_t bucket_size,
                      float* film_xyz, float* variance_xyz, int32_t* passes, void* stream);

/* The render autotuner (no counterpart in the reference): times psy_render
 * on every pixel at an spp that keeps to max_samples samples (<= 0: 2^23),
 * by coordinate descent over the render form (PSY_RENDER_STREAM or
 * PSY_RENDER_WAVEFRONT) with its paths in flight (render chunk,
 * 2^19..2^23), the main pass's launch form (psy_ctx_set_trace_mode), its
 * LDS stack (psy_ctx_set_trace_stack), the stream form's steps per graph
 * (psy_ctx_set_stream_steps) and the scene's dicing depth
 * (psy_scene_set_dice_depth: the cache 4 and 2 levels shallower and 2
 * deeper, within 16 GiB) and, for the wavefront form, 1 to 4 chunks in
 * flight (psy_ctx_set_render_streams; k are timed on a sample of at least k
 * chunks, up to the full spp), the whole image as 2 to 4 chunks all in
 * flight and the shading kernel's register budget
 * (psy_ctx_set_render_shade_waves), and leaves the fastest setting on the
 * context
 * and the scene.  `out` (or NULL) receives it. */
typedef struct psy_render_tuning {
    int32_t mode;             /* PSY_RENDER_STREAM or PSY_RENDER_WAVEFRONT */
    int32_t trace_mode;       /* PSY_TRACE_MODE_*                        */
    int32_t refill_min;
    int32_t stack;            /* main-pass stack entries                 */
    int32_t steps_per_graph;
    int32_t trials;           /* renders timed                           */
    int64_t pool;             /* paths in flight                         */
    int32_t dice_depth;       /* the scene's dice-cache depth             */
    int32_t streams;          /* wavefront chunks in flight (1..4)       */
    int32_t shade_waves;      /* psy_ctx_set_render_shade_waves          */
    double mrays_per_s;       /* of the best trial                       */
} psy_render_tuning;
int psy_render_autotune(psy_ctx* ctx, psy_scene* scene, const psy_render_params* params,
                        int64_t max_samples, psy_render_tuning* out);

/* ---- differentiation ---------------------------------------------------- */
/* Differentiable scene parameters (no counterpart in the reference, which
 * renders forward only):
 *   albedo[k]         grey albedo of surface shader k (the scene's shader
 *                     table; a Lambert shader's colour is (a, a, a); entries
 *                     of other shaders are ignored and get zero gradient);
 *   emission_scale[j] multiplier of the emission of the j-th light object,
 *                     counting SphereLight / RectangleLight objects in scene
 *                     order (assembly 0's objects first, then assembly 1's,
 *                     ...); 1 after psy_scene_create.
 * The light trees keep the energies of the scene as created: light
 * selection probabilities do not change with the scales, so renders stay
 * unbiased and every sample is a polynomial in the parameters.
 * psy_scene_set_params copies host arrays synchronously (NULL keeps that
 * set); do not call it while work that reads the scene is in flight. */
int psy_scene_param_counts(const psy_scene* scene, int64_t* n_albedo, int64_t* n_emission);
int psy_scene_get_params(const psy_scene* scene, float* albedo, float* emission_scale);
int psy_scene_set_params(psy_scene* scene, const float* albedo, int64_t n_albedo,
                         const float* emission_scale, int64_t n_emission);
/* Path-replay adjoint of psy_render: the gradients of a scalar image loss
 * with respect to the parameters above, given d_film_xyz = dLoss/dfilm
 * (device, width*height*3 floats, the layout of psy_render's film_xyz).
 * Every path is re-run from its sample index with the same sampler as the
 * forward render (same seed, spp, path_length: identical paths, nothing
 * stored between the calls), and each radiance contribution adds its
 * derivatives.  d_albedo (n_albedo floats) and d_emission (n_emission
 * floats) are device memory and are overwritten.  path_length <= 8.
 * rays_traced as in psy_render.  Synchronizes `stream` once per step. */
int psy_adjoint(psy_ctx* ctx, const psy_scene* scene, const psy_render_params* params,
                const float* d_film_xyz, float* d_albedo, float* d_emission,
                int64_t* rays_traced, void* stream);

/* ---- sampling ----------------------------------------------------------- */
/* Replaces: Halton::sample(dim, index) (sampling/halton.cpp:2154-2671). */
int psy_halton(psy_ctx* ctx, uint32_t dim, const uint32_t* index, int64_t n,
               float* out, void* stream);
/* Sobol' generator matrices for psy_sobol: num_dims * 52 columns of 52
 * bits, dimension-major, the layout of Sobol::Matrices::matrices
 * (sampling/sobol.hpp:28-32, sampling/sobol.cpp:36; the reference's table
 * has 1024 dimensions).  Host memory, copied to the device. */
int psy_sobol_set_matrices(psy_ctx* ctx, const uint64_t* matrices, uint32_t num_dims);
/* Replaces: Sobol::sample(dimension, index, scramble)
 * (sampling/sobol.hpp:40-53) over a stream of indices (device memory):
 * out[i] = (scramble_low52 ^ XOR of column (dim*52 + b) for each set bit b
 * of index[i]) * 2^-52.  As in the reference's flat indexing, index bits
 * 52 and up select the next dimension's columns; columns past the table
 * read as zero. */
int psy_sobol(psy_ctx* ctx, uint32_t dim, const uint64_t* index, int64_t n,
              uint64_t scramble, double* out, void* stream);
/* ---- colour ------------------------------------------------------------- */
/* The spectrum grid of tristimulus upsampling (Hanika et al.;
 * color/spectrum_grid.h, color/spectra_xyz_5nm_380_780_0.97.h), handed over
 * by the caller as data (the library embeds none):
 *   cells        : grid_width * grid_height cells, row-major, 8 ints each
 *                  (inside, num_points, idx[6]) -- spectrum_grid_cell_t,
 *                  i.e. (const int32_t*)spectrum_grid;
 *   points       : n_points data points of 4 + n_samples floats (xystar[2],
 *                  uv[2], spectrum[n_samples]) -- spectrum_data_point_t,
 *                  i.e. (const float*)spectrum_data_points;
 *   sample_min/max, n_samples: the spectra's wavelength range and knots
 *                  (spectrum_sample_min/max, spectrum_num_samples);
 *   mat_xy_to_uv : the 3x2 xy -> uv matrix (spectrum_mat_xy_to_uv).
 * Host memory, validated and copied to the device.  With a grid set,
 * psy_render upsamples every colour through it, as the reference does;
 * without one (or after a call with cells == NULL) it takes grey colours
 * only, on the grid's flat white-point spectrum. */
int psy_spectrum_set_grid(psy_ctx* ctx, const int32_t* cells, int32_t grid_width, int32_t grid_height,
                          const float* points, int32_t n_points, int32_t n_samples, float sample_min,
                          float sample_max, const float* mat_xy_to_uv);
/* Replaces: Color_to_SpectralSample(col, hero) (color/color.hpp:490-511,
 * color/spectrum_grid.h:21-135) over a stream (device memory): rgb 3*n,
 * hero_wavelength n (nm, in [380, 700]), out 4*n = the spectral samples at
 * the hero wavelength and its 3 rotations.  Without a grid: the flat
 * spectrum (exact for grey colours only). */
int psy_color_to_spectrum(psy_ctx* ctx, const float* rgb, const float* hero_wavelength, int64_t n, float* out,
                          void* stream);

/* Replaces: Sampler(x,y,s,seed) + Sampler::get_sample(dim)
 * (sampling/image_sampler.hpp:16-38): out[i*n_dims + k] = sample of dim
 * dim0+k for pixel sample i of the tile (x0..x0+w, y0..y0+h, spp). */
int psy_pixel_samples(psy_ctx* ctx, int32_t x0, int32_t y0, int32_t w, int32_t h,
                      int32_t spp, uint32_t seed, uint32_t dim0, int32_t n_dims,
                      float* out, void* stream);

/* ---- stream generation (integrator steps) ------------------------------ */
/* Camera rays: PathTraceIntegrator::init_path + next_ray_for_path step 0
 * (integrator/path_trace_integrator.cpp:91-125, basics/camera.hpp:72-104)
 * for every pixel sample of the image, ordered (y, x, s) row-major.
 * Writes world rays (PSY_WRAY_PLANES) and the hero wavelength per ray. */
int psy_camera_rays(psy_ctx* ctx, const psy_scene* scene, int32_t width,
                    int32_t height, int32_t spp, uint32_t seed,
                    float* world_rays, float* wavelength, void* stream);
/* Lambert bounce rays from the closest hits of `in_rays`
 * (integrator/path_trace_integrator.cpp:183-209,
 *  shading/surface_closure.hpp:233-331).  Sample dimensions
 * dim_u, dim_v of each path's Sampler; rays that missed (or hit a light)
 * are written with d = 0 and flagged in `valid` (uint32 0/1).  pix_*:
 * the generating pixel sample (same ordering as psy_camera_rays). */
int psy_bounce_rays(psy_ctx* ctx, const psy_scene* scene,
                    const float* in_world_rays, const psy_hits* in_hits,
                    int64_t n, int32_t width, int32_t spp, uint32_t seed,
                    uint32_t dim_u, uint32_t dim_v,
                    float* out_world_rays, uint32_t* valid, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PSY_GPU_H */
