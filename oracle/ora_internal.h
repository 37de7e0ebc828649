/*
 * ora_internal.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 * Internal types of the CPU restatement of the reference's hot path.
 */
#ifndef ORA_INTERNAL_H
#define ORA_INTERNAL_H

#include "ora_core.h"
#include "psy_oracle.h"
#include "../include/psy_gpu.h"

/* Ray (basics/ray.hpp:24-155) */
typedef struct {
    v3 o; float max_t;
    v3 d; float time;
    v3 d_inv;
    float owx, owy, dwx, dwy, fwx, fwy;
    uint32_t id_and_flags;
    uint64_t trav_stack;
} ora_ray;

#define RAY_OCCL (1u << 30)
#define RAY_DONE (1u << 31)
#define RAY_IDMASK (0xffffffffu >> 2)

static inline int ray_is_occl(const ora_ray* r) { return (r->id_and_flags & RAY_OCCL) != 0; }
static inline int ray_is_done(const ora_ray* r) { return (r->id_and_flags & RAY_DONE) != 0; }
static inline uint32_t ray_id(const ora_ray* r) { return r->id_and_flags & RAY_IDMASK; }

/* SubdivisionSurface::Node (object/subdivision_surface.hpp:16-20) with one
 * time sample's bounds; patch >= 0 marks a leaf */
typedef struct {
    bbox b;
    int child[2];
    int patch;
} ora_sd_node;

typedef struct ora_object {
    int kind, ts;
    float escale;     /* emission scale of a light object (1); not in the light tree */
    float* data;      /* ts * floats-per-kind; a subdivision surface's control points */
    bbox* bounds;     /* ts */
    int n_bounds;
    /* PSY_OBJ_SUBDIVISION: control point count, Bicubic patches, patch BVH */
    int n_verts;
    int n_patches;
    struct ora_object* patches;
    ora_sd_node* sd_nodes;
    int sd_root;
} ora_object;

/* BVH::Node (accel/bvh.hpp:46-55) */
typedef struct {
    size_t bbox_index;
    size_t child_or_data;
    size_t parent_index;
    uint16_t ts, flags;
} ora_bvh_node;

/* BVH4::Node (accel/bvh4.hpp:39-71); data_index shares storage with bounds
 * in the reference, kept separately here. */
typedef struct {
    bbox4 bounds;
    size_t data_index;
    size_t child[3];
    uint32_t ts;
} ora_bvh4_node;

/* LightTree::Node (accel/light_tree.hpp) */
typedef struct {
    bbox* bounds; int n_bounds;
    float energy;
    size_t index1, index2;
    int is_leaf;
    size_t instance_index;
} ora_ltree_node;

typedef struct ora_assembly {
    struct ora_scene* scene;
    int n_inst; psy_instance_desc* inst;
    int n_xf; mat4* xf;
    int n_obj; ora_object* obj;
    int n_sub; struct ora_assembly** sub;
    int finalized;
    /* BVH4 */
    ora_bvh4_node* nodes; size_t n_nodes;
    bbox* bounds; int n_bounds;            /* BVH4::_bounds */
    /* LightTree */
    ora_ltree_node* lnodes; size_t n_lnodes;
    bbox* lbounds; int n_lbounds;
    float total_energy;
    size_t total_lights;
} ora_assembly;

typedef struct ora_scene {
    int n_asm; ora_assembly* asms;
    int n_shaders; psy_shader_desc* shaders;
    /* camera */
    int n_cam_xf; mat4* cam_xf;
    int n_tfov; float* tfovs;
    int n_ap; float* aperture;
    int n_fd; float* focus;
    float background[3];
    float dice_rate, min_upoly;
    uint32_t flags;  /* PSY_SCENE_* */
} ora_scene;

/* element_id_bits (scene/assembly.hpp:245-254) */
uint32_t ora_intlog2(uint32_t v);
uint32_t ora_upper_pow2(uint32_t v);
static inline int asm_id_bits(const ora_assembly* a) {
    return (int)ora_intlog2(ora_upper_pow2((uint32_t)a->n_inst));
}

bbox ora_bbox_inverse_transformed(bbox b, const mat4* xf);
bbox ora_instance_bounds_at(const ora_assembly* a, float t, size_t i);

/* world ray -> ray (basics/ray.hpp:207-323) */
void ora_update_ray(const ora_wray* w, ora_ray* r);
void ora_update_ray_xf(const ora_wray* w, ora_ray* r, const mat4* xf);
float ora_ray_min_width(const ora_ray* r, float tnear, float tfar);

/* floats per object kind */
int ora_kind_floats(int kind);

#endif
