// psy_internal.hpp -- host-side state shared by the translation units of
// libpsy_gpu.so (psy_capi.cpp: context, scene upload, tracing;
// psy_render.cpp: light trees and the path-tracing integrator).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/psy_gpu.h"
#include "psy_device.hpp"

namespace psy {

// Records `msg` for psy_last_error() and returns `code`.
int fail(int code, const std::string& msg);
int floats_per_kind(int kind);

#define HIP_OK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::psy::fail(PSY_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Owning device allocation.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int alloc(size_t b) {
        release();
        bytes = b;
        if (b == 0) return 0;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            p = nullptr;
            bytes = 0;
            return fail(PSY_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        return 0;
    }
    // grows (never shrinks) to at least b bytes; contents are not kept
    int reserve(size_t b) { return b <= bytes && p ? 0 : alloc(b); }
    int upload(const void* src, size_t b) {
        int rc = alloc(b);
        if (rc) return rc;
        if (b) HIP_OK(hipMemcpy(p, src, b, hipMemcpyHostToDevice));
        return 0;
    }
    template <class T>
    int upload(const std::vector<T>& v) { return upload(v.data(), v.size() * sizeof(T)); }
};

// Words of a context's continuation buffer for streams of up to cap rays:
// room for every continued ray of a call that continues waves with fewer
// than 24 live lanes and the 16-entry stack (48 words an entry, 0.36
// entries per ray), or fewer than 16 and the 32-entry stack (80 words,
// 0.23 entries per ray); a call that could need more runs without
// continuations.
inline int64_t cont_area_words(int64_t cap) { return 19 * cap + 10240; }

}  // namespace psy

struct psy_ctx {
    int device = 0;
    psy::DBuf halton_dims, halton_perm;
    psy::DBuf sobol_m;         // generator matrices, 52 columns per dimension
    uint32_t sobol_dims = 0;
    psy::DBuf spec_cells, spec_points;  // the spectrum grid (psy_spectrum_set_grid)
    psy::DevSpectrum spec{};            // cells == NULL: none set
    // trace scratch: queue of rays for the second pass + its counters
    psy::DBuf queue, counters;
    int64_t queue_cap = 0;
    // Calls on one context share that scratch.  Each call records
    // scratch_done on its stream when its work is queued; a call on another
    // stream first makes its stream wait for that event, so calls made on
    // different streams run one after the other (scratch_order,
    // scratch_release in psy_capi.cpp).
    hipEvent_t scratch_done = nullptr;
    hipStream_t scratch_stream = nullptr;
    bool scratch_recorded = false;
    // the allocation-free trace's scratch on this device (static_scratch_address)
    uint32_t* static_scratch = nullptr;
    // second pass: queue length from which the wide build walks, node budget
    uint32_t deep_wide_min = psy::DEEP_WIDE_MIN;
    int32_t deep_budget = psy::DEEP_BUDGET;
    // main-pass launch form, per ray form (0: object rays, 1: world rays)
    int n_cu = 0;
    int occupancy[2] = {0, 0};    // 16-entry stack
    int occupancy32[2] = {0, 0};  // 32-entry stack
    int trace_stack = 16;         // main-pass LDS stack entries
    int cont_below = 24;          // main pass: continue a wave's rays once fewer lanes trace (0: never)
    int mode = PSY_TRACE_MODE_PLAIN;
    int refill_min = 16;
    int blocks_per_cu = 0;  // <= 0: from occupancy
    // pass timing: event triples of the trace calls since the last query
    bool timing = false;
    std::vector<void*> timing_events;
    psy::TraceTuning tuning(bool world, bool timed = true) {
        psy::TraceTuning t;
        t.persistent = mode == PSY_TRACE_MODE_PERSISTENT || mode == PSY_TRACE_MODE_COHERENT;
        t.coherent = mode == PSY_TRACE_MODE_COHERENT;
        t.plain_coherent = mode == PSY_TRACE_MODE_PLAIN_COHERENT;
        t.persist_blocks = n_cu * (blocks_per_cu > 0 ? blocks_per_cu : occupancy[world ? 1 : 0]);
        t.persist_blocks32 = n_cu * (blocks_per_cu > 0 ? blocks_per_cu : occupancy32[world ? 1 : 0]);
        t.refill_min = refill_min;
        t.stack = trace_stack;
        t.wide_min = deep_wide_min;
        t.deep_budget = deep_budget;
        if (queue.p) {
            t.cont = (uint32_t*)queue.p + 6 * queue_cap;
            t.cont_words = psy::cont_area_words(queue_cap);
            t.cont_below = cont_below;
        }
        if (timing && timed) {
            for (int k = 0; k < 3; ++k) {
                hipEvent_t e = nullptr;
                if (hipEventCreate(&e) != hipSuccess) {
                    for (int j = 0; j < k; ++j) {
                        (void)hipEventDestroy((hipEvent_t)t.ev[j]);
                        t.ev[j] = nullptr;
                    }
                    break;
                }
                t.ev[k] = e;
            }
            if (t.ev[0])
                for (int k = 0; k < 3; ++k) timing_events.push_back(t.ev[k]);
        }
        return t;
    }
    ~psy_ctx() {
        for (void* e : timing_events) (void)hipEventDestroy((hipEvent_t)e);
        if (scratch_done) (void)hipEventDestroy(scratch_done);
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
        for (hipStream_t x : slot_stream)
            if (x) (void)hipStreamDestroy(x);
        if (render_join) (void)hipEventDestroy(render_join);
        for (hipEvent_t e : render_done)
            if (e) (void)hipEventDestroy(e);
        if (host_counts) (void)hipHostFree(host_counts);
    }
    // integrator scratch (path states, ray / hit streams, active lists)
    psy::DBuf render_scratch;
    psy::DBuf film_state, film_samples;  // psy_film_adaptive / psy_render_adaptive
    int64_t render_chunk = 1 << 21;  // paths per chunk
    int render_mode = PSY_RENDER_WAVEFRONT;
    int render_refill = 16;          // fused mode: refill a wave at this many idle lanes
    int render_blocks_per_cu = 0;    // fused mode: <= 0 from occupancy
    int render_occupancy = 0;
    int render_graph_iters = 4;      // stream mode: steps per captured graph
    int render_shade_waves = 0;      // wavefront hit-step shading: 0 compiler's registers, 4 held to 4 waves
    hipStream_t capture_stream = nullptr;  // for graph capture (never runs work)
    // The wavefront form's chunks in flight (1..MAX_RENDER_SLOTS).  Chunk
    // k >= 1 runs on slot_stream[k - 1] with scratch, trace queue and
    // counters of its own; render_join orders those streams after the
    // caller's at the start of a render and the caller's after them at the
    // end; render_done[k] is recorded after each of chunk k's steps.
    static constexpr int MAX_RENDER_SLOTS = 4;
    int render_streams = 2;
    hipStream_t slot_stream[MAX_RENDER_SLOTS - 1] = {};
    hipEvent_t render_join = nullptr;
    hipEvent_t render_done[MAX_RENDER_SLOTS] = {};
    psy::DBuf slot_scratch[MAX_RENDER_SLOTS - 1], slot_queue[MAX_RENDER_SLOTS - 1],
        slot_counters[MAX_RENDER_SLOTS - 1];
    int64_t slot_queue_cap[MAX_RENDER_SLOTS - 1] = {};
    // pinned words the wavefront form reads each chunk's next active count
    // into (word 16 k for chunk k), so that one chunk's read back does not
    // wait for the other chunks' queued work
    uint32_t* host_counts = nullptr;
};

struct psy_scene {
    psy_ctx* ctx = nullptr;
    int device = 0;  // the context's device (the scene may outlive its context)
    psy::DevScene dev{};
    psy::DevCamera cam{};
    psy::DBuf bvh, bvh1, inst, obj, odata, xf, dice, shaders;
    // differentiable parameters: shader albedos (host copy of the table),
    // emission scales of the light objects
    std::vector<psy::DevShader> host_shaders;
    // leaves, objects and the patch objects in dice-cache slot order, kept
    // to re-dice the scene at another depth (psy_scene_set_dice_depth)
    std::vector<psy::DevInstance> host_insts;
    std::vector<psy::DevObject> host_objs;
    std::vector<int32_t> patch_ids;
    std::vector<int32_t> light_objects;
    std::vector<float> escale;
    psy::DBuf escale_dev, lparam_dev;
    int stack_need = 1;
    bool transformed_patches = false;
    std::vector<uint64_t> ids;
    psy_scene_info info{};
    // integrator data: per-assembly light trees and the assembly hierarchy
    psy::DevLights lights{};
    psy::DBuf l_nodes, l_bounds, l_asms, l_inst, l_xf, l_sub;
    float background[3] = {0.0f, 0.0f, 0.0f};
    bool chromatic = false;  // some colour is not grey (no spectral table)
};

namespace psy {

// Everything psy_scene_create derives from a scene description, in host
// memory, before any device call: the flattened leaves, the BVH4, objects,
// transforms, shaders, camera, the dice-cache layout (the cache itself is
// built on the device) and the light trees.
struct HostScene {
    std::vector<BvhNode> bvh;
    std::vector<BvhNode> bvh1;  // child box change to time 1 when some leaf moves
    int32_t root = CHILD_EMPTY;
    int stack_need = 1;
    std::vector<DevInstance> insts;
    std::vector<DevObject> objs;
    std::vector<float> odata;
    std::vector<M44> xf;
    std::vector<DevShader> shaders;
    std::vector<float> escale;         // per object: emission scale (1)
    std::vector<int32_t> light_param;  // per object: index among light objects, -1 else
    std::vector<int32_t> light_objects;
    std::vector<int32_t> patch_ids;  // objects with a dice-cache slot, in slot order
    int levels = 0;                  // cached split depth
    int64_t per = 2;                 // nodes per slot (1-based heap, entry 0 unused)
    std::vector<uint64_t> ids;       // InstanceID per leaf
    bool transformed_patches = false;
    bool chromatic = false;          // some colour is not grey (or a surface has no shader)
    float dice_rate = 0.25f, min_upoly = 1e-5f;
    DevCamera cam{};
    float background[3] = {0.0f, 0.0f, 0.0f};
    // light trees of every assembly and the assembly hierarchy
    std::vector<DevLNode> l_nodes;
    std::vector<float4> l_bounds;
    std::vector<DevAsm> l_asms;
    std::vector<DevAsmInst> l_inst;
    std::vector<M44> l_xf;
    std::vector<int32_t> l_sub;
};

// Catmull-Clark limit surface of a control mesh as bicubic Bezier patches,
// 48 floats each (psy_subdiv.cpp); err holds the reason of a nonzero status.
int subdivide_patches(const float* verts, int32_t n_verts, const int32_t* counts, int32_t n_faces,
                      const int32_t* indices, std::vector<float>& out, std::string& err);
// Dice-cache depth for n patch objects: dice_depth, or below 0 the memory
// budget's (up to 24 levels within 2 GiB, else up to 18 within 16 GiB);
// per = nodes per patch slot.
int dice_layout(size_t n_patches, int dice_depth, int& levels, int64_t& per, std::string& err);
// Rebuilds a scene's dice cache at another depth (psy_capi.cpp).
int rebuild_dice(psy_scene* sc, int dice_depth);
// Host part of psy_scene_create (psy_capi.cpp); no device calls.
int prepare_scene(const psy_scene_desc* d, int dice_depth, HostScene& H);
// Light trees of every assembly (psy_render.cpp); obj_base: first global
// object of each assembly.
int prepare_lights(const psy_scene_desc* d, const std::vector<int32_t>& obj_base, HostScene& H);
int upload_lights(const HostScene& H, psy_scene* sc);
// Faure-permuted Halton tables, 256 dimensions (sampling/halton.py:27-87).
void make_halton_tables(std::vector<HaltonDim>& dims, std::vector<uint16_t>& perm);
// Grows the context's trace scratch (second-pass queue) to n rays.
int ensure_queue(psy_ctx* ctx, int64_t n);
// The same for a given queue / counters pair and its capacity in rays.
int ensure_trace_scratch(DBuf& queue, DBuf& counters, int64_t& queue_cap, int64_t n);
// Orders a call that uses the context's trace scratch on stream st after
// the previous such call (made on any stream): scratch_order before the
// call's first launch, scratch_release after its last.  Streams under graph
// capture are left alone.
int scratch_order(psy_ctx* ctx, hipStream_t st);
int scratch_release(psy_ctx* ctx, hipStream_t st);
}  // namespace psy
