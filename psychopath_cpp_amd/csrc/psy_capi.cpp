// psy_capi.cpp -- host side of libpsy_gpu.so: the C ABI declared in
// include/psy_gpu.h, scene flattening, the instance BVH4 builder, uploads to
// HBM, Halton table generation.
//
// Scene preparation replaces Scene::finalize / Assembly::finalize /
// BVH4::build / LightTree::build (scene/assembly.hpp:206-238,
// accel/bvh4.cpp:17-159).  The reference keeps a per-assembly BVH4 and
// re-derives rays per instance while recursing (tracer/tracer.cpp:125-219).
// Here the assembly tree is flattened once: every leaf object instance gets
// the reference's merged transform list (math/transform.hpp:241-272) and
// InstanceID (basics/instance_id.hpp:9-40), and one BVH4 is built over all
// leaves.  A ray at a leaf is re-derived from the world ray with the merged
// transform exactly as the reference does, so closest hits are unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/psy_gpu.h"
#include "psy_internal.hpp"

using namespace psy;

namespace psy {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int floats_per_kind(int kind) {
    switch (kind) {
        case PSY_OBJ_SPHERE: return 4;
        case PSY_OBJ_BILINEAR: return 12;
        case PSY_OBJ_BICUBIC: return 48;
        case PSY_OBJ_SPHERE_LIGHT: return 7;
        case PSY_OBJ_RECTANGLE_LIGHT: return 5;
        default: return 0;
    }
}
}  // namespace psy

namespace {

// ---- Halton tables (sampling/halton.py:27-87 generation rule) --------------
}  // namespace

namespace psy {
void make_halton_tables(std::vector<HaltonDim>& t_dims, std::vector<uint16_t>& t_perm) {
    t_dims.clear();
    t_perm.clear();
    const int D = 256;
    std::vector<uint32_t> primes;
    uint32_t cand = 1;
    while ((int)primes.size() < D) {
        ++cand;
        bool prime = true;
        for (uint32_t k = 2; k * k <= cand; ++k)
            if (cand % k == 0) { prime = false; break; }
        if (prime) primes.push_back(cand);
    }
    const uint32_t maxb = primes.back();
    std::vector<std::vector<uint16_t>> faure(maxb + 1);
    for (uint32_t b = 0; b <= maxb; ++b) {
        if (b < 2) { faure[b] = {0}; continue; }
        if (b == 2) { faure[b] = {0, 1}; continue; }
        faure[b].resize(b);
        if (b & 1) {
            const uint32_t c = (b - 1) / 2;
            for (uint32_t i = 0; i < b; ++i) {
                if (i == c) { faure[b][i] = (uint16_t)c; continue; }
                const uint32_t f = faure[b - 1][i - (i > c ? 1 : 0)];
                faure[b][i] = (uint16_t)(f + (f >= c ? 1 : 0));
            }
        } else {
            const uint32_t c = b / 2;
            for (uint32_t i = 0; i < b; ++i)
                faure[b][i] = (uint16_t)(i < c ? 2 * faure[c][i] : 2 * faure[c][i - c] + 1);
        }
    }
    t_dims.assign(D, HaltonDim{});
    for (int d = 1; d < D; ++d) {
        HaltonDim& h = t_dims[d];
        const uint64_t base = primes[d];
        uint32_t digits = 1;
        uint64_t pow_base = base;
        while (pow_base * base <= 500) { pow_base *= base; ++digits; }
        uint64_t max_power = pow_base;
        while (max_power * pow_base < (1ull << 32)) max_power *= pow_base;
        h.perm_offset = (uint32_t)t_perm.size();
        h.pow_base = (uint32_t)pow_base;
        for (uint64_t j = 0; j < pow_base; ++j) {
            uint64_t idx = j, result = 0;
            for (uint32_t k = 0; k < digits; ++k) {
                const uint64_t rem = idx % base;
                idx /= base;
                result = result * base + faure[base][rem];
            }
            t_perm.push_back((uint16_t)result);
        }
        uint64_t power = max_power / pow_base, div = 1;
        uint32_t n = 0;
        h.div[n] = 1; h.power[n] = (uint32_t)power; ++n;
        while (power / pow_base > 1) {
            div *= pow_base; power /= pow_base;
            h.div[n] = (uint32_t)div; h.power[n] = (uint32_t)power; ++n;
        }
        h.div[n] = (uint32_t)(div * pow_base); h.power[n] = 1; ++n;
        h.n_terms = n;
        h.scale = 0x1.fffffep-1 / (double)max_power;
    }
}
}  // namespace psy

namespace {

}  // namespace

namespace psy {
// Grows the per-context trace scratch to n entries.  Calls on one context
// must not overlap in time (they share this scratch).
// GTRClosure::validate and normalization (shading/surface_closure.hpp:
// 351-382): roughness clamped to [0, 0.9999] and taken as 0 below 2^-12 (a
// perfect mirror), tail shape at least 1e-4 and pushed off 1 by 1e-4, and
// the distribution's normalization (t - 1)(r^2 - 1) / (pi (1 - r^2^(1-t))),
// its pi product in double as in the reference.
void gtr_validate(float roughness, float tail_shape, float fresnel, DevShader& o) {
    float r = std::max(0.0f, std::min(0.9999f, roughness));
    float t = std::max(0.0001f, tail_shape);
    if (r < 0.000244140625f) r = 0.0f;
    if (std::abs(t - 1.0f) < 0.0001f) t = 1.0f + 0.0001f;
    const float r2 = r * r;
    const float top = (t - 1.0f) * (r2 - 1);
    const double bottom = M_PI * (double)(1.0f - std::pow(r2, 1.0f - t));
    o.roughness = r;
    o.tail_shape = t;
    o.fresnel = fresnel;
    o.norm = (float)((double)top / bottom);
}

int ensure_trace_scratch(DBuf& queue, DBuf& counters, int64_t& queue_cap, int64_t n) {
    if (!counters.p) {
        int rc = counters.alloc(sizeof(uint32_t) * 8);
        if (rc) return rc;
    }
    if (n <= queue_cap) return PSY_OK;
    HIP_OK(hipDeviceSynchronize());
    queue.release();
    const int64_t cap = std::max<int64_t>(n, 1 << 16);
    // for up to cap rays: the second pass's queue, the third pass's queue,
    // the second-pass rays' deep-node lists and the deep-node pool (3 words
    // an entry), then the main pass's continuation buffer
    int rc = queue.alloc((size_t)(6 * cap + psy::cont_area_words(cap)) * sizeof(uint32_t));
    if (rc) { queue_cap = 0; return rc; }
    queue_cap = cap;
    return PSY_OK;
}

int ensure_queue(psy_ctx* ctx, int64_t n) { return ensure_trace_scratch(ctx->queue, ctx->counters, ctx->queue_cap, n); }

static int capturing(hipStream_t st, bool* yes) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_OK(hipStreamIsCapturing(st, &cs));
    *yes = cs != hipStreamCaptureStatusNone;
    return PSY_OK;
}

int scratch_order(psy_ctx* ctx, hipStream_t st) {
    bool cap = false;
    if (int rc = capturing(st, &cap)) return rc;
    if (cap || !ctx->scratch_recorded || ctx->scratch_stream == st) return PSY_OK;
    HIP_OK(hipStreamWaitEvent(st, ctx->scratch_done, 0));
    return PSY_OK;
}

int scratch_release(psy_ctx* ctx, hipStream_t st) {
    bool cap = false;
    if (int rc = capturing(st, &cap)) return rc;
    if (cap) return PSY_OK;
    HIP_OK(hipEventRecord(ctx->scratch_done, st));
    ctx->scratch_stream = st;
    ctx->scratch_recorded = true;
    return PSY_OK;
}
}  // namespace psy

extern "C" {

const char* psy_last_error(void) { return g_err.c_str(); }
int psy_abi_version(void) { return PSY_ABI_VERSION; }

int psy_ctx_create(int device, psy_ctx** out) {
    if (!out) return fail(PSY_ERR_INVALID, "psy_ctx_create: out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(PSY_ERR_NO_DEVICE, "psy_ctx_create: no HIP device visible");
    if (device < 0 || device >= count) return fail(PSY_ERR_INVALID, "psy_ctx_create: bad device index");
    HIP_OK(hipSetDevice(device));
    std::unique_ptr<psy_ctx> c(new psy_ctx());
    c->device = device;
    std::vector<HaltonDim> hdims;
    std::vector<uint16_t> hperm;
    make_halton_tables(hdims, hperm);
    int rc = c->halton_dims.upload(hdims);
    if (rc) return rc;
    rc = c->halton_perm.upload(hperm);
    if (rc) return rc;
    HIP_OK(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device));
    if (trace_persist_occupancy(false, 16, &c->occupancy[0]) || trace_persist_occupancy(true, 16, &c->occupancy[1]) ||
        trace_persist_occupancy(false, 32, &c->occupancy32[0]) ||
        trace_persist_occupancy(true, 32, &c->occupancy32[1]) || render_persist_occupancy(&c->render_occupancy))
        return fail(PSY_ERR_HIP, "psy_ctx_create: occupancy query failed");
    HIP_OK(hipEventCreateWithFlags(&c->scratch_done, hipEventDisableTiming));
    if (static_scratch_address(device, &c->static_scratch))
        return fail(PSY_ERR_HIP, "psy_ctx_create: device address of the static trace scratch not found");
    *out = c.release();
    return PSY_OK;
}

int psy_ctx_set_second_pass(psy_ctx* ctx, uint32_t wide_min, int32_t budget) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_ctx_set_second_pass: NULL context");
    if (budget < 1) return fail(PSY_ERR_INVALID, "psy_ctx_set_second_pass: budget must be at least 1");
    ctx->deep_wide_min = wide_min;
    ctx->deep_budget = budget;
    return PSY_OK;
}

int psy_ctx_set_continuation(psy_ctx* ctx, int below) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_ctx_set_continuation: NULL context");
    if (below < 0 || below > 64) return fail(PSY_ERR_INVALID, "psy_ctx_set_continuation: below must be 0..64");
    ctx->cont_below = below;
    return PSY_OK;
}

int psy_ctx_set_trace_stack(psy_ctx* ctx, int entries) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_ctx_set_trace_stack: NULL context");
    if (entries != 16 && entries != 32) return fail(PSY_ERR_INVALID, "psy_ctx_set_trace_stack: entries must be 16 or 32");
    ctx->trace_stack = entries;
    return PSY_OK;
}

int psy_ctx_set_trace_mode(psy_ctx* ctx, int mode, int refill_min, int blocks_per_cu) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_ctx_set_trace_mode: NULL context");
    if (mode != PSY_TRACE_MODE_PLAIN && mode != PSY_TRACE_MODE_PERSISTENT && mode != PSY_TRACE_MODE_COHERENT &&
        mode != PSY_TRACE_MODE_PLAIN_COHERENT)
        return fail(PSY_ERR_INVALID, "psy_ctx_set_trace_mode: unknown mode");
    if (refill_min < 1 || refill_min > 64) return fail(PSY_ERR_INVALID, "psy_ctx_set_trace_mode: refill_min not in 1..64");
    if (blocks_per_cu > 64) return fail(PSY_ERR_INVALID, "psy_ctx_set_trace_mode: blocks_per_cu above 64");
    ctx->mode = mode;
    ctx->refill_min = refill_min;
    ctx->blocks_per_cu = blocks_per_cu;
    return PSY_OK;
}

int psy_ctx_set_render_mode(psy_ctx* ctx, int mode, int refill_min, int blocks_per_cu) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_ctx_set_render_mode: NULL context");
    if (mode != PSY_RENDER_WAVEFRONT && mode != PSY_RENDER_FUSED && mode != PSY_RENDER_STREAM)
        return fail(PSY_ERR_INVALID, "psy_ctx_set_render_mode: unknown mode");
    if (refill_min < 1 || refill_min > 64) return fail(PSY_ERR_INVALID, "psy_ctx_set_render_mode: refill_min not in 1..64");
    if (blocks_per_cu > 64) return fail(PSY_ERR_INVALID, "psy_ctx_set_render_mode: blocks_per_cu above 64");
    ctx->render_mode = mode;
    ctx->render_refill = refill_min;
    ctx->render_blocks_per_cu = blocks_per_cu;
    return PSY_OK;
}

int psy_ctx_set_stream_steps(psy_ctx* ctx, int steps_per_graph) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_ctx_set_stream_steps: NULL context");
    if (steps_per_graph < 2 || steps_per_graph > 64 || (steps_per_graph & 1))
        return fail(PSY_ERR_INVALID, "psy_ctx_set_stream_steps: steps must be even, 2..64");
    ctx->render_graph_iters = steps_per_graph;
    return PSY_OK;
}

int psy_ctx_set_render_chunk(psy_ctx* ctx, int64_t paths) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_ctx_set_render_chunk: NULL context");
    if (paths < 1 || paths > (int64_t)INT32_MAX) return fail(PSY_ERR_INVALID, "psy_ctx_set_render_chunk: paths not in 1..2^31-1");
    ctx->render_chunk = paths;
    return PSY_OK;
}

int psy_ctx_set_render_shade_waves(psy_ctx* ctx, int waves) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_ctx_set_render_shade_waves: NULL context");
    if (waves != 0 && waves != 4) return fail(PSY_ERR_INVALID, "psy_ctx_set_render_shade_waves: waves must be 0 or 4");
    ctx->render_shade_waves = waves;
    return PSY_OK;
}

int psy_ctx_set_render_streams(psy_ctx* ctx, int streams) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_ctx_set_render_streams: NULL context");
    if (streams < 1 || streams > psy_ctx::MAX_RENDER_SLOTS)
        return fail(PSY_ERR_INVALID, "psy_ctx_set_render_streams: streams must be in 1..4");
    ctx->render_streams = streams;
    return PSY_OK;
}

int psy_ctx_retrace_count(psy_ctx* ctx, int64_t* out) {
    if (!ctx || !out) return fail(PSY_ERR_INVALID, "psy_ctx_retrace_count: NULL argument");
    *out = 0;
    if (!ctx->counters.p) return PSY_OK;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipDeviceSynchronize());
    uint32_t c[4];
    HIP_OK(hipMemcpy(c, ctx->counters.p, sizeof c, hipMemcpyDeviceToHost));
    *out = (int64_t)c[0];
    return PSY_OK;
}

int psy_ctx_trace_pass_counts(psy_ctx* ctx, int64_t* second, int64_t* third) {
    if (!ctx || !second || !third) return fail(PSY_ERR_INVALID, "psy_ctx_trace_pass_counts: NULL argument");
    *second = *third = 0;
    if (!ctx->counters.p) return PSY_OK;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipDeviceSynchronize());
    uint32_t c[4];
    HIP_OK(hipMemcpy(c, ctx->counters.p, sizeof c, hipMemcpyDeviceToHost));
    *second = (int64_t)c[0];
    *third = (int64_t)c[1];
    return PSY_OK;
}

int psy_ctx_set_trace_timing(psy_ctx* ctx, int enable) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_ctx_set_trace_timing: NULL context");
    ctx->timing = enable != 0;
    return PSY_OK;
}

int psy_ctx_trace_timing(psy_ctx* ctx, double* main_ms, double* second_ms, int64_t* calls) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_ctx_trace_timing: NULL context");
    HIP_OK(hipSetDevice(ctx->device));
    double a = 0.0, b = 0.0;
    int64_t n = 0;
    int rc = PSY_OK;
    for (size_t i = 0; i + 2 < ctx->timing_events.size(); i += 3) {
        hipEvent_t e0 = (hipEvent_t)ctx->timing_events[i], e1 = (hipEvent_t)ctx->timing_events[i + 1],
                   e2 = (hipEvent_t)ctx->timing_events[i + 2];
        float x = 0.0f, y = 0.0f;
        if (hipEventSynchronize(e2) != hipSuccess || hipEventElapsedTime(&x, e0, e1) != hipSuccess ||
            hipEventElapsedTime(&y, e1, e2) != hipSuccess)
            rc = PSY_ERR_HIP;
        a += x;
        b += y;
        ++n;
    }
    for (void* e : ctx->timing_events) (void)hipEventDestroy((hipEvent_t)e);
    ctx->timing_events.clear();
    if (main_ms) *main_ms = a;
    if (second_ms) *second_ms = b;
    if (calls) *calls = n;
    return rc ? fail(rc, "psy_ctx_trace_timing: an event failed") : PSY_OK;
}

int psy_malloc(psy_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(PSY_ERR_INVALID, "psy_malloc: NULL argument");
    *out = nullptr;
    HIP_OK(hipSetDevice(ctx->device));
    if (bytes == 0) return PSY_OK;
    HIP_OK(hipMalloc(out, bytes));
    return PSY_OK;
}

int psy_free(psy_ctx* ctx, void* p) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_free: NULL context");
    if (!p) return PSY_OK;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipFree(p));
    return PSY_OK;
}

int psy_copy_to_device(psy_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || ((!dst || !src) && bytes)) return fail(PSY_ERR_INVALID, "psy_copy_to_device: NULL argument");
    HIP_OK(hipSetDevice(ctx->device));
    if (bytes) HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return PSY_OK;
}

int psy_copy_to_host(psy_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || ((!dst || !src) && bytes)) return fail(PSY_ERR_INVALID, "psy_copy_to_host: NULL argument");
    HIP_OK(hipSetDevice(ctx->device));
    if (bytes) HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return PSY_OK;
}

int psy_ctx_destroy(psy_ctx* ctx) {
    if (!ctx) return PSY_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) (void)hipGetLastError();
    delete ctx;
    return PSY_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Flattening
// ---------------------------------------------------------------------------
namespace {

M44 load_m44(const float* f) {
    M44 m;
    std::memcpy(m.m, f, sizeof(float) * 16);
    return m;
}

// merge(Transform* dest, a, b) (math/transform.hpp:241-272), a = parent
std::vector<M44> merge_frames(const std::vector<M44>& a, const M44* b, size_t nb) {
    const size_t na = a.size();
    std::vector<M44> d;
    if (na == nb) {
        for (size_t i = 0; i < na; ++i) d.push_back(m_mul(a[i], b[i]));
    } else if (na > nb) {
        if (nb == 0) d = a;
        else {
            const float inv_s = 1.0f / (float)(na - 1);
            for (size_t i = 0; i < na; ++i) d.push_back(m_mul(a[i], m_lerp_seq((float)i * inv_s, b, (int)nb)));
        }
    } else {
        if (na == 0) d.assign(b, b + nb);
        else {
            const float inv_s = 1.0f / (float)(nb - 1);
            for (size_t i = 0; i < nb; ++i) d.push_back(m_mul(b[i], m_lerp_seq((float)i * inv_s, a.data(), (int)na)));
        }
    }
    return d;
}

uint32_t intlog2(uint32_t v) {
    uint32_t r, shift;
    r = (uint32_t)(v > 0xFFFF) << 4; v >>= r;
    shift = (uint32_t)(v > 0xFF) << 3; v >>= shift; r |= shift;
    shift = (uint32_t)(v > 0xF) << 2; v >>= shift; r |= shift;
    shift = (uint32_t)(v > 0x3) << 1; v >>= shift; r |= shift;
    r |= (v >> 1);
    return r;
}
uint32_t upper_pow2(uint32_t v) {
    v--; v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16; v++;
    return v;
}

struct FlatLeaf {
    int32_t object;   // global object index
    int32_t frame;    // index into frames (-1 none)
    int32_t shader;
    uint64_t id;
};

struct Flattener {
    const psy_scene_desc* d;
    std::vector<int32_t> obj_base;           // first global object of each assembly
    // per global object: its first patch object and patch count when it is
    // a subdivision surface (first -1 otherwise)
    std::vector<int32_t> sub_first, sub_count;
    std::vector<std::vector<M44>> frames;
    std::vector<FlatLeaf> leaves;
    int depth_guard = 0;
    std::string err;

    bool run(int a, int frame, uint64_t prefix, int shader) {
        if (++depth_guard > 64) { err = "assembly nesting deeper than 64 (cycle?)"; return false; }
        const psy_assembly_desc& ad = d->assemblies[a];
        const int bits = (int)intlog2(upper_pow2((uint32_t)ad.n_instances));
        for (int i = 0; i < ad.n_instances; ++i) {
            const psy_instance_desc& in = ad.instances[i];
            const uint64_t id = (prefix << bits) | ((uint64_t)i & (uint64_t)((1 << bits) - 1));
            int fr = frame;
            if (in.transform_count > 0) {
                if (in.transform_index < 0 || in.transform_index + in.transform_count > ad.n_transforms) {
                    err = "instance transform range out of bounds";
                    return false;
                }
                std::vector<M44> xs;
                for (int k = 0; k < in.transform_count; ++k)
                    xs.push_back(load_m44(ad.transforms + 16 * (in.transform_index + k)));
                const std::vector<M44> parent = frame >= 0 ? frames[frame] : std::vector<M44>();
                frames.push_back(merge_frames(parent, xs.data(), xs.size()));
                fr = (int)frames.size() - 1;
            }
            const int sh = in.shader_index >= 0 ? in.shader_index : shader;
            if (in.shader_index >= d->n_shaders) { err = "shader index out of range"; return false; }
            if (in.type == PSY_INSTANCE_OBJECT) {
                if (in.data_index < 0 || in.data_index >= ad.n_objects) { err = "object index out of range"; return false; }
                const int g = obj_base[a] + in.data_index;
                if (sub_first[g] >= 0) {
                    // SubdivisionSurface::intersect_rays passes the instance's
                    // id to every patch (object/subdivision_surface.cpp:41-83)
                    for (int k = 0; k < sub_count[g]; ++k) leaves.push_back(FlatLeaf{sub_first[g] + k, fr, sh, id});
                } else {
                    leaves.push_back(FlatLeaf{g, fr, sh, id});
                }
            } else if (in.type == PSY_INSTANCE_ASSEMBLY) {
                if (in.data_index < 0 || in.data_index >= ad.n_assemblies) { err = "sub-assembly index out of range"; return false; }
                const int sub = ad.assemblies[in.data_index];
                if (sub <= 0 || sub >= d->n_assemblies) { err = "sub-assembly scene index out of range"; return false; }
                if (!run(sub, fr, id, sh)) return false;
            } else {
                err = "unknown instance type";
                return false;
            }
        }
        --depth_guard;
        return true;
    }
};

// ---- BVH4 builder: binned-SAH binary tree collapsed to 4-wide --------------
struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
    }
    void growp(const float* p) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
    }
    float area() const {
        const float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
        if (!(x >= 0) || !(y >= 0) || !(z >= 0)) return 0.0f;
        return 2.0f * (x * y + x * z + y * z);
    }
};

struct B2 {
    Box box;         // over the whole shutter (the split heuristic's box)
    Box box0, box1;  // at time 0 and 1: their lerp bounds the subtree at any time
    int left = -1, right = -1;
    int leaf = -1;
    int count = 0;   // leaves below
};

struct Builder2 {
    std::vector<B2> nodes;
    std::vector<Box> boxes;
    std::vector<Box> boxes0, boxes1;  // per prim at time 0 and 1
    std::vector<float> cent;  // 3 per prim
    std::vector<int> idx;

    // Binned SAH down to depth SAH_DEPTH, median splits below it, so a
    // skewed layout (e.g. instances spaced geometrically, which SAH peels
    // off one at a time) stays within SAH_DEPTH + log2(n) levels.  With
    // median_only every split is a median one (a balanced tree).
    static constexpr int SAH_DEPTH = 20;
    bool median_only = false;

    int build(int begin, int end, int depth = 0) {
        const int me = (int)nodes.size();
        nodes.emplace_back();
        Box b, cb, b0, b1;
        for (int i = begin; i < end; ++i) {
            b.grow(boxes[idx[i]]);
            b0.grow(boxes0[idx[i]]);
            b1.grow(boxes1[idx[i]]);
            cb.growp(&cent[3 * idx[i]]);
        }
        nodes[me].box = b;
        nodes[me].box0 = b0;
        nodes[me].box1 = b1;
        nodes[me].count = end - begin;
        if (end - begin == 1) { nodes[me].leaf = idx[begin]; return me; }
        int axis = 0;
        float ext[3];
        for (int k = 0; k < 3; ++k) ext[k] = cb.hi[k] - cb.lo[k];
        if (ext[1] > ext[axis]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;
        int mid = (begin + end) / 2;
        if (ext[axis] > 0.0f && (median_only || depth >= SAH_DEPTH)) {
            std::nth_element(idx.begin() + begin, idx.begin() + mid, idx.begin() + end,
                             [&](int a, int c) { return cent[3 * a + axis] < cent[3 * c + axis]; });
        } else if (ext[axis] > 0.0f) {
            const int NB = 16;
            Box bb[NB];
            int bc[NB] = {0};
            const float s = (float)NB / ext[axis];
            auto bin_of = [&](int p) {
                int k = (int)((cent[3 * p + axis] - cb.lo[axis]) * s);
                return std::min(NB - 1, std::max(0, k));
            };
            for (int i = begin; i < end; ++i) { const int k = bin_of(idx[i]); bb[k].grow(boxes[idx[i]]); bc[k]++; }
            float best = INFINITY;
            int best_k = -1;
            Box lacc[NB];
            int lcnt[NB];
            Box acc; int cnt = 0;
            for (int k = 0; k < NB; ++k) { acc.grow(bb[k]); cnt += bc[k]; lacc[k] = acc; lcnt[k] = cnt; }
            Box racc; int rcnt = 0;
            for (int k = NB - 1; k > 0; --k) {
                racc.grow(bb[k]); rcnt += bc[k];
                const int lc = lcnt[k - 1];
                if (lc == 0 || rcnt == 0) continue;
                const float cost = lacc[k - 1].area() * lc + racc.area() * rcnt;
                if (cost < best) { best = cost; best_k = k; }
            }
            if (best_k > 0) {
                auto it = std::partition(idx.begin() + begin, idx.begin() + end,
                                         [&](int p) { return bin_of(p) < best_k; });
                mid = (int)(it - idx.begin());
                if (mid == begin || mid == end) mid = (begin + end) / 2;
            } else {
                std::nth_element(idx.begin() + begin, idx.begin() + mid, idx.begin() + end,
                                 [&](int a, int c) { return cent[3 * a + axis] < cent[3 * c + axis]; });
            }
        }
        const int l = build(begin, mid, depth + 1);
        const int r = build(mid, end, depth + 1);
        nodes[me].left = l;
        nodes[me].right = r;
        return me;
    }
};

struct Collapser {
    const std::vector<B2>* n2;
    std::vector<BvhNode> out;   // child boxes at time 0
    std::vector<BvhNode> out1;  // the same nodes: child box change from time 0 to 1
    int max_need = 0;
    // which binary child a 4-wide node opens up next: the largest surface
    // area (the SAH's choice), or with by_count the most leaves (halves the
    // depth of a balanced tree)
    bool by_count = false;

    // returns child ref, and the traversal stack need of the subtree
    int32_t emit(int b2, int* need) {
        const B2& nd = (*n2)[b2];
        if (nd.leaf >= 0) { *need = 0; return ~nd.leaf; }
        std::vector<int> kids = {nd.left, nd.right};
        while (kids.size() < 4) {
            int best = -1;
            float ba = -1.0f;
            for (int k = 0; k < (int)kids.size(); ++k) {
                const B2& c = (*n2)[kids[k]];
                if (c.leaf >= 0) continue;
                const float a = by_count ? (float)c.count : c.box.area();
                if (a > ba) { ba = a; best = k; }
            }
            if (best < 0) break;
            const int c = kids[best];
            kids.erase(kids.begin() + best);
            kids.push_back((*n2)[c].left);
            kids.push_back((*n2)[c].right);
        }
        const int me = (int)out.size();
        out.emplace_back();
        out1.emplace_back();
        int32_t refs[4];
        int sub_need = 0;
        for (int k = 0; k < 4; ++k) {
            if (k < (int)kids.size()) {
                int cn = 0;
                refs[k] = emit(kids[k], &cn);
                sub_need = std::max(sub_need, cn);
            } else {
                refs[k] = CHILD_EMPTY;
            }
        }
        // out: child boxes at time 0; out1: their change up to time 1
        // (0 for empty slots, whose +inf / -inf then stay put)
        BvhNode& o = out[me];
        BvhNode& o1 = out1[me];
        for (int k = 0; k < 4; ++k) {
            o.child[k] = o1.child[k] = refs[k];
            if (k < (int)kids.size()) {
                const Box& b = (*n2)[kids[k]].box0;
                const Box& e = (*n2)[kids[k]].box1;
                o.lo_x[k] = b.lo[0]; o.hi_x[k] = b.hi[0];
                o.lo_y[k] = b.lo[1]; o.hi_y[k] = b.hi[1];
                o.lo_z[k] = b.lo[2]; o.hi_z[k] = b.hi[2];
                o1.lo_x[k] = e.lo[0] - b.lo[0]; o1.hi_x[k] = e.hi[0] - b.hi[0];
                o1.lo_y[k] = e.lo[1] - b.lo[1]; o1.hi_y[k] = e.hi[1] - b.hi[1];
                o1.lo_z[k] = e.lo[2] - b.lo[2]; o1.hi_z[k] = e.hi[2] - b.hi[2];
            } else {
                o.lo_x[k] = o.lo_y[k] = o.lo_z[k] = INFINITY;
                o.hi_x[k] = o.hi_y[k] = o.hi_z[k] = -INFINITY;
                o1.lo_x[k] = o1.lo_y[k] = o1.lo_z[k] = 0.0f;
                o1.hi_x[k] = o1.hi_y[k] = o1.hi_z[k] = 0.0f;
            }
        }
        const int kc = (int)kids.size();
        *need = std::max(kc, kc - 1 + sub_need);
        return me;
    }
};

// exact inverse in double for conservative world bounds
bool invert_d(const M44& m, double inv[16]) {
    double a[16];
    for (int i = 0; i < 16; ++i) a[i] = m.m[i];
    double r[16];
    r[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    r[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    r[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    r[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    r[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    r[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    r[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    r[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    r[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    r[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    r[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    r[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    r[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    r[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    r[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    r[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    const double det = a[0] * r[0] + a[1] * r[4] + a[2] * r[8] + a[3] * r[12];
    if (det == 0.0 || !std::isfinite(det)) return false;
    for (int i = 0; i < 16; ++i) inv[i] = r[i] / det;
    return true;
}

}  // namespace

namespace psy {

int dice_layout(size_t n_patches, int dice_depth, int& levels, int64_t& per, std::string& err) {
    levels = dice_depth;
    if (levels < 0) {
        // Up to 24 levels while the cache stays within 2 GiB (scenes of a few
        // large patches, e.g. a ground plane diced to pixel size), else up to
        // 18 levels within 16 GiB (under 6% of an MI355X's HBM); rays that
        // need deeper splits take the second pass
        const double gib = 1024.0 * 1024.0 * 1024.0;
        auto bytes = [&](int l) { return (double)((int64_t)1 << (l + 1)) * 32.0 * (double)n_patches; };
        levels = 24;
        while (levels > 18 && bytes(levels) > 2.0 * gib) --levels;
        while (levels > 0 && bytes(levels) > 16.0 * gib) --levels;
    }
    if (levels > 24) {
        err = "dice_depth above 24";
        return PSY_ERR_INVALID;
    }
    per = (int64_t)1 << (levels + 1);  // slot: 1-based heap, entry 0 unused
    // the main pass addresses the cache with 32-bit node indices
    if ((double)per * (double)n_patches > 4294967295.0) {
        err = "dice cache above 2^32 nodes; lower dice_depth";
        return PSY_ERR_INVALID;
    }
    return PSY_OK;
}

int rebuild_dice(psy_scene* sc, int dice_depth) {
    int levels = 0;
    int64_t per = 2;
    std::string err;
    if (int rc = dice_layout(sc->patch_ids.size(), dice_depth, levels, per, err))
        return fail(rc, "psy_scene_set_dice_depth: " + err);
    HIP_OK(hipSetDevice(sc->device));
    // the new leaves, objects and cache are built beside the old ones and
    // swapped in only once all of them exist, so a failure leaves the scene
    // as it was
    std::vector<DevObject> objs = sc->host_objs;
    std::vector<DevInstance> insts = sc->host_insts;
    for (size_t k = 0; k < sc->patch_ids.size(); ++k) {
        objs[sc->patch_ids[k]].dice_levels = levels;
        objs[sc->patch_ids[k]].dice = (int64_t)k * per;
    }
    for (DevInstance& I : insts) {
        I.dice_levels = objs[I.object].dice_levels;
        I.dice = objs[I.object].dice;
    }
    DBuf inst, obj, dice;
    int rc;
    if ((rc = inst.upload(insts))) return rc;
    if ((rc = obj.upload(objs))) return rc;
    const size_t dice_bytes = (size_t)per * sc->patch_ids.size() * 32;
    if ((rc = dice.alloc(dice_bytes))) return rc;
    if (!sc->patch_ids.empty()) {
        DBuf pid;
        if ((rc = pid.upload(sc->patch_ids.data(), sc->patch_ids.size() * sizeof(int32_t)))) return rc;
        rc = launch_build_dice((const float*)sc->odata.p, (const DevObject*)obj.p, (const int32_t*)pid.p,
                               (int)sc->patch_ids.size(), levels, (float4*)dice.p, nullptr);
        if (rc) return fail(rc, "psy_scene_set_dice_depth: dice cache build launch failed");
    }
    // no launch of any stream may still read the old cache or leaves, and
    // the new cache must be complete
    HIP_OK(hipDeviceSynchronize());
    const size_t old_bytes = sc->dice.bytes;
    std::swap(sc->inst.p, inst.p); std::swap(sc->inst.bytes, inst.bytes);
    std::swap(sc->obj.p, obj.p); std::swap(sc->obj.bytes, obj.bytes);
    std::swap(sc->dice.p, dice.p); std::swap(sc->dice.bytes, dice.bytes);
    sc->host_objs.swap(objs);
    sc->host_insts.swap(insts);
    sc->dev.inst = (const DevInstance*)sc->inst.p;
    sc->dev.obj = (const DevObject*)sc->obj.p;
    sc->dev.dice = (const float4*)sc->dice.p;
    sc->info.dice_depth = levels;
    sc->info.dice_cache_bytes = (int64_t)dice_bytes;
    sc->info.device_bytes += (int64_t)dice_bytes - (int64_t)old_bytes;
    return PSY_OK;
}

int prepare_scene(const psy_scene_desc* d, int dice_depth, HostScene& H) {
    if (!d) return fail(PSY_ERR_INVALID, "psy_scene_create: NULL argument");
    if (d->n_assemblies < 1 || !d->assemblies) return fail(PSY_ERR_INVALID, "psy_scene_create: no root assembly");
    if (d->n_shaders < 0 || (d->n_shaders > 0 && !d->shaders)) return fail(PSY_ERR_INVALID, "psy_scene_create: bad shader table");
    H = HostScene{};
    std::vector<DevObject>& objs = H.objs;
    std::vector<float>& odata = H.odata;
    std::vector<M44>& xf_pool = H.xf;

    // ---- objects (global numbering: assemblies in order) -------------------
    Flattener fl;
    fl.d = d;
    std::vector<Box> obox;  // union over time samples
    std::vector<std::pair<int32_t, std::vector<float>>> subdiv;  // (global object, its patches)
    for (int a = 0; a < d->n_assemblies; ++a) {
        const psy_assembly_desc& ad = d->assemblies[a];
        fl.obj_base.push_back((int32_t)objs.size());
        for (int i = 0; i < ad.n_objects; ++i) {
            const psy_object_desc& od = ad.objects[i];
            if (od.kind == PSY_OBJ_SUBDIVISION) {
                if (od.time_samples != 1)
                    return fail(PSY_ERR_UNSUPPORTED, "psy_scene_create: deforming subdivision surfaces (time_samples > 1) are not supported on the device path");
                if (od.n_vertices < 0 || od.n_faces < 0 || (od.n_vertices > 0 && !od.data))
                    return fail(PSY_ERR_INVALID, "psy_scene_create: subdivision surface without its control mesh");
                std::vector<float> patches;
                std::string err;
                if (int rc = subdivide_patches(od.data, od.n_vertices, od.face_vert_counts, od.n_faces, od.face_vert_indices, patches, err))
                    return fail(rc, "psy_scene_create: subdivision surface: " + err);
                subdiv.emplace_back((int32_t)objs.size(), std::move(patches));
                DevObject o{};
                o.kind = od.kind;
                o.ts = 1;
                obox.push_back(Box{});
                objs.push_back(o);
                continue;
            }
            const int nf = floats_per_kind(od.kind);
            if (nf == 0) return fail(PSY_ERR_INVALID, "psy_scene_create: unknown object kind");
            if (od.time_samples < 1 || !od.data) return fail(PSY_ERR_INVALID, "psy_scene_create: object without data");
            if ((od.kind == PSY_OBJ_BILINEAR || od.kind == PSY_OBJ_BICUBIC) && od.time_samples != 1)
                return fail(PSY_ERR_UNSUPPORTED, "psy_scene_create: deforming patches (time_samples > 1) are not supported on the device path");
            DevObject o{};
            o.kind = od.kind;
            o.ts = od.time_samples;
            o.data = (int32_t)odata.size();
            o.dice_levels = 0;
            o.dice = 0;
            odata.insert(odata.end(), od.data, od.data + (size_t)nf * od.time_samples);
            Box b;
            for (int t = 0; t < od.time_samples; ++t) {
                const float* p = od.data + (size_t)t * nf;
                float lo[3], hi[3];
                switch (od.kind) {
                    case PSY_OBJ_SPHERE:
                    case PSY_OBJ_SPHERE_LIGHT:
                        for (int k = 0; k < 3; ++k) { lo[k] = p[k] - p[3]; hi[k] = p[k] + p[3]; }
                        break;
                    case PSY_OBJ_RECTANGLE_LIGHT:
                        lo[0] = -p[0] * 0.5f; hi[0] = p[0] * 0.5f;
                        lo[1] = -p[1] * 0.5f; hi[1] = p[1] * 0.5f;
                        lo[2] = 0.0f; hi[2] = 0.0f;
                        break;
                    default: {
                        const int np = nf / 3;
                        for (int k = 0; k < 3; ++k) { lo[k] = p[k]; hi[k] = p[k]; }
                        for (int q = 1; q < np; ++q)
                            for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[3 * q + k]); hi[k] = std::max(hi[k], p[3 * q + k]); }
                    }
                }
                b.growp(lo);
                b.growp(hi);
            }
            obox.push_back(b);
            objs.push_back(o);
        }
    }
    // the patches of the subdivision surfaces, after every described object
    fl.sub_first.assign(objs.size(), -1);
    fl.sub_count.assign(objs.size(), 0);
    for (const auto& sp : subdiv) {
        const size_t np = sp.second.size() / 48;
        fl.sub_first[sp.first] = (int32_t)objs.size();
        fl.sub_count[sp.first] = (int32_t)np;
        for (size_t k = 0; k < np; ++k) {
            const float* p = sp.second.data() + 48 * k;
            DevObject o{};
            o.kind = PSY_OBJ_BICUBIC;
            o.ts = 1;
            o.data = (int32_t)odata.size();
            odata.insert(odata.end(), p, p + 48);
            Box b;
            for (int q = 0; q < 16; ++q) b.growp(p + 3 * q);
            obox.push_back(b);
            objs.push_back(o);
        }
    }
    if (odata.size() > (size_t)INT32_MAX) return fail(PSY_ERR_UNSUPPORTED, "psy_scene_create: object data above 2^31 floats");
    if (!fl.run(0, -1, 0, -1)) return fail(PSY_ERR_INVALID, "psy_scene_create: " + fl.err);
    if (int rc = prepare_lights(d, fl.obj_base, H)) return rc;
    for (const FlatLeaf& L : fl.leaves) {
        const int k = objs[L.object].kind;
        // a surface without a shader gets the reference's magenta emitter
        if (L.shader < 0 && k != PSY_OBJ_SPHERE_LIGHT && k != PSY_OBJ_RECTANGLE_LIGHT) H.chromatic = true;
    }

    // ---- leaf instances, transform pool, world bounds ----------------------
    std::vector<int32_t> frame_off(fl.frames.size());
    for (size_t f = 0; f < fl.frames.size(); ++f) {
        frame_off[f] = (int32_t)xf_pool.size();
        xf_pool.insert(xf_pool.end(), fl.frames[f].begin(), fl.frames[f].end());
    }
    const size_t nl = fl.leaves.size();
    if (nl > (size_t)INT32_MAX / 2) return fail(PSY_ERR_UNSUPPORTED, "psy_scene_create: too many instances");
    std::vector<DevInstance>& insts = H.insts;
    insts.resize(nl);
    Builder2 b2;
    bool moving = false;
    b2.boxes.resize(nl);
    b2.boxes0.resize(nl);
    b2.boxes1.resize(nl);
    b2.cent.resize(3 * nl);
    b2.idx.resize(nl);
    for (size_t i = 0; i < nl; ++i) {
        const FlatLeaf& L = fl.leaves[i];
        DevInstance& I = insts[i];
        I = DevInstance{};
        I.object = L.object;
        I.kind = objs[L.object].kind;
        I.shader = L.shader;
        I.xf_offset = L.frame >= 0 ? frame_off[L.frame] : 0;
        I.xf_count = L.frame >= 0 ? (int32_t)fl.frames[L.frame].size() : 0;
        if (I.xf_count > 0 && (I.kind == PSY_OBJ_BILINEAR || I.kind == PSY_OBJ_BICUBIC)) H.transformed_patches = true;
        const Box& ob = obox[L.object];
        Box wb;
        Box wb0, wb1;
        if (I.xf_count == 0) {
            wb = ob;
            wb0 = wb1 = ob;
        } else {
            // the object's box corners carried to world space at time
            // samples (64 steps per motion segment).  wb: over the whole
            // shutter; wb0, wb1: at time 0 and 1, padded by the largest
            // amount the sampled boxes leave their lerp, plus the largest
            // corner displacement between neighbouring steps (the corners
            // move smoothly, so between steps they stay within it)
            const std::vector<M44>& fr = fl.frames[L.frame];
            const int steps = fr.size() > 1 ? 64 * (int)(fr.size() - 1) : 0;
            std::vector<Box> at((size_t)steps + 1);
            float prev[8][3] = {};
            float move = 0.0f;
            for (int s = 0; s <= steps; ++s) {
                const float t = steps ? (float)s / steps : 0.0f;
                const M44 m = m_lerp_seq(t, fr.data(), (int)fr.size());
                double inv[16];
                if (!invert_d(m, inv)) return fail(PSY_ERR_INVALID, "psy_scene_create: singular instance transform");
                for (int c = 0; c < 8; ++c) {
                    const double pc[3] = {(c & 1) ? ob.hi[0] : ob.lo[0], (c & 2) ? ob.hi[1] : ob.lo[1],
                                          (c & 4) ? ob.hi[2] : ob.lo[2]};
                    double r[4];
                    for (int j = 0; j < 4; ++j) r[j] = pc[0] * inv[0 * 4 + j] + pc[1] * inv[1 * 4 + j] + pc[2] * inv[2 * 4 + j] + inv[3 * 4 + j];
                    const float q[3] = {(float)(r[0] / r[3]), (float)(r[1] / r[3]), (float)(r[2] / r[3])};
                    wb.growp(q);
                    at[(size_t)s].growp(q);
                    if (s > 0)
                        for (int k = 0; k < 3; ++k) move = std::max(move, std::fabs(q[k] - prev[c][k]));
                    for (int k = 0; k < 3; ++k) prev[c][k] = q[k];
                }
            }
            float dev = 0.0f;
            for (int s = 0; s <= steps; ++s) {
                const float t = steps ? (float)s / steps : 0.0f;
                for (int k = 0; k < 3; ++k) {
                    const float lo = at[0].lo[k] * (1.0f - t) + at[(size_t)steps].lo[k] * t;
                    const float hi = at[0].hi[k] * (1.0f - t) + at[(size_t)steps].hi[k] * t;
                    dev = std::max(dev, std::max(lo - at[(size_t)s].lo[k], at[(size_t)s].hi[k] - hi));
                }
            }
            float ext = 0.0f;
            for (int k = 0; k < 3; ++k) ext = std::max(ext, std::max(std::fabs(wb.lo[k]), std::fabs(wb.hi[k])));
            const float pad = move + 1e-5f * ext + 1e-6f;
            wb0 = at[0];
            wb1 = at[(size_t)steps];
            for (int k = 0; k < 3; ++k) {
                wb.lo[k] -= pad; wb.hi[k] += pad;
                // between samples the lerp moves by at most one step's
                // displacement too, hence the second `move`
                wb0.lo[k] -= pad + dev + move; wb0.hi[k] += pad + dev + move;
                wb1.lo[k] -= pad + dev + move; wb1.hi[k] += pad + dev + move;
            }
            if (steps) moving = true;
        }
        b2.boxes[i] = wb;
        b2.boxes0[i] = wb0;
        b2.boxes1[i] = wb1;
        for (int k = 0; k < 3; ++k) b2.cent[3 * i + k] = 0.5f * (wb.lo[k] + wb.hi[k]);
        b2.idx[i] = (int)i;
        H.ids.push_back(L.id);
    }
    std::vector<BvhNode>& bvh = H.bvh;
    int32_t root = CHILD_EMPTY;
    int stack_need = 1;
    if (nl == 1) {
        root = ~0;
    } else if (nl > 1) {
        b2.nodes.reserve(2 * nl);
        b2.build(0, (int)nl);
        Collapser cl;
        cl.n2 = &b2.nodes;
        int need = 0;
        root = cl.emit(0, &need);
        stack_need = std::max(1, need) + 1;
        if (stack_need > 64) {
            // A layout the SAH tree nests too deeply for the traversal
            // stacks: a balanced tree (median splits, 4-wide nodes opened up
            // by leaf count) needs at most 3 entries per 4-wide level, 49
            // for 2^32 leaves.
            b2.nodes.clear();
            b2.median_only = true;
            b2.build(0, (int)nl);
            cl = Collapser{};
            cl.n2 = &b2.nodes;
            cl.by_count = true;
            root = cl.emit(0, &need);
            stack_need = std::max(1, need) + 1;
        }
        if (stack_need > 64)
            return fail(PSY_ERR_UNSUPPORTED, "psy_scene_create: the instance BVH needs a traversal stack of " +
                                                 std::to_string(stack_need) + " entries; at most 64 are supported");
        bvh.swap(cl.out);
        if (moving) H.bvh1.swap(cl.out1);
    }
    H.root = root;
    H.stack_need = stack_need;

    // ---- dice cache --------------------------------------------------------
    std::vector<int32_t>& patch_ids = H.patch_ids;
    for (size_t i = 0; i < objs.size(); ++i)
        if (objs[i].kind == PSY_OBJ_BILINEAR || objs[i].kind == PSY_OBJ_BICUBIC) patch_ids.push_back((int32_t)i);
    int levels = 0;
    int64_t per = 2;
    std::string derr;
    if (int rc = dice_layout(patch_ids.size(), dice_depth, levels, per, derr)) return fail(rc, "psy_scene_create: " + derr);
    H.levels = levels;
    H.per = per;
    for (size_t k = 0; k < patch_ids.size(); ++k) {
        objs[patch_ids[k]].dice_levels = levels;
        objs[patch_ids[k]].dice = (int64_t)k * per;
    }
    for (DevInstance& I : insts) {
        I.dice_levels = objs[I.object].dice_levels;
        I.dice = objs[I.object].dice;
    }

    // ---- emission scales of the light objects ------------------------------
    H.escale.assign(objs.size(), 1.0f);
    H.light_param.assign(objs.size(), -1);
    for (size_t i = 0; i < objs.size(); ++i)
        if (objs[i].kind == PSY_OBJ_SPHERE_LIGHT || objs[i].kind == PSY_OBJ_RECTANGLE_LIGHT) {
            H.light_param[i] = (int32_t)H.light_objects.size();
            H.light_objects.push_back((int32_t)i);
        }

    // ---- shaders, camera (basics/camera.hpp:23-69) -------------------------
    H.shaders.resize(d->n_shaders);
    for (int i = 0; i < d->n_shaders; ++i) {
        const psy_shader_desc& sd = d->shaders[i];
        if (sd.kind != PSY_SHADER_LAMBERT && sd.kind != PSY_SHADER_EMIT && sd.kind != PSY_SHADER_GTR)
            return fail(PSY_ERR_INVALID, "psy_scene_create: shader " + std::to_string(i) + " has an unknown kind");
        DevShader& o = H.shaders[i];
        o.kind = sd.kind;
        for (int k = 0; k < 3; ++k) o.color[k] = sd.color[k];
        o.roughness = o.tail_shape = o.fresnel = o.norm = 0.0f;
        if (sd.kind == PSY_SHADER_GTR) gtr_validate(sd.roughness, sd.tail_shape, sd.fresnel, o);
    }
    const psy_camera_desc& c = d->camera;
    DevCamera& cam = H.cam;
    if (c.n_transforms > MAX_CAM_SAMPLES || c.n_fovs > MAX_CAM_SAMPLES || c.n_aperture_radii > MAX_CAM_SAMPLES ||
        c.n_focus_distances > MAX_CAM_SAMPLES)
        return fail(PSY_ERR_UNSUPPORTED, "psy_scene_create: more than 8 camera motion samples");
    cam.n_xf = std::max(1, (int)c.n_transforms);
    if (c.n_transforms == 0) cam.xf[0] = m_identity();
    for (int i = 0; i < c.n_transforms; ++i) cam.xf[i] = load_m44(c.transforms + 16 * i);
    cam.n_tfov = std::max(1, (int)c.n_fovs);
    cam.tfov[0] = 1.0f;
    for (int i = 0; i < c.n_fovs; ++i) {
        const float f = c.fovs[i];
        cam.tfov[i] = (float)(std::sin((double)(f / 2)) / std::cos((double)(f / 2)));
    }
    bool focal_ok = c.n_aperture_radii > 0 && c.n_focus_distances > 0;
    for (int i = 0; focal_ok && i < c.n_focus_distances; ++i) if (c.focus_distances[i] <= 0.0f) focal_ok = false;
    if (focal_ok) {
        cam.n_ap = c.n_aperture_radii;
        for (int i = 0; i < cam.n_ap; ++i) cam.aperture[i] = c.aperture_radii[i];
        cam.n_fd = c.n_focus_distances;
        for (int i = 0; i < cam.n_fd; ++i) cam.focus[i] = c.focus_distances[i];
    } else {
        cam.n_ap = 1; cam.aperture[0] = 0.0f;
        cam.n_fd = 1; cam.focus[0] = 1.0f;
    }
    H.dice_rate = d->dice_rate;
    H.min_upoly = d->min_upoly_size;
    for (int k = 0; k < 3; ++k) H.background[k] = d->background[k];
    return PSY_OK;
}

}  // namespace psy

extern "C" {

int psy_scene_create(psy_ctx* ctx, const psy_scene_desc* d, int dice_depth, psy_scene** out) {
    if (!ctx || !d || !out) return fail(PSY_ERR_INVALID, "psy_scene_create: NULL argument");
    *out = nullptr;
    HIP_OK(hipSetDevice(ctx->device));
    std::unique_ptr<psy_scene> sc(new psy_scene());
    sc->ctx = ctx;
    sc->device = ctx->device;
    HostScene H;
    if (int rc = prepare_scene(d, dice_depth, H)) return rc;
    sc->stack_need = H.stack_need;
    sc->transformed_patches = H.transformed_patches;
    sc->chromatic = H.chromatic;
    sc->ids = H.ids;
    for (int k = 0; k < 3; ++k) sc->background[k] = H.background[k];
    sc->cam = H.cam;
    const std::vector<int32_t>& patch_ids = H.patch_ids;
    const int levels = H.levels;
    const int64_t per = H.per;
    const size_t nl = H.insts.size();

    // ---- upload ------------------------------------------------------------
    int rc;
    if ((rc = sc->bvh.upload(H.bvh))) return rc;
    if (!H.bvh1.empty() && (rc = sc->bvh1.upload(H.bvh1))) return rc;
    if ((rc = sc->inst.upload(H.insts))) return rc;
    if ((rc = sc->obj.upload(H.objs))) return rc;
    if ((rc = sc->odata.upload(H.odata))) return rc;
    if ((rc = sc->xf.upload(H.xf))) return rc;
    if ((rc = sc->shaders.upload(H.shaders))) return rc;
    if ((rc = sc->escale_dev.upload(H.escale))) return rc;
    if ((rc = sc->lparam_dev.upload(H.light_param))) return rc;
    sc->host_shaders = H.shaders;
    sc->host_insts = H.insts;
    sc->host_objs = H.objs;
    sc->patch_ids = H.patch_ids;
    sc->light_objects = H.light_objects;
    sc->escale = H.escale;
    if ((rc = upload_lights(H, sc.get()))) return rc;
    const size_t dice_bytes = (size_t)per * patch_ids.size() * 32;
    if ((rc = sc->dice.alloc(dice_bytes))) return rc;
    if (!patch_ids.empty()) {
        DBuf pid;
        if ((rc = pid.upload(patch_ids.data(), patch_ids.size() * sizeof(int32_t)))) return rc;
        rc = launch_build_dice((const float*)sc->odata.p, (const DevObject*)sc->obj.p, (const int32_t*)pid.p,
                               (int)patch_ids.size(), levels, (float4*)sc->dice.p, nullptr);
        if (rc) return fail(rc, "psy_scene_create: dice cache build launch failed");
        HIP_OK(hipDeviceSynchronize());
    }

    DevScene& s = sc->dev;
    s.bvh = (const BvhNode*)sc->bvh.p;
    s.bvh1 = H.bvh1.empty() ? nullptr : (const BvhNode*)sc->bvh1.p;
    s.n_bvh = (int32_t)H.bvh.size();
    s.root = H.root;
    s.inst = (const DevInstance*)sc->inst.p;
    s.obj = (const DevObject*)sc->obj.p;
    s.odata = (const float*)sc->odata.p;
    s.xf = (const M44*)sc->xf.p;
    s.dice = (const float4*)sc->dice.p;
    s.shaders = (const DevShader*)sc->shaders.p;
    s.escale = (const float*)sc->escale_dev.p;
    s.dice_rate = H.dice_rate;
    s.min_upoly = H.min_upoly;
    s.n_inst = (int32_t)nl;
    s.flags = d->flags;

    psy_scene_info& info = sc->info;
    info.n_leaf_instances = (int64_t)nl;
    info.n_objects = (int64_t)H.objs.size();
    info.n_bvh_nodes = (int64_t)H.bvh.size();
    info.n_patches = (int64_t)patch_ids.size();
    info.dice_depth = levels;
    info.bvh_stack_need = H.stack_need;
    info.dice_cache_bytes = (int64_t)dice_bytes;
    info.device_bytes = (int64_t)(sc->bvh.bytes + sc->bvh1.bytes + sc->inst.bytes + sc->obj.bytes + sc->odata.bytes + sc->xf.bytes +
                                  sc->shaders.bytes + dice_bytes);
    *out = sc.release();
    return PSY_OK;
}

int psy_scene_set_dice_depth(psy_scene* s, int dice_depth) {
    if (!s) return fail(PSY_ERR_INVALID, "psy_scene_set_dice_depth: NULL scene");
    return rebuild_dice(s, dice_depth);
}

int psy_scene_param_counts(const psy_scene* s, int64_t* n_albedo, int64_t* n_emission) {
    if (!s) return fail(PSY_ERR_INVALID, "psy_scene_param_counts: NULL scene");
    if (n_albedo) *n_albedo = (int64_t)s->host_shaders.size();
    if (n_emission) *n_emission = (int64_t)s->light_objects.size();
    return PSY_OK;
}

int psy_scene_get_params(const psy_scene* s, float* albedo, float* emission_scale) {
    if (!s) return fail(PSY_ERR_INVALID, "psy_scene_get_params: NULL scene");
    if (albedo)
        for (size_t i = 0; i < s->host_shaders.size(); ++i) albedo[i] = s->host_shaders[i].color[0];
    if (emission_scale)
        for (size_t j = 0; j < s->light_objects.size(); ++j) emission_scale[j] = s->escale[s->light_objects[j]];
    return PSY_OK;
}

int psy_scene_set_params(psy_scene* s, const float* albedo, int64_t n_albedo, const float* emission_scale,
                         int64_t n_emission) {
    if (!s) return fail(PSY_ERR_INVALID, "psy_scene_set_params: NULL scene");
    if (albedo && n_albedo != (int64_t)s->host_shaders.size())
        return fail(PSY_ERR_INVALID, "psy_scene_set_params: albedo count differs from the scene's shader count");
    if (emission_scale && n_emission != (int64_t)s->light_objects.size())
        return fail(PSY_ERR_INVALID, "psy_scene_set_params: emission count differs from the scene's light objects");
    if (albedo)
        for (int64_t i = 0; i < n_albedo; ++i)
            if (!(albedo[i] >= 0.0f && albedo[i] < 3.402823466e+38f))
                return fail(PSY_ERR_INVALID, "psy_scene_set_params: albedos must be finite and >= 0");
    if (emission_scale)
        for (int64_t j = 0; j < n_emission; ++j)
            if (!(emission_scale[j] >= 0.0f && emission_scale[j] < 3.402823466e+38f))
                return fail(PSY_ERR_INVALID, "psy_scene_set_params: emission scales must be finite and >= 0");
    HIP_OK(hipSetDevice(s->device));
    if (albedo) {
        for (int64_t i = 0; i < n_albedo; ++i)
            if (s->host_shaders[i].kind == PSY_SHADER_LAMBERT)
                for (int k = 0; k < 3; ++k) s->host_shaders[i].color[k] = albedo[i];
        if (!s->host_shaders.empty())
            HIP_OK(hipMemcpy(s->shaders.p, s->host_shaders.data(), s->host_shaders.size() * sizeof(DevShader),
                             hipMemcpyHostToDevice));
    }
    if (emission_scale) {
        for (int64_t j = 0; j < n_emission; ++j) s->escale[s->light_objects[j]] = emission_scale[j];
        if (!s->escale.empty())
            HIP_OK(hipMemcpy(s->escale_dev.p, s->escale.data(), s->escale.size() * sizeof(float),
                             hipMemcpyHostToDevice));
    }
    return PSY_OK;
}

int psy_scene_destroy(psy_scene* s) {
    if (!s) return PSY_OK;
    // a failure here must not linger as the thread's last HIP error
    if (hipSetDevice(s->device) != hipSuccess) (void)hipGetLastError();
    delete s;
    return PSY_OK;
}

int psy_scene_get_info(const psy_scene* s, psy_scene_info* info) {
    if (!s || !info) return fail(PSY_ERR_INVALID, "psy_scene_get_info: NULL argument");
    *info = s->info;
    return PSY_OK;
}

int psy_scene_instance_ids(const psy_scene* s, uint64_t* out, int64_t n) {
    if (!s || !out) return fail(PSY_ERR_INVALID, "psy_scene_instance_ids: NULL argument");
    if (n < (int64_t)s->ids.size()) return fail(PSY_ERR_INVALID, "psy_scene_instance_ids: output too small");
    std::memcpy(out, s->ids.data(), s->ids.size() * sizeof(uint64_t));
    return PSY_OK;
}

static int check_hits(const psy_hits* h, int64_t n) {
    if (!h) return fail(PSY_ERR_INVALID, "hit buffer struct is NULL");
    if (n > 0 && (!h->t || !h->inst || !h->u || !h->v))
        return fail(PSY_ERR_INVALID, "hit buffers must be non-NULL (aux may be NULL)");
    return PSY_OK;
}

static DevHits to_dev(const psy_hits* h) { return DevHits{h->t, h->inst, h->u, h->v, h->aux}; }

int psy_trace_world(psy_ctx* ctx, const psy_scene* s, const float* rays, int64_t n, const psy_hits* hits,
                    uint32_t flags, void* stream) {
    if (!ctx || !s || (!rays && n > 0)) return fail(PSY_ERR_INVALID, "psy_trace_world: NULL argument");
    if (n < 0) return fail(PSY_ERR_INVALID, "psy_trace_world: negative ray count");
    if (n > (int64_t)INT32_MAX) return fail(PSY_ERR_INVALID, "psy_trace_world: more than 2^31-1 rays in one call");
    if (int rc = check_hits(hits, n)) return rc;
    HIP_OK(hipSetDevice(ctx->device));
    if (int rc = ensure_queue(ctx, n)) return rc;
    const hipStream_t st = (hipStream_t)stream;
    if (int rc = scratch_order(ctx, st)) return rc;
    int rc = launch_trace(s->dev, true, (flags & PSY_TRACE_OCCLUSION_ONLY) != 0, s->stack_need, rays, n, n, to_dev(hits),
                          (uint32_t*)ctx->queue.p, (uint32_t*)ctx->counters.p, ctx->tuning(true), st);
    if (rc) return fail(rc, "psy_trace_world: launch failed");
    return scratch_release(ctx, st);
}

int psy_trace_rays(psy_ctx* ctx, const psy_scene* s, const float* rays, int64_t n, const psy_hits* hits,
                   uint32_t flags, void* stream) {
    if (!ctx || !s || (!rays && n > 0)) return fail(PSY_ERR_INVALID, "psy_trace_rays: NULL argument");
    if (n < 0) return fail(PSY_ERR_INVALID, "psy_trace_rays: negative ray count");
    if (n > (int64_t)INT32_MAX) return fail(PSY_ERR_INVALID, "psy_trace_rays: more than 2^31-1 rays in one call");
    if (s->transformed_patches)
        return fail(PSY_ERR_UNSUPPORTED, "psy_trace_rays: scene has transformed patch instances; their footprint needs world rays (psy_trace_world)");
    if (int rc = check_hits(hits, n)) return rc;
    HIP_OK(hipSetDevice(ctx->device));
    if (int rc = ensure_queue(ctx, n)) return rc;
    const hipStream_t st = (hipStream_t)stream;
    if (int rc = scratch_order(ctx, st)) return rc;
    int rc = launch_trace(s->dev, false, (flags & PSY_TRACE_OCCLUSION_ONLY) != 0, s->stack_need, rays, n, n,
                          to_dev(hits), (uint32_t*)ctx->queue.p, (uint32_t*)ctx->counters.p, ctx->tuning(false), st);
    if (rc) return fail(rc, "psy_trace_rays: launch failed");
    return scratch_release(ctx, st);
}

int psy_trace_rays_static(psy_ctx* ctx, const psy_scene* s, const float* rays, int64_t n, const psy_hits* hits,
                          uint32_t flags, void* stream) {
    if (!ctx || !s || (!rays && n > 0)) return fail(PSY_ERR_INVALID, "psy_trace_rays_static: NULL argument");
    if (n < 0) return fail(PSY_ERR_INVALID, "psy_trace_rays_static: negative ray count");
    if (n > (int64_t)PSY_STATIC_MAX_RAYS)
        return fail(PSY_ERR_INVALID, "psy_trace_rays_static: " + std::to_string(n) + " rays; at most " +
                                         std::to_string((int64_t)PSY_STATIC_MAX_RAYS) +
                                         " (PSY_STATIC_MAX_RAYS) per call");
    if (s->transformed_patches)
        return fail(PSY_ERR_UNSUPPORTED, "psy_trace_rays_static: scene has transformed patch instances; their footprint needs world rays (psy_trace_world)");
    if (int rc = check_hits(hits, n)) return rc;
    if (!ctx->static_scratch) return fail(PSY_ERR_HIP, "psy_trace_rays_static: no static scratch on this device");
    HIP_OK(hipSetDevice(ctx->device));
    // 6 * PSY_STATIC_MAX_RAYS words for an n-ray call's queues, deep-node
    // lists and pool (laid out from queue by 6 n words), then the counters,
    // which the trace's first kernel clears
    uint32_t* queue = ctx->static_scratch;
    uint32_t* counters = ctx->static_scratch + 6 * (int64_t)PSY_STATIC_MAX_RAYS;
    psy::TraceTuning tune = ctx->tuning(false, false);
    tune.cont = nullptr;  // no continuation buffer in the fixed array: waves run their rays to the end
    int rc = launch_trace(s->dev, false, (flags & PSY_TRACE_OCCLUSION_ONLY) != 0, s->stack_need, rays, n, n,
                          to_dev(hits), queue, counters, tune, (hipStream_t)stream);
    return rc ? fail(rc, "psy_trace_rays_static: launch failed") : PSY_OK;
}

int psy_intersections(psy_ctx* ctx, const psy_scene* s, const float* wrays, int64_t n, const psy_hits* hits,
                      float* out, void* stream) {
    if (!ctx || !s || ((!wrays || !out) && n > 0) || n < 0)
        return fail(PSY_ERR_INVALID, "psy_intersections: bad argument");
    if (int rc = check_hits(hits, n)) return rc;
    if (n > 0 && !hits->aux) return fail(PSY_ERR_INVALID, "psy_intersections: hits.aux is needed (patch offsets)");
    HIP_OK(hipSetDevice(ctx->device));
    int rc = launch_intersections(s->dev, wrays, n, to_dev(hits), out, (hipStream_t)stream);
    return rc ? fail(rc, "psy_intersections: launch failed") : PSY_OK;
}

int psy_world_to_rays(psy_ctx* ctx, const float* wrays, int64_t n, float* rays, void* stream) {
    if (!ctx || ((!wrays || !rays) && n > 0) || n < 0) return fail(PSY_ERR_INVALID, "psy_world_to_rays: bad argument");
    HIP_OK(hipSetDevice(ctx->device));
    int rc = launch_world_to_rays(wrays, n, rays, (hipStream_t)stream);
    return rc ? fail(rc, "psy_world_to_rays: launch failed") : PSY_OK;
}

int psy_halton(psy_ctx* ctx, uint32_t dim, const uint32_t* index, int64_t n, float* out, void* stream) {
    if (!ctx || ((!index || !out) && n > 0) || n < 0) return fail(PSY_ERR_INVALID, "psy_halton: bad argument");
    HIP_OK(hipSetDevice(ctx->device));
    int rc = launch_halton((const HaltonDim*)ctx->halton_dims.p, (const uint16_t*)ctx->halton_perm.p, dim, index, n,
                           out, (hipStream_t)stream);
    return rc ? fail(rc, "psy_halton: launch failed") : PSY_OK;
}

int psy_sobol_set_matrices(psy_ctx* ctx, const uint64_t* matrices, uint32_t num_dims) {
    if (!ctx || !matrices || num_dims == 0 || num_dims > (1u << 20))
        return fail(PSY_ERR_INVALID, "psy_sobol_set_matrices: bad argument");
    HIP_OK(hipSetDevice(ctx->device));
    std::vector<uint64_t> m(matrices, matrices + (size_t)num_dims * 52u);
    for (uint64_t v : m)
        if (v >> 52) return fail(PSY_ERR_INVALID, "psy_sobol_set_matrices: a column has bits above bit 51");
    if (int rc = ctx->sobol_m.upload(m)) return rc;
    ctx->sobol_dims = num_dims;
    return PSY_OK;
}

int psy_sobol(psy_ctx* ctx, uint32_t dim, const uint64_t* index, int64_t n, uint64_t scramble, double* out,
              void* stream) {
    if (!ctx || ((!index || !out) && n > 0) || n < 0) return fail(PSY_ERR_INVALID, "psy_sobol: bad argument");
    if (ctx->sobol_dims == 0) return fail(PSY_ERR_INVALID, "psy_sobol: no matrices (psy_sobol_set_matrices)");
    if (dim >= ctx->sobol_dims) return fail(PSY_ERR_INVALID, "psy_sobol: dimension out of range");
    HIP_OK(hipSetDevice(ctx->device));
    int rc = launch_sobol((const uint64_t*)ctx->sobol_m.p, ctx->sobol_dims * 52u, dim, index, n, scramble, out,
                          (hipStream_t)stream);
    return rc ? fail(rc, "psy_sobol: launch failed") : PSY_OK;
}

int psy_spectrum_set_grid(psy_ctx* ctx, const int32_t* cells, int32_t grid_width, int32_t grid_height,
                          const float* points, int32_t n_points, int32_t n_samples, float sample_min,
                          float sample_max, const float* mat_xy_to_uv) {
    if (!ctx) return fail(PSY_ERR_INVALID, "psy_spectrum_set_grid: NULL context");
    HIP_OK(hipSetDevice(ctx->device));
    if (!cells) {
        ctx->spec = DevSpectrum{};
        return PSY_OK;
    }
    if (grid_width < 1 || grid_height < 1 || grid_width > 4096 || grid_height > 4096 || n_points < 1 ||
        n_samples < 2 || n_samples > 4096 || !points || !mat_xy_to_uv || !(sample_min < sample_max) ||
        !(sample_min <= 380.0f) || !(sample_max >= 700.0f))
        return fail(PSY_ERR_INVALID, "psy_spectrum_set_grid: bad sizes (the grid must cover 380..700 nm)");
    const int64_t n_cells = (int64_t)grid_width * grid_height;
    for (int64_t c = 0; c < n_cells; ++c) {
        const int32_t* e = cells + 8 * c;
        const int inside = e[0], num = e[1];
        if (inside < 0 || inside > 1 || num < 0 || num > 6 || (inside && num != 4))
            return fail(PSY_ERR_INVALID, "psy_spectrum_set_grid: bad cell (inside, num_points)");
        for (int j = 0; j < num; ++j)
            if (e[2 + j] < 0 || e[2 + j] >= n_points)
                return fail(PSY_ERR_INVALID, "psy_spectrum_set_grid: cell point index out of range");
    }
    for (int k = 0; k < 6; ++k)
        if (!std::isfinite(mat_xy_to_uv[k])) return fail(PSY_ERR_INVALID, "psy_spectrum_set_grid: bad matrix");
    if (int rc = ctx->spec_cells.upload(cells, sizeof(int32_t) * 8 * (size_t)n_cells)) return rc;
    if (int rc = ctx->spec_points.upload(points, sizeof(float) * (size_t)(4 + n_samples) * (size_t)n_points))
        return rc;
    DevSpectrum& S = ctx->spec;
    S.cells = (const int32_t*)ctx->spec_cells.p;
    S.points = (const float*)ctx->spec_points.p;
    S.w = grid_width;
    S.h = grid_height;
    S.n_points = n_points;
    S.n_samples = n_samples;
    S.smin = sample_min;
    S.smax = sample_max;
    for (int k = 0; k < 6; ++k) S.m[k] = mat_xy_to_uv[k];
    return PSY_OK;
}

int psy_color_to_spectrum(psy_ctx* ctx, const float* rgb, const float* hero_wavelength, int64_t n, float* out,
                          void* stream) {
    if (!ctx || n < 0 || ((!rgb || !hero_wavelength || !out) && n > 0))
        return fail(PSY_ERR_INVALID, "psy_color_to_spectrum: bad argument");
    HIP_OK(hipSetDevice(ctx->device));
    int rc = launch_color_spectrum(ctx->spec, rgb, hero_wavelength, n, out, (hipStream_t)stream);
    return rc ? fail(rc, "psy_color_to_spectrum: launch failed") : PSY_OK;
}

int psy_pixel_samples(psy_ctx* ctx, int32_t x0, int32_t y0, int32_t w, int32_t h, int32_t spp, uint32_t seed,
                      uint32_t dim0, int32_t n_dims, float* out, void* stream) {
    if (!ctx || w < 0 || h < 0 || spp < 1 || n_dims < 0 || (!out && (int64_t)w * h * n_dims > 0))
        return fail(PSY_ERR_INVALID, "psy_pixel_samples: bad argument");
    HIP_OK(hipSetDevice(ctx->device));
    int rc = launch_pixel_samples((const HaltonDim*)ctx->halton_dims.p, (const uint16_t*)ctx->halton_perm.p, x0, y0, w,
                                  h, spp, seed, dim0, n_dims, out, (hipStream_t)stream);
    return rc ? fail(rc, "psy_pixel_samples: launch failed") : PSY_OK;
}

int psy_camera_rays(psy_ctx* ctx, const psy_scene* s, int32_t width, int32_t height, int32_t spp, uint32_t seed,
                    float* wrays, float* wavelength, void* stream) {
    if (!ctx || !s || width <= 0 || height <= 0 || spp <= 0 || !wrays) return fail(PSY_ERR_INVALID, "psy_camera_rays: bad argument");
    if (width > 32767 || height > 32767) return fail(PSY_ERR_INVALID, "psy_camera_rays: resolution above 32767");
    HIP_OK(hipSetDevice(ctx->device));
    int rc = launch_camera_rays(s->cam, (const HaltonDim*)ctx->halton_dims.p, (const uint16_t*)ctx->halton_perm.p, width,
                                height, spp, seed, wrays, wavelength, (hipStream_t)stream);
    return rc ? fail(rc, "psy_camera_rays: launch failed") : PSY_OK;
}

int psy_bounce_rays(psy_ctx* ctx, const psy_scene* s, const float* in_wrays, const psy_hits* in_hits, int64_t n,
                    int32_t width, int32_t spp, uint32_t seed, uint32_t dim_u, uint32_t dim_v, float* out_wrays,
                    uint32_t* valid, void* stream) {
    if (!ctx || !s || n < 0 || width <= 0 || spp <= 0 || ((!in_wrays || !out_wrays || !valid) && n > 0))
        return fail(PSY_ERR_INVALID, "psy_bounce_rays: bad argument");
    if (int rc = check_hits(in_hits, n)) return rc;
    HIP_OK(hipSetDevice(ctx->device));
    int rc = launch_bounce_rays(s->dev, (const HaltonDim*)ctx->halton_dims.p, (const uint16_t*)ctx->halton_perm.p,
                                in_wrays, to_dev(in_hits), n, width, spp, seed, dim_u, dim_v, out_wrays, valid,
                                (hipStream_t)stream);
    return rc ? fail(rc, "psy_bounce_rays: launch failed") : PSY_OK;
}

}  // extern "C"
