// psy_kernels.hip -- HIP kernels for gfx950 (MI355X) and their launchers.
// The per-ray and per-path logic lives in psy_trace.hpp and
// psy_render_dev.hpp; the kernels here map it onto the grid (one lane per
// ray or path), own the LDS traversal stacks and do the wave-level parts
// (ballot compaction, atomics, persistent refill).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "psy_trace.hpp"
#include "psy_render_dev.hpp"

namespace psy {

__global__ void k_halton(const HaltonDim* dims, const uint16_t* perm, uint32_t dim, const uint32_t* idx,
                         int64_t n, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = halton_sample(dims, perm, dim, idx[i]);
}

// Sobol::sample (sampling/sobol.hpp:40-53): the XOR of the generator
// columns of `dim` selected by the set bits of the index, plus the scramble's
// low 52 bits, as a double in [0,1).  The reference indexes its flat table
// at dim*52 + bit, so index bits 52 and up read the next dimension's
// columns; columns past the end of the table read as zero here.
__global__ void k_sobol(const uint64_t* m, uint32_t n_cols, uint32_t dim, const uint64_t* idx, int64_t n,
                        uint64_t scramble, double* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = sobol_value(m, n_cols, dim, idx[i], scramble);
}

__global__ void k_pixel_samples(const HaltonDim* dims, const uint16_t* perm, int x0, int y0, int w, int h,
                                int spp, uint32_t seed, uint32_t dim0, int n_dims, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)w * h * spp;
    if (i >= total) return;
    const int s = (int)(i % spp);
    const int64_t pix = i / spp;
    const int px = x0 + (int)(pix % w);
    const int py = y0 + (int)(pix / w);
    const uint32_t off = sampler_offset((uint32_t)px, (uint32_t)py, (uint32_t)s, seed);
    for (int k = 0; k < n_dims; ++k) out[i * n_dims + k] = sampler_get(dims, perm, off, dim0 + (uint32_t)k);
}

__global__ void k_build_dice(const float* odata, const DevObject* objs, const int32_t* patch_ids, int n_patches,
                             int levels, float4* dice) {
    const int64_t per = (int64_t)1 << (levels + 1);
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= per * n_patches) return;
    const int pi = (int)(gid / per);
    const uint32_t k = (uint32_t)(gid % per);
    const DevObject ob = objs[patch_ids[pi]];
    float4* out = dice + 2 * (ob.dice + k);
    if (k == 0) {  // unused slot of the 1-based heap
        out[0] = make_float4(0.f, 0.f, 0.f, 0.f);
        out[1] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    if (ob.kind == K_BILINEAR) build_node<4>(odata, ob, k, out);
    else build_node<16>(odata, ob, k, out);
}

// Main pass, one thread per ray.  n_dev (or NULL): a device-side ray
// count, read at launch, that lowers n (streams whose length only the
// device knows, e.g. in a captured graph).  COHERENT: each wave step
// advances only the lanes of the wave's majority kind of step (split-tree
// walk, or BVH node / leaf), so a step issues one kind of work.
template <bool WORLD, int STACK, bool COHERENT>
__global__ __launch_bounds__(TRACE_BLOCK, main_waves(WORLD, STACK, COHERENT)) void k_trace(
    DevScene s, const float* __restrict__ rays, int64_t n, int64_t rs, DevHits h, uint32_t flags,
    uint32_t* __restrict__ queue, uint32_t* __restrict__ heads, uint32_t* __restrict__ counters, DeepPool pool,
    const uint32_t* __restrict__ n_dev, ContBuf cont) {
    __shared__ int32_t stk_ref[STACK][TRACE_BLOCK];
    __shared__ float stk_near[STACK][TRACE_BLOCK];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * TRACE_BLOCK + tid;
    if (n_dev) n = std::min<int64_t>(n, (int64_t)*n_dev);
    if (i >= n) return;
    Lane L;
    lane_init<WORLD>(s, rays, rs, (uint32_t)i, flags, L);
    if constexpr (!COHERENT) {
        // Rays of one wave end after very different numbers of steps; once
        // fewer than cont.below lanes are still tracing, they are written
        // to the continuation buffer and finished by k_trace_cont, packed
        // 64 to a wave.
        if (lane_run<WORLD, STACK>(s, L, stk_ref, stk_near, tid, pool, cont)) return;
    } else {
        bool done = false;
        for (;;) {
            const unsigned long long busy = __ballot(!done);
            if (busy == 0ull) break;
            const bool walk = !done && (L.state & LANE_WALKING) != 0;
            const bool walk_round = 2 * __popcll(__ballot(walk)) >= __popcll(busy);  // wave-uniform
            if (!done && walk == walk_round && lane_step<WORLD, STACK>(s, L, stk_ref, stk_near, tid, pool))
                done = true;
        }
    }
    if (lane_finish(s, L, h)) enqueue_second(queue, heads, counters, L);
}

// The continued rays of the previous launch (`in`), one per lane: entry e
// restores its lane, which runs to the end or is continued again into
// `out`.  The grid covers `in`'s capacity; lanes past its entry count exit
// at once.
template <bool WORLD, int STACK>
__global__ __launch_bounds__(TRACE_BLOCK, main_waves(WORLD, STACK)) void k_trace_cont(
    DevScene s, const float* __restrict__ rays, int64_t rs, DevHits h, uint32_t flags, uint32_t* __restrict__ queue,
    uint32_t* __restrict__ heads, uint32_t* __restrict__ counters, DeepPool pool, ContBuf in, ContBuf out) {
    __shared__ int32_t stk_ref[STACK][TRACE_BLOCK];
    __shared__ float stk_near[STACK][TRACE_BLOCK];
    const int tid = threadIdx.x;
    const uint32_t e = blockIdx.x * TRACE_BLOCK + tid;
    const uint32_t count = min(*in.count, in.cap);
    if (e >= count) return;
    Lane L;
    cont_load<WORLD, STACK>(s, rays, rs, flags, in.c, in.cap, e, L, stk_ref, stk_near, tid);
    if (lane_run<WORLD, STACK>(s, L, stk_ref, stk_near, tid, pool, out)) return;
    if (lane_finish(s, L, h)) enqueue_second(queue, heads, counters, L);
}

constexpr int PERSIST_CHUNK = 256;

template <bool WORLD, int STACK>
__global__ __launch_bounds__(TRACE_BLOCK, persist_waves(WORLD, STACK)) void k_trace_persist(
    DevScene s, const float* __restrict__ rays, int64_t n, int64_t rs, DevHits h, uint32_t flags,
    uint32_t* __restrict__ queue, uint32_t* __restrict__ heads, uint32_t* __restrict__ counters, DeepPool pool,
    int refill_min, const uint32_t* __restrict__ n_dev, int coherent) {
    __shared__ int32_t stk_ref[STACK][TRACE_BLOCK];
    __shared__ float stk_near[STACK][TRACE_BLOCK];
    const int tid = threadIdx.x;
    if (n_dev) n = std::min<int64_t>(n, (int64_t)*n_dev);
    const uint32_t lane = __lane_id();
    unsigned long long* next_ray = reinterpret_cast<unsigned long long*>(counters + 2);
    Lane L;
    bool active = false;
    bool exhausted = false;  // wave-uniform
    // the wave's claimed rays [cbase, cend): one atomic per PERSIST_CHUNK
    // rays, so the waves do not queue on the counter at every refill
    int64_t cbase = 0, cend = 0;
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (!exhausted && nidle >= refill_min) {
            if (cbase >= cend) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(next_ray, (unsigned long long)PERSIST_CHUNK);
                base = ((unsigned long long)(uint32_t)__shfl((int)(uint32_t)(base >> 32), 0) << 32) |
                       (unsigned long long)(uint32_t)__shfl((int)(uint32_t)base, 0);
                if ((int64_t)base >= n) exhausted = true;
                cbase = (int64_t)base;
                cend = std::min<int64_t>(cbase + PERSIST_CHUNK, n);
            }
            if (!exhausted) {
                const int64_t take = std::min<int64_t>(nidle, cend - cbase);
                const int rank = __popcll(idle & ((1ull << lane) - 1ull));
                if (!active && rank < take) {
                    lane_init<WORLD>(s, rays, rs, (uint32_t)(cbase + rank), flags, L);
                    active = true;
                }
                cbase += take;
            }
        }
        const unsigned long long busy = __ballot(active);
        if (exhausted && busy == 0ull) break;
        bool go = active;
        if (coherent) {
            // the majority kind of step this round (wave-uniform)
            const bool walk = active && (L.state & LANE_WALKING) != 0;
            const bool walk_round = 2 * __popcll(__ballot(walk)) >= __popcll(busy);
            go = active && walk == walk_round;
        }
        if (go && lane_step<WORLD, STACK>(s, L, stk_ref, stk_near, tid, pool)) {
            active = false;
            if (lane_finish(s, L, h)) enqueue_second(queue, heads, counters, L);
        }
    }
}

template <bool WORLD, int STACK, bool WIDE>
__global__ __launch_bounds__(TRACE_BLOCK, WIDE ? DEEP_WIDE_WAVES : 1)
void k_trace_deep(DevScene s, const float* __restrict__ rays, int64_t n, int64_t rs, DevHits h, uint32_t flags,
                  const uint32_t* __restrict__ queue, const uint32_t* __restrict__ heads,
                  const uint32_t* __restrict__ pool, uint32_t* __restrict__ counters,
                  uint32_t* __restrict__ coop_queue, uint32_t wide_min, int32_t budget) {
    __shared__ int32_t stk_ref[STACK][TRACE_BLOCK];
    __shared__ float stk_near[STACK][TRACE_BLOCK];
    const int tid = threadIdx.x;
    const uint32_t count = counters[0];
    if (WIDE != (count >= wide_min)) return;  // the other build walks this queue
    // Queue entry k goes to wave k % n_waves, lane k / n_waves: a short queue
    // puts one ray in each wave, so rays that take different paths through
    // the split trees run side by side instead of serializing in one wave.
    const uint32_t n_waves = gridDim.x * (TRACE_BLOCK / 64);
    const uint32_t wave = blockIdx.x * (TRACE_BLOCK / 64) + (uint32_t)(tid >> 6);
    const uint32_t lane = (uint32_t)(tid & 63);
    for (uint64_t k = (uint64_t)wave + (uint64_t)lane * n_waves; k < count; k += 64ull * n_waves) {
        const uint32_t e = queue[k];
        const uint32_t i = e & ~QUEUE_RESUME;
        if ((int64_t)i >= n) continue;  // entries are ray indices of this call
        // a ray that runs out of node budget goes to the third pass
        bool more;
        if (e & QUEUE_RESUME) {
            more = resume_list<WORLD>(s, rays, rs, (int64_t)i, heads[k], pool, h, flags, budget);
        } else {
            more = trace_one<WORLD, STACK, D_SERIAL>(s, rays, rs, (int64_t)i, h, flags, stk_ref, stk_near, tid, budget);
        }
        if (more) coop_queue[atomicAdd(&counters[1], 1u)] = i;
    }
}

// Third pass: the rays whose split trees outran the second pass's budget,
// one per wave, every lane of the wave on the same ray; the split-dice
// below the cache runs across the lanes from a node stack in LDS.
template <bool WORLD, int STACK>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_coop(DevScene s, const float* __restrict__ rays, int64_t n,
                                                             int64_t rs, DevHits h, uint32_t flags,
                                                             const uint32_t* __restrict__ counters,
                                                             const uint32_t* __restrict__ coop_queue) {
    __shared__ int32_t stk_ref[STACK][TRACE_BLOCK];
    __shared__ float stk_near[STACK][TRACE_BLOCK];
    __shared__ uint64_t nodes[TRACE_BLOCK / 64][COOP_CAP];
    const int tid = threadIdx.x;
    const uint32_t count = counters[1];
    const uint32_t n_waves = gridDim.x * (TRACE_BLOCK / 64);
    const uint32_t wib = (uint32_t)(tid >> 6);
    for (uint32_t k = blockIdx.x * (TRACE_BLOCK / 64) + wib; k < count; k += n_waves) {
        const uint32_t i = coop_queue[k];  // wave-uniform
        if ((int64_t)i >= n) continue;
        (void)trace_one<WORLD, STACK, D_COOP>(s, rays, rs, (int64_t)i, h, flags, stk_ref, stk_near, tid, INT32_MAX,
                                              nodes[wib]);
    }
}

// counters of a trace call: [0] second-pass queue, [1] third-pass queue,
// [2..3] next ray of the persistent form (u64), [4] deep-node pool
__global__ void k_zero_n(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_zero_counters(uint32_t* p) {
    if (threadIdx.x < 8 && blockIdx.x == 0) p[threadIdx.x] = 0u;
}

// Intersection records of the hits of world rays (PSY_ISECT_PLANES planes
// of n entries; see psy_gpu.h).  Misses get closure kind 0.
__global__ void k_intersections(DevScene s, const float* __restrict__ w, int64_t n, DevHits h,
                                float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t li = h.inst[i];
    float* o = out + i;
    if (li == 0xffffffffu) {
        // a miss keeps Intersection's defaults (basics/intersection.hpp:24-49)
        for (int p = 0; p < PSY_ISECT_PLANES; ++p) o[p * n] = 0.0f;
        for (int k = 0; k < 4; ++k) o[(21 + 5 * k) * n] = 1.0f;
        o[37 * n] = 9999.0f;
        o[43 * n] = __int_as_float(-1);
        return;
    }
    const V3 wo = mk(w[0 * n + i], w[1 * n + i], w[2 * n + i]);
    const V3 wd = mk(w[3 * n + i], w[4 * n + i], w[5 * n + i]);
    const float time = w[18 * n + i];
    IsectRec R;
    intersection_record(s, li, h.t[i], h.u[i], h.v[i], h.aux ? h.aux[i] : 0.0f, wo, wd, time, R);
    const V3 v3s[7] = {R.p, R.n, R.dpdu, R.dpdv, R.dndu, R.dndv, R.offset};
    for (int k = 0; k < 7; ++k) {
        o[(3 * k + 0) * n] = v3s[k].x;
        o[(3 * k + 1) * n] = v3s[k].y;
        o[(3 * k + 2) * n] = v3s[k].z;
    }
    for (int k = 0; k < 16; ++k) o[(21 + k) * n] = R.space.m[k];
    o[37 * n] = R.light_pdf;
    o[38 * n] = __uint_as_float(R.backfacing);
    o[39 * n] = __uint_as_float(R.closure);
    o[40 * n] = R.col[0];
    o[41 * n] = R.col[1];
    o[42 * n] = R.col[2];
    o[43 * n] = __int_as_float(R.shader);
}

__global__ void k_world_to_rays(const float* __restrict__ w, int64_t n, float* __restrict__ r) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 o = mk(w[0 * n + i], w[1 * n + i], w[2 * n + i]);
    const V3 d = mk(w[3 * n + i], w[4 * n + i], w[5 * n + i]);
    const V3 odx = mk(w[6 * n + i], w[7 * n + i], w[8 * n + i]);
    const V3 ody = mk(w[9 * n + i], w[10 * n + i], w[11 * n + i]);
    const V3 ddx = mk(w[12 * n + i], w[13 * n + i], w[14 * n + i]);
    const V3 ddy = mk(w[15 * n + i], w[16 * n + i], w[17 * n + i]);
    const float time = w[18 * n + i];
    const uint32_t type = __float_as_uint(w[19 * n + i]);
    const bool occl = type == PSY_RAY_OCCLUSION;
    const RayW rw = ray_widths(o, d, odx, ody, ddx, ddy);
    r[0 * n + i] = o.x; r[1 * n + i] = o.y; r[2 * n + i] = o.z;
    r[3 * n + i] = d.x; r[4 * n + i] = d.y; r[5 * n + i] = d.z;
    r[6 * n + i] = occl ? 1.0f : INFINITY;
    r[7 * n + i] = time;
    r[8 * n + i] = rw.owx; r[9 * n + i] = rw.owy;
    r[10 * n + i] = rw.dwx; r[11 * n + i] = rw.dwy;
    r[12 * n + i] = rw.fwx; r[13 * n + i] = rw.fwy;
    r[14 * n + i] = __uint_as_float(occl ? (1u << 30) : 0u);
}

__global__ void k_camera_rays(DevCamera cam, const HaltonDim* dims, const uint16_t* perm, int width, int height,
                              int spp, uint32_t seed, float* wrays, float* wavelength) {
    const int64_t n = (int64_t)width * height * spp;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = (int)(i % spp);
    const int64_t pix = i / spp;
    const int px = (int)(pix % width);
    const int py = (int)(pix / width);
    const uint32_t off = sampler_offset((uint32_t)px, (uint32_t)py, (uint32_t)s, seed);
    uint32_t dim = 0;
    const float wl = (sampler_get(dims, perm, off, dim++) * (700.0f - 380.0f)) + 380.0f;
    const float sx = sampler_get(dims, perm, off, dim++);
    const float sy = sampler_get(dims, perm, off, dim++);
    const float sa = sampler_get(dims, perm, off, dim++);
    const float sb = sampler_get(dims, perm, off, dim++);
    const float time = sampler_get(dims, perm, off, dim++);
    V3 o, d, ddx, ddy, z;
    camera_ray(cam, width, height, px, py, sx, sy, sa, sb, time, &o, &d, &ddx, &ddy, &z);
    store_wray(wrays, n, i, o, d, z, z, ddx, ddy, time, 1u /* WorldRay::CAMERA */);
    if (wavelength) wavelength[i] = wl;
}

__global__ void k_bounce_rays(DevScene s, const HaltonDim* dims, const uint16_t* perm, const float* __restrict__ in,
                              DevHits hits, int64_t n, int width, int spp, uint32_t seed, uint32_t dim_u,
                              uint32_t dim_v, float* out, uint32_t* valid) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t li = hits.inst[i];
    const V3 zero = splat(0.0f);
    bool ok = li != 0xffffffffu;
    int shader = -1;
    if (ok) {
        shader = s.inst[li].shader;
        ok = shader >= 0 && s.shaders[shader].kind == PSY_SHADER_LAMBERT;
    }
    const V3 wo = mk(in[0 * n + i], in[1 * n + i], in[2 * n + i]);
    const V3 wd = mk(in[3 * n + i], in[4 * n + i], in[5 * n + i]);
    const float time = in[18 * n + i];
    M44 space;
    V3 pl, nl, offl;
    bool backfacing = false;
    if (ok) ok = hit_geometry(s, li, hits.t[i], hits.u[i], hits.v[i], hits.aux ? hits.aux[i] : 0.0f, wo, wd,
                              time, &space, &pl, &nl, &offl, &backfacing);
    if (!ok) {
        valid[i] = 0;
        store_wray(out, n, i, zero, zero, zero, zero, zero, zero, time, 2u);
        return;
    }
    valid[i] = 1;
    // DifferentialGeometry::transformed_from(space)
    const M44 inv = m_inverse(space);
    const V3 gp = m_pos(inv, pl);
    V3 gn = m_dir_transpose(space, nl);
    const float il = 1.0f / length(gn);
    gn = scale(gn, il);
    V3 pos_offset = m_dir(inv, offl);
    if (backfacing) pos_offset = scale(pos_offset, -1.0f);
    // LambertClosure::sample
    V3 nn = normalized(gn);
    if (dot(nn, wd) > 0.0f) nn = scale(nn, -1.0f);
    const int sidx = (int)(i % spp);
    const int64_t pix = i / spp;
    const uint32_t off = sampler_offset((uint32_t)(pix % width), (uint32_t)(pix / width), (uint32_t)sidx, seed);
    float su = sampler_get(dims, perm, off, dim_u);
    float sv = sampler_get(dims, perm, off, dim_v);
    su = (su * 2) - 1;
    sv = (sv * 2) - 1;
    square_to_circle(&su, &sv);
    const double zz = 1.0 - (double)((su * su) + (sv * sv));
    const float z = (float)sqrt(zz > 0.0 ? zz : 0.0);
    const V3 outd = zup_to_vec(mk(su, sv, z), nn);
    // LambertClosure::propagate_differentials
    const float len = length(outd);
    const V3 gnn = normalized(gn);
    V3 x, y;
    coordinate_system_from_vec3(outd, &x, &y);
    x = normalized(x);
    y = normalized(y);
    const V3 podx = mk(in[6 * n + i], in[7 * n + i], in[8 * n + i]);
    const V3 pody = mk(in[9 * n + i], in[10 * n + i], in[11 * n + i]);
    const V3 pddx = mk(in[12 * n + i], in[13 * n + i], in[14 * n + i]);
    const V3 pddy = mk(in[15 * n + i], in[16 * n + i], in[17 * n + i]);
    const float t = hits.t[i];
    store_wray(out, n, i, add(gp, pos_offset), outd, transfer_origin_diff(t, gnn, wd, podx, pddx),
               transfer_origin_diff(t, gnn, wd, pody, pddy), sdiv(scale(x, 0.15f), len),
               sdiv(scale(y, 0.15f), len), time, 2u /* WorldRay::R_DIFFUSE */);
}

// ---------------------------------------------------------------------------
// Path tracer kernels (per-path logic in psy_render_dev.hpp)
// ---------------------------------------------------------------------------
constexpr int RENDER_BLOCK = 256;

__global__ __launch_bounds__(RENDER_BLOCK) void k_render_init(RenderLaunch A, int64_t first, int64_t n, PathBuf P,
                                                              float* __restrict__ rays, int64_t rcap,
                                                              uint32_t* __restrict__ act) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) render_init_one(A, first, i, P, rays, rcap, act);
}

// Paths that go on are appended to act_out / rays_out with one atomic per
// wave (ballot + prefix count): the stream compaction that replaces the
// reference's std::partition of the path array.  Every lane of the wave
// calls it.
__device__ __forceinline__ void compact_push(bool alive, uint32_t i, const WRay& nr, uint32_t* __restrict__ act_out,
                                             float* __restrict__ rays_out, int64_t rcap,
                                             uint32_t* __restrict__ count_out) {
    const unsigned long long m = __ballot(alive);
    const uint32_t lane = __lane_id();
    uint32_t base = 0;
    if (lane == 0 && m) base = atomicAdd(count_out, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, 0);
    if (alive) {
        const uint32_t k = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        act_out[k] = i;
        store_wray(rays_out, rcap, k, nr.o, nr.d, nr.odx, nr.ody, nr.ddx, nr.ddy, nr.time, nr.type);
    }
}

// The shading step of one stream slot.  PAR: the step parity of every path
// in the stream (0 or 1), or -1 (mixed).
template <int PAR>
__device__ __forceinline__ void render_update_slot(const RenderLaunch& A, int64_t first, const PathBuf& P,
                                                   const uint32_t* __restrict__ act_in,
                                                   const float* __restrict__ rays_in, const DevHits& hits,
                                                   int64_t n_in, int64_t rcap, uint32_t* __restrict__ act_out,
                                                   float* __restrict__ rays_out, uint32_t* __restrict__ count_out,
                                                   float* __restrict__ sample_xyz, int64_t sbase) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool alive = false;
    uint32_t i = 0;
    WRay nr;
    if (j < n_in) {
        i = act_in[j];
        alive = render_update_one<NoAdjoint, PAR>(A, first, P, i, j, rays_in, hits, rcap, sample_xyz, sbase, nr);
    }
    compact_push(alive, i, nr, act_out, rays_out, rcap, count_out);
}

template <int PAR>
__global__ __launch_bounds__(RENDER_BLOCK) void k_render_update(RenderLaunch A, int64_t first, PathBuf P,
                                                                const uint32_t* __restrict__ act_in,
                                                                const float* __restrict__ rays_in, DevHits hits,
                                                                int64_t n_in, int64_t rcap,
                                                                uint32_t* __restrict__ act_out,
                                                                float* __restrict__ rays_out,
                                                                uint32_t* __restrict__ count_out,
                                                                float* __restrict__ sample_xyz, int64_t sbase) {
    render_update_slot<PAR>(A, first, P, act_in, rays_in, hits, n_in, rcap, act_out, rays_out, count_out, sample_xyz,
                            sbase);
}

// The even (hit, then light sample) step with registers held to 128 for 4
// waves per SIMD, spilling the rest to scratch: the light tree's descent
// is a chain of dependent loads, which more waves hide where a scene has
// many lights.  Same arithmetic as k_render_update<0>.
__global__ __launch_bounds__(RENDER_BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_render_update_w4(
    RenderLaunch A, int64_t first, PathBuf P, const uint32_t* __restrict__ act_in, const float* __restrict__ rays_in,
    DevHits hits, int64_t n_in, int64_t rcap, uint32_t* __restrict__ act_out, float* __restrict__ rays_out,
    uint32_t* __restrict__ count_out, float* __restrict__ sample_xyz, int64_t sbase) {
    render_update_slot<0>(A, first, P, act_in, rays_in, hits, n_in, rcap, act_out, rays_out, count_out, sample_xyz,
                          sbase);
}

// ---- stream form ---------------------------------------------------------------
// A pool of `cap` path slots works through the samples [first, end) of the
// image: the shading step hands a finished path's slot to the next sample
// (one atomic per wave), whose camera ray joins the same compacted stream.
// So one stream mixes the paths of many image tiles at every bounce depth,
// it stays full until the samples run out, and each step's length lives
// only on the device (the trace reads it there), which lets the host capture
// a run of steps as one graph.
// Paths start only on even steps (cur == 0 streams), so every path's step
// parity is the stream's: even streams hold camera and bounce rays (closest
// hit), odd streams shadow rays (any hit) -- no wave mixes the two kinds.  A
// path that ends after an even step parks its slot on the wait list for one
// step; the odd step's shading refills finished and parked slots alike.
__global__ void k_stream_init(RenderLaunch A, PathBuf P, int64_t first, int64_t n, int64_t end,
                              float* __restrict__ rays, int64_t rcap, uint32_t* __restrict__ act,
                              StreamState* __restrict__ S, int reset_rays) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) render_init_one(A, first, i, P, rays, rcap, act);
    if (i == 0) {
        S->count[0] = (uint32_t)n;
        S->count[1] = 0u;
        S->n_wait = 0u;
        S->next = (unsigned long long)(first + n);
        S->end = (unsigned long long)end;
        if (reset_rays) S->rays = 0ull;
    }
}

__global__ __launch_bounds__(RENDER_BLOCK) void k_stream_shade(RenderLaunch A, PathBuf P, int cur,
                                                               const uint32_t* __restrict__ act_in,
                                                               const float* __restrict__ rays_in, DevHits hits,
                                                               int64_t rcap, uint32_t* __restrict__ act_out,
                                                               float* __restrict__ rays_out,
                                                               uint32_t* __restrict__ wait,
                                                               StreamState* __restrict__ S,
                                                               float* __restrict__ sample_xyz, int64_t sbase) {
    const int64_t n_in = (int64_t)S->count[cur];
    // odd step: the parked slots are refilled too (entries past the stream)
    const int64_t n_wait = cur ? (int64_t)S->n_wait : 0;
    const int64_t n_all = n_in + n_wait;
    const int64_t end = (int64_t)S->end;
    if (blockIdx.x == 0 && threadIdx.x == 0) S->rays += (unsigned long long)n_in;
    uint32_t* count_out = &S->count[cur ^ 1];
    const uint32_t lane = __lane_id();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j0 = (int64_t)blockIdx.x * blockDim.x; j0 < n_all; j0 += stride) {
        const int64_t j = j0 + threadIdx.x;
        bool alive = false;
        uint32_t i = 0;
        WRay nr;
        if (j < n_in) {
            i = act_in[j];
            alive = render_update_one(A, 0, P, i, j, rays_in, hits, rcap, sample_xyz, sbase, nr);
        } else if (j < n_all) {
            i = wait[j - n_in];
        }
        bool want = j < n_all && !alive;
        if (!cur) {
            // even step: finished slots wait for the odd step
            const unsigned long long mw = __ballot(want);
            uint32_t wb = 0;
            if (lane == 0 && mw) wb = atomicAdd(&S->n_wait, (uint32_t)__popcll(mw));
            wb = (uint32_t)__shfl((int)wb, 0);
            if (want) wait[wb + (uint32_t)__popcll(mw & ((1ull << lane) - 1ull))] = i;
            want = false;
        }
        const unsigned long long m = __ballot(want);
        unsigned long long base = 0;
        if (lane == 0 && m) base = atomicAdd(&S->next, (unsigned long long)__popcll(m));
        base = ((unsigned long long)(uint32_t)__shfl((int)(uint32_t)(base >> 32), 0) << 32) |
               (unsigned long long)(uint32_t)__shfl((int)(uint32_t)base, 0);
        if (want) {
            const int64_t g = (int64_t)base + __popcll(m & ((1ull << lane) - 1ull));
            if (g < end) {
                render_init_path(A, g, i, P, nr);
                alive = true;
            }
        }
        compact_push(alive, i, nr, act_out, rays_out, rcap, count_out);
    }
}

// ---- adjoint ----------------------------------------------------------------
// Gradient accumulator of the adjoint pass: float atomics into a block's LDS
// row of n_params sums, or straight into the global gradient vector.
struct AdjAcc {
    static constexpr bool on = true;
    PathBuf Q;
    int32_t n_albedo;
    const int32_t* lparam;  // per object: light parameter index or -1
    float* acc;
    __device__ int light_param(int obj) const { return lparam[obj]; }
    __device__ void add(int p, float v) const { atomicAdd(acc + p, v); }
};

__global__ __launch_bounds__(RENDER_BLOCK) void k_adjoint_init(RenderLaunch A, PathBuf P, PathBuf Q, int64_t n,
                                                               const float* __restrict__ d_film) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) adjoint_init_one(A, P, Q, i, d_film);
}

// The shading step of the adjoint pass: update_path + next_ray_for_path for
// every active path (the forward step, identical decisions) plus the
// gradients of its contributions.  A fixed grid walks the stream; with LDS
// sums each block adds its row into partials[blockIdx.x] at the end (rows are
// owned by block index, and launches on one stream do not overlap).  PAR:
// the step parity of every path in the stream, as in k_render_update.
// Held to 2 waves per SIMD: the even step otherwise takes 256 VGPRs and 24
// AGPRs (1 wave); at 256 registers it spills 80-112 B per lane and runs
// the C4 adjoint in 8.9 ms against 10.8.
template <bool LDS, int PAR>
__global__ __launch_bounds__(RENDER_BLOCK) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_render_adjoint(
    RenderLaunch A, int64_t first, PathBuf P, PathBuf Q, const uint32_t* __restrict__ act_in,
    const float* __restrict__ rays_in, DevHits hits, int64_t n_in, int64_t rcap, uint32_t* __restrict__ act_out,
    float* __restrict__ rays_out, uint32_t* __restrict__ count_out, int32_t n_albedo, int32_t n_params,
    const int32_t* __restrict__ lparam, float* __restrict__ grad, float* __restrict__ partials) {
    extern __shared__ float sacc[];
    AdjAcc acc{Q, n_albedo, lparam, LDS ? sacc : grad};
    if (LDS) {
        for (int p = threadIdx.x; p < n_params; p += blockDim.x) sacc[p] = 0.0f;
        __syncthreads();
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j0 = (int64_t)blockIdx.x * blockDim.x; j0 < n_in; j0 += stride) {
        const int64_t j = j0 + threadIdx.x;
        bool alive = false;
        uint32_t i = 0;
        WRay nr;
        if (j < n_in) {
            i = act_in[j];
            alive = render_update_one<AdjAcc, PAR>(A, first, P, i, j, rays_in, hits, rcap, nullptr, 0, nr, acc);
        }
        compact_push(alive, i, nr, act_out, rays_out, rcap, count_out);
    }
    if (LDS) {
        __syncthreads();
        float* row = partials + (int64_t)blockIdx.x * n_params;
        for (int p = threadIdx.x; p < n_params; p += blockDim.x) row[p] += sacc[p];
    }
}

// d_albedo[p] / d_emission[p - n_albedo] = grad[p] + the sum of the rows.
__global__ void k_adjoint_reduce(const float* __restrict__ grad, const float* __restrict__ partials, int rows,
                                 int n_albedo, int n_params, float* __restrict__ d_albedo,
                                 float* __restrict__ d_emission) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_params) return;
    float v = grad[p];
    for (int r = 0; r < rows; ++r) v += partials[(int64_t)r * n_params + p];
    if (p < n_albedo) {
        if (d_albedo) d_albedo[p] = v;
    } else if (d_emission) {
        d_emission[p - n_albedo] = v;
    }
}

__global__ void k_fill_f32(float* p, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Fused path kernel (the batched form of the integrator): persistent waves
// in which every lane owns one path at a time and runs trace -> shade ->
// next ray itself, with no stream in memory between the steps.  A wave
// refills its idle lanes (one atomic per wave) from one index space: first
// the `n_resume` paths handed back by the previous round (their next ray in
// resume_rays), then the chunk's new paths.  A ray that needs the second
// trace pass (splits below the dice cache, a deep BVH stack) parks its path
// in the defer list; the host traces those rays with the two-pass tracer,
// shades them and hands the survivors to the next round.
// counters: [0..1] next index (u64), [2] defer count, [3] n_resume.
__global__ __launch_bounds__(TRACE_BLOCK) void k_render_persist(RenderLaunch A, int64_t first, int64_t n_new,
                                                                PathBuf P, uint32_t* __restrict__ counters,
                                                                const uint32_t* __restrict__ resume_ids,
                                                                const float* __restrict__ resume_rays,
                                                                uint32_t* __restrict__ defer_ids,
                                                                float* __restrict__ defer_rays, int64_t rcap,
                                                                float* __restrict__ sample_xyz, int64_t sbase,
                                                                int refill_min) {
    __shared__ int32_t stk_ref[MAIN_STACK][TRACE_BLOCK];
    __shared__ float stk_near[MAIN_STACK][TRACE_BLOCK];
    const int tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    unsigned long long* next = reinterpret_cast<unsigned long long*>(counters);
    const int64_t n_resume = (int64_t)counters[3];
    const int64_t n_total = n_resume + n_new;
    const DevScene& s = A.s;
    Lane L;
    uint32_t pi = 0, rtype = 0;
    bool active = false;
    bool exhausted = false;  // wave-uniform
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (!exhausted && nidle >= refill_min) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(next, (unsigned long long)nidle);
            base = ((unsigned long long)(uint32_t)__shfl((int)(uint32_t)(base >> 32), 0) << 32) |
                   (unsigned long long)(uint32_t)__shfl((int)(uint32_t)base, 0);
            if ((int64_t)base >= n_total) exhausted = true;
            if (!active) {
                const int64_t e = (int64_t)base + __popcll(idle & ((1ull << lane) - 1ull));
                if (e < n_total) {
                    WRay w;
                    if (e < n_resume) {
                        pi = resume_ids[e];
                        load_wray(resume_rays, rcap, e, w);
                    } else {
                        pi = (uint32_t)(e - n_resume);
                        render_init_path(A, first + pi, pi, P, w);
                    }
                    rtype = w.type;
                    lane_init_wray(s, w, pi, L);
                    active = true;
                }
            }
        }
        if (exhausted && __ballot(active) == 0ull) break;
        // no deep-node pool: a ray that needs splits below the cache is
        // handed back whole and traced by the second pass
        if (active && lane_step<true>(s, L, stk_ref, stk_near, tid, DeepPool{nullptr, nullptr, 0u})) {
            const bool again = lane_result(s, L);
            WRay w;
            lane_wray(L, rtype, w);
            if (again) {
                const uint32_t k = atomicAdd(&counters[2], 1u);
                defer_ids[k] = pi;
                store_wray(defer_rays, rcap, k, w);
                active = false;
            } else {
                HitRec hr;
                hr.t = L.b.t;
                hr.inst = L.b.inst;
                hr.u = L.b.u;
                hr.v = L.b.v;
                hr.aux = L.b.aux;
                WRay nr;
                if (render_update_path(A, first, P, pi, w, hr, sample_xyz, sbase, nr, NoAdjoint())) {
                    rtype = nr.type;
                    lane_init_wray(s, nr, pi, L);
                } else {
                    active = false;
                }
            }
        }
    }
}

// Rays traced by a chunk's paths: every step of a path traced one ray.
__global__ void k_render_count_steps(PathBuf P, int64_t n, unsigned long long* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t steps = i < n ? (P.u(P_STEP, i) & 0xffu) : 0u;
    for (int o = 32; o > 0; o >>= 1) steps += (uint32_t)__shfl_xor((int)steps, o);
    if (__lane_id() == 0 && steps) atomicAdd(out, (unsigned long long)steps);
}

__global__ void k_set_counters(uint32_t* p, uint32_t n_resume) {
    if (threadIdx.x < 8 && blockIdx.x == 0) p[threadIdx.x] = threadIdx.x == 3 ? n_resume : 0u;
}

__global__ void k_render_resolve(const float* __restrict__ sample_xyz, int64_t sbase, int64_t first_pixel,
                                 int64_t n_pixels, int spp, float* __restrict__ film) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n_pixels) render_resolve_one(sample_xyz, sbase, first_pixel, p, spp, film);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

static int check_launch() { return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP; }

int launch_build_dice(const float* odata, const DevObject* objs, const int32_t* patch_ids, int n_patches,
                      int levels, float4* dice, hipStream_t st) {
    const int64_t per = (int64_t)1 << (levels + 1);
    const int64_t total = per * n_patches;
    if (total <= 0) return 0;
    hipLaunchKernelGGL(k_build_dice, dim3(grid_for(total, 256)), dim3(256), 0, st, odata, objs, patch_ids,
                       n_patches, levels, dice);
    return check_launch();
}

template <bool W, int S>
static void launch_deep(const DevScene& s, const float* rays, int64_t n, int64_t rs, const DevHits& h, uint32_t flags,
                        const uint32_t* queue, const uint32_t* heads, const uint32_t* pool, uint32_t* counters,
                        const TraceTuning& tune, hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<int64_t>(grid_for(n, TRACE_BLOCK), DEEP_ROUND);
    uint32_t* coop_queue = const_cast<uint32_t*>(queue) + n;
    const int32_t budget = std::max<int32_t>(1, tune.deep_budget);
    if (tune.wide_min > 0)
        hipLaunchKernelGGL((k_trace_deep<W, S, false>), dim3(blocks), dim3(TRACE_BLOCK), 0, st, s, rays, n, rs, h,
                           flags, queue, heads, pool, counters, coop_queue, tune.wide_min, budget);
    if (n >= (int64_t)tune.wide_min) {
        const unsigned wide = (unsigned)std::min<int64_t>(grid_for(n, TRACE_BLOCK), DEEP_ROUND * DEEP_WIDE_WAVES);
        hipLaunchKernelGGL((k_trace_deep<W, S, true>), dim3(wide), dim3(TRACE_BLOCK), 0, st, s, rays, n, rs, h,
                           flags, queue, heads, pool, counters, coop_queue, tune.wide_min, budget);
    }
    hipLaunchKernelGGL((k_trace_coop<W, S>), dim3(blocks), dim3(TRACE_BLOCK), 0, st, s, rays, n, rs, h, flags,
                       (const uint32_t*)counters, (const uint32_t*)coop_queue);
}

// The main pass one thread per ray, then the rays it continued (when the
// call has a continuation buffer: tune.cont, cont_words(STACK) words per
// entry).
template <bool W, int STACK>
static void launch_plain(const DevScene& s, const float* rays, int64_t n, int64_t rs, const DevHits& h,
                         uint32_t flags, uint32_t* queue, uint32_t* heads, uint32_t* counters, const DeepPool& pool,
                         const TraceTuning& tune, int64_t blocks, hipStream_t st) {
    // Up to CONT_LEVELS continuation buffers, each after the previous in
    // tune.cont, each counted by its own counter (counters[5 + level]).
    // Fewer than `below` of a wave's 64 rays are continued, so a level
    // needs at most this many entries for a launch of m rays.
    const int below = tune.cont_below;
    auto need = [&](int64_t m) { return (m * (int64_t)(below - 1) + 63) / 64 + 64; };
    ContBuf lv[CONT_LEVELS + 1];
    int levels = 0;
    if (tune.cont && below > 1 && below <= 64) {
        int64_t used = 0, m = n;
        for (; levels < CONT_LEVELS; ++levels) {
            const int64_t cap = need(m);
            if ((used + cap) * cont_words(STACK) > tune.cont_words || cap > (int64_t)UINT32_MAX) break;
            lv[levels] = ContBuf{tune.cont + used * cont_words(STACK), (uint32_t)cap, below, counters + 5 + levels};
            used += cap;
            m = cap;
        }
    }
    lv[levels] = ContBuf{nullptr, 0u, 0, counters + 5};
    hipLaunchKernelGGL((k_trace<W, STACK, false>), dim3((unsigned)blocks), dim3(TRACE_BLOCK), 0, st, s, rays, n, rs,
                       h, flags, queue, heads, counters, pool, tune.n_dev, lv[0]);
    for (int l = 0; l < levels; ++l)
        hipLaunchKernelGGL((k_trace_cont<W, STACK>), dim3((unsigned)((lv[l].cap + TRACE_BLOCK - 1) / TRACE_BLOCK)),
                           dim3(TRACE_BLOCK), 0, st, s, rays, rs, h, flags, queue, heads, counters, pool, lv[l],
                           lv[l + 1]);
}

template <bool W>
static int launch_trace_w(const DevScene& s, int stack_need, const float* rays, int64_t n, int64_t rs, const DevHits& h,
                          uint32_t flags, uint32_t* queue, uint32_t* counters, const TraceTuning& tune,
                          hipStream_t st) {
    // scratch layout for an n-ray call (6 n words): second-pass queue (n),
    // third-pass queue (n), the second-pass rays' deep-node lists (n), the
    // deep-node pool (n entries of 3 words)
    uint32_t* heads = queue + 2 * n;
    const DeepPool pool{queue + 3 * n, counters + 4, (uint32_t)std::min<int64_t>(n, (int64_t)UINT32_MAX - 2)};
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, counters);
    if (tune.ev[0]) (void)hipEventRecord((hipEvent_t)tune.ev[0], st);
    const ContBuf no_cont{nullptr, 0u, 0, counters + 5};
    const int64_t plain_blocks = grid_for(n, TRACE_BLOCK);
    const int64_t persist_blocks = (int64_t)(tune.stack > 16 ? tune.persist_blocks32 : tune.persist_blocks);
    const int refill = std::max(1, std::min(64, tune.refill_min));
    if (tune.persistent && persist_blocks > 0 && persist_blocks < plain_blocks) {
        if (tune.stack > 16)
            hipLaunchKernelGGL((k_trace_persist<W, 32>), dim3((unsigned)persist_blocks), dim3(TRACE_BLOCK), 0, st, s,
                               rays, n, rs, h, flags, queue, heads, counters, pool, refill, tune.n_dev, tune.coherent ? 1 : 0);
        else
            hipLaunchKernelGGL((k_trace_persist<W, MAIN_STACK>), dim3((unsigned)persist_blocks), dim3(TRACE_BLOCK), 0,
                               st, s, rays, n, rs, h, flags, queue, heads, counters, pool, refill, tune.n_dev,
                               tune.coherent ? 1 : 0);
    } else if (tune.stack > 16) {
        if (tune.plain_coherent)
            hipLaunchKernelGGL((k_trace<W, 32, true>), dim3((unsigned)plain_blocks), dim3(TRACE_BLOCK), 0, st, s, rays,
                               n, rs, h, flags, queue, heads, counters, pool, tune.n_dev, no_cont);
        else
            launch_plain<W, 32>(s, rays, n, rs, h, flags, queue, heads, counters, pool, tune, plain_blocks, st);
    } else {
        if (tune.plain_coherent)
            hipLaunchKernelGGL((k_trace<W, MAIN_STACK, true>), dim3((unsigned)plain_blocks), dim3(TRACE_BLOCK), 0, st, s,
                               rays, n, rs, h, flags, queue, heads, counters, pool, tune.n_dev, no_cont);
        else
            launch_plain<W, MAIN_STACK>(s, rays, n, rs, h, flags, queue, heads, counters, pool, tune, plain_blocks, st);
    }
    if (tune.ev[1]) (void)hipEventRecord((hipEvent_t)tune.ev[1], st);
    if (stack_need <= 16) launch_deep<W, 16>(s, rays, n, rs, h, flags, queue, heads, pool.e, counters, tune, st);
    else if (stack_need <= 32) launch_deep<W, 32>(s, rays, n, rs, h, flags, queue, heads, pool.e, counters, tune, st);
    else if (stack_need <= 64) launch_deep<W, 64>(s, rays, n, rs, h, flags, queue, heads, pool.e, counters, tune, st);
    else return PSY_ERR_UNSUPPORTED;
    if (tune.ev[2]) (void)hipEventRecord((hipEvent_t)tune.ev[2], st);
    return check_launch();
}

int launch_trace(const DevScene& s, bool world, bool occl_only, int stack_need, const float* rays, int64_t n,
                 int64_t rs, const DevHits& h, uint32_t* queue, uint32_t* counters, const TraceTuning& tune, hipStream_t st) {
    if (n <= 0) return 0;
    if (stack_need > 64) return PSY_ERR_UNSUPPORTED;
    const uint32_t flags = occl_only ? PSY_TRACE_OCCLUSION_ONLY : 0u;
    return world ? launch_trace_w<true>(s, stack_need, rays, n, rs, h, flags, queue, counters, tune, st)
                 : launch_trace_w<false>(s, stack_need, rays, n, rs, h, flags, queue, counters, tune, st);
}

int trace_persist_occupancy(bool world, int stack, int* blocks_per_cu) {
    int nb = 0;
    hipError_t e;
    if (stack > 16)
        e = world ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace_persist<true, 32>, TRACE_BLOCK, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace_persist<false, 32>, TRACE_BLOCK, 0);
    else
        e = world ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace_persist<true, MAIN_STACK>, TRACE_BLOCK, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace_persist<false, MAIN_STACK>, TRACE_BLOCK,
                                                                 0);
    *blocks_per_cu = nb;
    return e == hipSuccess ? 0 : PSY_ERR_HIP;
}


int launch_zero_u32(uint32_t* p, int n, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_zero_n, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    return check_launch();
}

int launch_intersections(const DevScene& s, const float* wrays, int64_t n, const DevHits& h, float* out,
                         hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_intersections, dim3(grid_for(n, 128)), dim3(128), 0, st, s, wrays, n, h, out);
    return check_launch();
}

int launch_world_to_rays(const float* wrays, int64_t n, float* rays, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_world_to_rays, dim3(grid_for(n, 256)), dim3(256), 0, st, wrays, n, rays);
    return check_launch();
}

int launch_halton(const HaltonDim* dims, const uint16_t* perm, uint32_t dim, const uint32_t* idx, int64_t n,
                  float* out, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_halton, dim3(grid_for(n, 256)), dim3(256), 0, st, dims, perm, dim, idx, n, out);
    return check_launch();
}

int launch_sobol(const uint64_t* m, uint32_t n_cols, uint32_t dim, const uint64_t* idx, int64_t n,
                 uint64_t scramble, double* out, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_sobol, dim3(grid_for(n, 256)), dim3(256), 0, st, m, n_cols, dim, idx, n, scramble, out);
    return check_launch();
}

__global__ void k_color_spectrum(DevSpectrum S, const float* rgb, const float* hero, int64_t n, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float e[NSPEC];
    color_spec4(S, rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], hero[i], e);
    for (int k = 0; k < NSPEC; ++k) out[NSPEC * i + k] = e[k];
}

int launch_color_spectrum(const DevSpectrum& S, const float* rgb, const float* hero, int64_t n, float* out,
                          hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_color_spectrum, dim3(grid_for(n, 256)), dim3(256), 0, st, S, rgb, hero, n, out);
    return check_launch();
}

int launch_pixel_samples(const HaltonDim* dims, const uint16_t* perm, int x0, int y0, int w, int h, int spp,
                         uint32_t seed, uint32_t dim0, int n_dims, float* out, hipStream_t st) {
    const int64_t total = (int64_t)w * h * spp;
    if (total <= 0) return 0;
    hipLaunchKernelGGL(k_pixel_samples, dim3(grid_for(total, 256)), dim3(256), 0, st, dims, perm, x0, y0, w, h,
                       spp, seed, dim0, n_dims, out);
    return check_launch();
}

int launch_camera_rays(const DevCamera& cam, const HaltonDim* dims, const uint16_t* perm, int width, int height,
                       int spp, uint32_t seed, float* wrays, float* wavelength, hipStream_t st) {
    const int64_t n = (int64_t)width * height * spp;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_camera_rays, dim3(grid_for(n, 256)), dim3(256), 0, st, cam, dims, perm, width, height,
                       spp, seed, wrays, wavelength);
    return check_launch();
}

int launch_bounce_rays(const DevScene& s, const HaltonDim* dims, const uint16_t* perm, const float* in_wrays,
                       const DevHits& hits, int64_t n, int width, int spp, uint32_t seed, uint32_t dim_u,
                       uint32_t dim_v, float* out_wrays, uint32_t* valid, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_bounce_rays, dim3(grid_for(n, 256)), dim3(256), 0, st, s, dims, perm, in_wrays, hits, n,
                       width, spp, seed, dim_u, dim_v, out_wrays, valid);
    return check_launch();
}

int launch_render_init(const RenderLaunch& R, int64_t first, int64_t n, const RenderBuffers& B, hipStream_t st) {
    if (n <= 0) return 0;
    PathBuf P{B.paths, B.cap};
    hipLaunchKernelGGL(k_render_init, dim3((unsigned)((n + RENDER_BLOCK - 1) / RENDER_BLOCK)), dim3(RENDER_BLOCK), 0,
                       st, R, first, n, P, B.rays[0], B.cap, B.act[0]);
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

int launch_render_update(const RenderLaunch& R, int64_t first, const RenderBuffers& B, int cur, int64_t n_in,
                         int step, int shade_waves, float* sample_xyz, int64_t sbase, hipStream_t st) {
    if (n_in <= 0) return 0;
    PathBuf P{B.paths, B.cap};
    const int nxt = cur ^ 1;
    const dim3 grid((unsigned)((n_in + RENDER_BLOCK - 1) / RENDER_BLOCK));
    if (!(step & 1) && shade_waves == 4)
        hipLaunchKernelGGL(k_render_update_w4, grid, dim3(RENDER_BLOCK), 0, st, R, first, P,
                           (const uint32_t*)B.act[cur], (const float*)B.rays[cur], B.hits, n_in, B.cap, B.act[nxt],
                           B.rays[nxt], B.count + nxt, sample_xyz, sbase);
    else if (step & 1)
        hipLaunchKernelGGL(k_render_update<1>, grid, dim3(RENDER_BLOCK), 0, st, R, first, P,
                           (const uint32_t*)B.act[cur], (const float*)B.rays[cur], B.hits, n_in, B.cap, B.act[nxt],
                           B.rays[nxt], B.count + nxt, sample_xyz, sbase);
    else
        hipLaunchKernelGGL(k_render_update<0>, grid, dim3(RENDER_BLOCK), 0, st, R, first, P,
                           (const uint32_t*)B.act[cur], (const float*)B.rays[cur], B.hits, n_in, B.cap, B.act[nxt],
                           B.rays[nxt], B.count + nxt, sample_xyz, sbase);
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

int launch_fill_f32(float* p, int64_t n, float v, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

int launch_stream_init(const RenderLaunch& R, const RenderBuffers& B, int64_t first, int64_t n, int64_t end,
                       StreamState* S, bool reset_rays, hipStream_t st) {
    const int64_t g = std::max<int64_t>(1, (n + RENDER_BLOCK - 1) / RENDER_BLOCK);
    hipLaunchKernelGGL(k_stream_init, dim3((unsigned)g), dim3(RENDER_BLOCK), 0, st, R, PathBuf{B.paths, B.cap}, first, n,
                       end, B.rays[0], B.cap, B.act[0], S, reset_rays ? 1 : 0);
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

int launch_stream_shade(const RenderLaunch& R, const RenderBuffers& B, int cur, StreamState* S, uint32_t* wait,
                        int blocks, float* sample_xyz, int64_t sbase, hipStream_t st) {
    const int nxt = cur ^ 1;
    hipLaunchKernelGGL(k_stream_shade, dim3((unsigned)std::max(1, blocks)), dim3(RENDER_BLOCK), 0, st, R,
                       PathBuf{B.paths, B.cap}, cur, (const uint32_t*)B.act[cur], (const float*)B.rays[cur], B.hits,
                       B.cap, B.act[nxt], B.rays[nxt], wait, S, sample_xyz, sbase);
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

int launch_adjoint_init(const RenderLaunch& R, const RenderBuffers& B, const AdjointBuffers& J, int64_t n,
                        hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_adjoint_init, dim3((unsigned)((n + RENDER_BLOCK - 1) / RENDER_BLOCK)), dim3(RENDER_BLOCK), 0,
                       st, R, PathBuf{B.paths, B.cap}, PathBuf{J.q, B.cap}, n, J.d_film);
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

template <bool LDS, int PAR>
static void launch_adjoint_kernel(const RenderLaunch& R, int64_t first, const RenderBuffers& B,
                                  const AdjointBuffers& J, int cur, int64_t n_in, unsigned grid, hipStream_t st) {
    const int nxt = cur ^ 1;
    const PathBuf P{B.paths, B.cap}, Q{J.q, B.cap};
    hipLaunchKernelGGL((k_render_adjoint<LDS, PAR>), dim3(grid), dim3(RENDER_BLOCK),
                       LDS ? (size_t)J.n_params * sizeof(float) : 0, st, R, first, P, Q, (const uint32_t*)B.act[cur],
                       (const float*)B.rays[cur], B.hits, n_in, B.cap, B.act[nxt], B.rays[nxt], B.count + nxt,
                       J.n_albedo, J.n_params, J.lparam, J.grad, J.partials);
}

int launch_render_adjoint(const RenderLaunch& R, int64_t first, const RenderBuffers& B, const AdjointBuffers& J,
                          int cur, int64_t n_in, int step, hipStream_t st) {
    if (n_in <= 0) return 0;
    const int64_t want = (n_in + RENDER_BLOCK - 1) / RENDER_BLOCK;
    const unsigned grid = (unsigned)std::min<int64_t>(want, J.rows);
    if (J.lds) {
        if (step & 1) launch_adjoint_kernel<true, 1>(R, first, B, J, cur, n_in, grid, st);
        else launch_adjoint_kernel<true, 0>(R, first, B, J, cur, n_in, grid, st);
    } else {
        if (step & 1) launch_adjoint_kernel<false, 1>(R, first, B, J, cur, n_in, grid, st);
        else launch_adjoint_kernel<false, 0>(R, first, B, J, cur, n_in, grid, st);
    }
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

int launch_adjoint_reduce(const AdjointBuffers& J, float* d_albedo, float* d_emission, hipStream_t st) {
    if (J.n_params <= 0) return 0;
    hipLaunchKernelGGL(k_adjoint_reduce, dim3((unsigned)((J.n_params + 255) / 256)), dim3(256), 0, st, J.grad,
                       J.partials, J.lds ? J.rows : 0, J.n_albedo, J.n_params, d_albedo, d_emission);
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

int launch_render_persist(const RenderLaunch& R, int64_t first, int64_t n_new, const PersistBuffers& B,
                          int blocks, int refill_min, hipStream_t st) {
    PathBuf P{B.paths, B.cap};
    hipLaunchKernelGGL(k_render_persist, dim3((unsigned)blocks), dim3(TRACE_BLOCK), 0, st, R, first, n_new, P,
                       B.counters, (const uint32_t*)B.resume_ids, (const float*)B.resume_rays, B.defer_ids,
                       B.defer_rays, B.cap, B.sample_xyz, B.sbase, std::max(1, std::min(64, refill_min)));
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

int launch_render_update_io(const RenderLaunch& R, int64_t first, float* paths, int64_t cap, const uint32_t* act_in,
                            const float* rays_in, const DevHits& hits, int64_t n_in, uint32_t* act_out,
                            float* rays_out, uint32_t* count_out, float* sample_xyz, int64_t sbase, hipStream_t st) {
    if (n_in <= 0) return 0;
    PathBuf P{paths, cap};
    hipLaunchKernelGGL(k_render_update<-1>, dim3((unsigned)((n_in + RENDER_BLOCK - 1) / RENDER_BLOCK)),
                       dim3(RENDER_BLOCK), 0, st, R, first, P, act_in, rays_in, hits, n_in, cap, act_out, rays_out,
                       count_out, sample_xyz, sbase);
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

int launch_render_count_steps(float* paths, int64_t cap, int64_t n, unsigned long long* out, hipStream_t st) {
    if (n <= 0) return 0;
    PathBuf P{paths, cap};
    hipLaunchKernelGGL(k_render_count_steps, dim3((unsigned)((n + RENDER_BLOCK - 1) / RENDER_BLOCK)), dim3(RENDER_BLOCK),
                       0, st, P, n, out);
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

int launch_set_counters(uint32_t* p, uint32_t n_resume, hipStream_t st) {
    hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, st, p, n_resume);
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

int render_persist_occupancy(int* blocks_per_cu) {
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_persist, TRACE_BLOCK, 0);
    *blocks_per_cu = nb;
    return e == hipSuccess ? 0 : PSY_ERR_HIP;
}

int launch_render_resolve(const float* sample_xyz, int64_t sbase, int64_t first_pixel, int64_t n_pixels, int spp,
                          float* film, hipStream_t st) {
    if (n_pixels <= 0) return 0;
    hipLaunchKernelGGL(k_render_resolve, dim3((unsigned)((n_pixels + RENDER_BLOCK - 1) / RENDER_BLOCK)),
                       dim3(RENDER_BLOCK), 0, st, sample_xyz, sbase, first_pixel, n_pixels, spp, film);
    return hipGetLastError() == hipSuccess ? 0 : PSY_ERR_HIP;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free trace (psy_trace_rays_static): the second-
// and third-pass queues, the second-pass rays' deep-node lists and the
// deep-node pool for up to PSY_STATIC_MAX_RAYS rays (6 words a ray), then
// the pass counters, in one file-scope device array.  Each GPU the library
// runs on holds its own copy (48 MiB + 256 B) until the process ends.  In
// host code
// `scratch` names a host-side stand-in, never the device memory: the device
// address comes from hipGetSymbolAddress, once per device, below.
// ---------------------------------------------------------------------------
constexpr int64_t STATIC_COUNTER_WORDS = 64;
static __device__ __align__(256) uint32_t scratch[6 * (int64_t)PSY_STATIC_MAX_RAYS + STATIC_COUNTER_WORDS];

namespace {
constexpr int STATIC_MAX_DEVICES = 64;
std::mutex g_static_mu;
uint32_t* g_static_ptr[STATIC_MAX_DEVICES] = {};
}  // namespace

int static_scratch_address(int device, uint32_t** out) {
    *out = nullptr;
    if (device < 0 || device >= STATIC_MAX_DEVICES) return PSY_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_static_mu);
    if (!g_static_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PSY_ERR_HIP;
        g_static_ptr[device] = static_cast<uint32_t*>(p);
    }
    *out = g_static_ptr[device];
    return 0;
}


// ---- adaptive sampling film ---------------------------------------------------
// Film::add_sample and Film::variance_estimate (film/film.hpp:135-165) in
// the loop of PathTraceIntegrator::render_blocks
// (integrator/path_trace_integrator.cpp:296-376): a bucket takes passes of
// spp samples per pixel while the largest X variance estimate over its
// pixels exceeds variance_max and fewer than spp_max samples were taken.
// Every pass of the reference draws its samples from Sampler(x, y, s,
// seed) with s = 0 .. spp - 1, the same sampler offsets every time, so each
// pass adds the same samples: they are rendered once and added pass after
// pass here.  One block per bucket; per-pixel state in 10 planes of W*H
// (XYZ sum, count as uint16 bits, last gamma mean, the running variance sum).
constexpr int FILM_THREADS = 256;

PSY_DEV float film_hcol(float n) {
    if (n < 0.0f) n = 0.0f;
    return powf(n, 1.0f / 2.2f);
}
PSY_DEV float film_mmax(float a, float b) { return a > b ? a : b; }  // film/film.hpp:56-61

__global__ __launch_bounds__(FILM_THREADS) void k_film_adaptive(const float* __restrict__ samples, int W, int H,
                                                                int spp, int bs, int nbx, int spp_max, float var_max,
                                                                float* __restrict__ st, float* __restrict__ film,
                                                                float* __restrict__ var_out, int32_t* __restrict__ passes) {
    __shared__ float red[FILM_THREADS];
    const int b = blockIdx.x;
    const int x0 = (b % nbx) * bs, y0 = (b / nbx) * bs;
    const int bw = min(W - x0, bs), bh = min(H - y0, bs);
    const int npix = bw * bh;
    const int64_t P = (int64_t)W * H;
    for (int q = threadIdx.x; q < npix; q += FILM_THREADS) {
        const int64_t pix = (int64_t)(y0 + q / bw) * W + x0 + q % bw;
        for (int k = 0; k < 10; ++k) st[k * P + pix] = 0.0f;
    }
    float max_var = 9999999999.0f;  // X component of the bucket's max_variance
    float samp_it = 0.0f;
    int np = 0;
    while (max_var > var_max && samp_it < (float)spp_max) {
        float m = 0.0f;
        for (int q = threadIdx.x; q < npix; q += FILM_THREADS) {
            const int64_t pix = (int64_t)(y0 + q / bw) * W + x0 + q % bw;
            float sum[3], vp[3], vf[3];
            for (int c = 0; c < 3; ++c) {
                sum[c] = st[c * P + pix];
                vp[c] = st[(4 + c) * P + pix];
                vf[c] = st[(7 + c) * P + pix];
            }
            uint16_t acc = (uint16_t)f2u(st[3 * P + pix]);
            const float* sp = samples + pix * spp * 3;
            for (int s = 0; s < spp; ++s) {
                const float v0 = sp[3 * s], v1 = sp[3 * s + 1], v2 = sp[3 * s + 2];
                // skip NaN and infinite samples
                if (!(fabsf(v0) < INFINITY && fabsf(v1) < INFINITY && fabsf(v2) < INFINITY)) continue;
                sum[0] += v0; sum[1] += v1; sum[2] += v2;
                acc = (uint16_t)(acc + 1);
                const float k = (float)acc;
                for (int c = 0; c < 3; ++c) {
                    const float avg = film_hcol(sum[c] / k);
                    if (acc > 1) vf[c] += fabsf(vp[c] - avg) * (float)(acc - 1);
                    vp[c] = avg;
                }
            }
            for (int c = 0; c < 3; ++c) {
                st[c * P + pix] = sum[c];
                st[(4 + c) * P + pix] = vp[c];
                st[(7 + c) * P + pix] = vf[c];
            }
            st[3 * P + pix] = u2f((uint32_t)acc);
            const float ve = acc < 2 ? INFINITY : (vf[0] / (float)(acc - 1)) / (float)sqrt((double)acc);
            m = film_mmax(m, ve);
        }
        red[threadIdx.x] = m;
        __syncthreads();
        for (int d = FILM_THREADS / 2; d > 0; d >>= 1) {
            if ((int)threadIdx.x < d) red[threadIdx.x] = film_mmax(red[threadIdx.x], red[threadIdx.x + d]);
            __syncthreads();
        }
        max_var = red[0];
        __syncthreads();
        samp_it += (float)spp;
        ++np;
    }
    for (int q = threadIdx.x; q < npix; q += FILM_THREADS) {
        const int64_t pix = (int64_t)(y0 + q / bw) * W + x0 + q % bw;
        const uint16_t acc = (uint16_t)f2u(st[3 * P + pix]);
        for (int c = 0; c < 3; ++c) {
            if (film) film[pix * 3 + c] = acc ? st[c * P + pix] / (float)acc : 0.0f;
            if (var_out)
                var_out[pix * 3 + c] =
                    acc < 2 ? INFINITY : (st[(7 + c) * P + pix] / (float)(acc - 1)) / (float)sqrt((double)acc);
        }
    }
    if (threadIdx.x == 0 && passes) passes[b] = np;
}

int launch_film_adaptive(const float* samples, int W, int H, int spp, int bs, int spp_max, float var_max, float* state,
                         float* film, float* var_out, int32_t* passes, hipStream_t st) {
    const int nbx = (W + bs - 1) / bs, nby = (H + bs - 1) / bs;
    hipLaunchKernelGGL(k_film_adaptive, dim3((unsigned)(nbx * nby)), dim3(FILM_THREADS), 0, st, samples, W, H, spp,
                       bs, nbx, spp_max, var_max, state, film, var_out, passes);
    return check_launch();
}

}  // namespace psy
