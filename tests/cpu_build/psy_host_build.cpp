// psy_host_build.cpp -- TEST INFRASTRUCTURE: the device code of
// libpsy_gpu.so (psy_trace.hpp, psy_render_dev.hpp) compiled for the host
// CPU, so CPU tests can run the exact traversal and integrator logic the
// kernels run, lane by lane, against the oracle.  The product never loads
// this library; it has no CPU path.
//
// Scene preparation is the product's own host code (psy::prepare_scene,
// linked from libpsy_gpu.so); the dice cache is built here with the same
// per-node function the k_build_dice kernel runs.  Grid-level behaviour
// (wave-ballot compaction, the LDS stack per lane) is emulated one lane at
// a time, in order.
#define PSY_DEV __host__ __device__ __forceinline__
#define PSY_DEV_NOINLINE __host__ __device__ __noinline__
#define PSY_DEV_FN __host__ __device__

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../../psychopath_cpp_amd/csrc/psy_render_dev.hpp"
#include "../../psychopath_cpp_amd/csrc/psy_internal.hpp"

using namespace psy;

namespace {

struct HostBuilt {
    HostScene H;
    std::vector<float4> dice;
    DevScene s{};
    DevLights L{};
};

int build(const psy_scene_desc* d, int dice_depth, HostBuilt& B) {
    if (int rc = prepare_scene(d, dice_depth, B.H)) return rc;
    HostScene& H = B.H;
    B.dice.assign((size_t)(2 * H.per * (int64_t)H.patch_ids.size()), make_float4(0.f, 0.f, 0.f, 0.f));
    for (size_t pi = 0; pi < H.patch_ids.size(); ++pi) {
        const DevObject& ob = H.objs[H.patch_ids[pi]];
        for (int64_t k = 1; k < H.per; ++k) {
            float4* out = B.dice.data() + 2 * (ob.dice + k);
            if (ob.kind == K_BILINEAR) build_node<4>(H.odata.data(), ob, (uint32_t)k, out);
            else build_node<16>(H.odata.data(), ob, (uint32_t)k, out);
        }
    }
    DevScene& s = B.s;
    s.bvh = H.bvh.data();
    s.bvh1 = H.bvh1.empty() ? nullptr : H.bvh1.data();
    s.n_bvh = (int32_t)H.bvh.size();
    s.root = H.root;
    s.inst = H.insts.data();
    s.obj = H.objs.data();
    s.odata = H.odata.data();
    s.xf = H.xf.data();
    s.dice = B.dice.data();
    s.shaders = H.shaders.data();
    s.escale = H.escale.data();
    s.dice_rate = H.dice_rate;
    s.min_upoly = H.min_upoly;
    s.n_inst = (int32_t)H.insts.size();
    s.flags = d->flags;
    DevLights& L = B.L;
    L.nodes = H.l_nodes.data();
    L.bounds = H.l_bounds.data();
    L.asms = H.l_asms.data();
    L.inst = H.l_inst.data();
    L.xf = H.l_xf.data();
    L.sub = H.l_sub.data();
    L.n_asms = (int32_t)H.l_asms.size();
    return PSY_OK;
}

// One ray exactly as the kernels trace it: the main pass (k_trace), then
// the second pass (k_trace_deep) if the main pass hands the ray over.
template <bool W>
bool trace_ray(const DevScene& s, const float* rays, int64_t rs, int64_t i, uint32_t flags, const DevHits& h) {
    static thread_local int32_t stk_ref[64][TRACE_BLOCK];
    static thread_local float stk_near[64][TRACE_BLOCK];
    // the ray's own deep-node pool (the kernels share one per call)
    static thread_local std::vector<uint32_t> pool_e(3 * 4096);
    uint32_t pool_n = 0;
    const DeepPool pool{pool_e.data(), &pool_n, 4096u};
    Lane L;
    lane_init<W>(s, rays, rs, (uint32_t)i, flags, L);
    while (!lane_step<W>(s, L, stk_ref, stk_near, 0, pool)) {
    }
    if (!lane_finish(s, L, h)) return false;
    // the second pass resumes a ray from its listed cache nodes, or
    // retraces it; a resumed ray that runs out of budget is retraced whole
    // (the third pass)
    if ((L.state & (LANE_RETRACE | LANE_DEEP_MANY)) == 0 &&
        !resume_list<W>(s, rays, rs, i, L.deep_head, pool_e.data(), h, flags, DEEP_BUDGET))
        return true;
    (void)trace_one<W, 64, D_SERIAL>(s, rays, rs, i, h, flags, stk_ref, stk_near, 0, INT32_MAX);
    return true;
}

template <class F>
void parallel_for(int64_t n, int threads, F f) {
    if (threads <= 1 || n < 256) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t)
        th.emplace_back([&]() {
            for (;;) {
                const int64_t b = next.fetch_add(1024);
                if (b >= n) break;
                const int64_t e = std::min(n, b + 1024);
                for (int64_t i = b; i < e; ++i) f(i);
            }
        });
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

// psy_scene_create's host preparation alone: its status and the instance
// BVH's traversal stack need.
int hb_scene_stack_need(const psy_scene_desc* d, int dice_depth, int* need) {
    HostScene H;
    const int rc = prepare_scene(d, dice_depth, H);
    *need = rc ? 0 : H.stack_need;
    return rc;
}

// psy_trace_world / psy_trace_rays on the CPU.  Returns a psy_status;
// *retraced = rays that took the second pass.
int hb_trace(const psy_scene_desc* d, int dice_depth, int world, const float* rays, int64_t n, uint32_t flags,
             float* t, uint32_t* inst, float* u, float* v, float* aux, int64_t* retraced, int threads) {
    HostBuilt B;
    if (int rc = build(d, dice_depth, B)) return rc;
    if (B.H.stack_need > 64) return PSY_ERR_UNSUPPORTED;
    const DevHits h{t, inst, u, v, aux};
    std::atomic<int64_t> rt{0};
    parallel_for(n, threads, [&](int64_t i) {
        const bool r = world ? trace_ray<true>(B.s, rays, n, i, flags, h) : trace_ray<false>(B.s, rays, n, i, flags, h);
        if (r) rt.fetch_add(1);
    });
    if (retraced) *retraced = rt.load();
    return PSY_OK;
}

}  // extern "C"

namespace {
// The adjoint accumulator of the kernels, with plain adds (the update step
// runs on one thread here).
struct HostAdj {
    static constexpr bool on = true;
    PathBuf Q;
    int32_t n_albedo;
    const int32_t* lparam;
    float* acc;
    PSY_DEV int light_param(int obj) const { return lparam[obj]; }
    PSY_DEV void add(int p, float v) const { acc[p] += v; }
};
}  // namespace

extern "C" {

// psy_render (d_film == NULL) or psy_adjoint on the CPU: one chunk holding
// every path, the wavefront loop with in-order compaction.  albedo /
// escale: psy_scene_set_params values or NULL.  sample_xyz:
// The spectrum grid the host-built integrator upsamples colours with (as
// psy_spectrum_set_grid; cells == NULL clears it).
static std::vector<int32_t> g_spec_cells;
static std::vector<float> g_spec_points;
static DevSpectrum g_spec{};

void hb_set_spectrum_grid(const int32_t* cells, int w, int h, const float* points, int n_points, int n_samples,
                          float smin, float smax, const float* mat) {
    g_spec = DevSpectrum{};
    if (!cells) return;
    g_spec_cells.assign(cells, cells + 8 * (size_t)w * h);
    g_spec_points.assign(points, points + (size_t)(4 + n_samples) * n_points);
    g_spec.cells = g_spec_cells.data();
    g_spec.points = g_spec_points.data();
    g_spec.w = w;
    g_spec.h = h;
    g_spec.n_points = n_points;
    g_spec.n_samples = n_samples;
    g_spec.smin = smin;
    g_spec.smax = smax;
    for (int k = 0; k < 6; ++k) g_spec.m[k] = mat[k];
}

// width*height*spp*3 floats (render); grad: n_albedo + n_lights (adjoint).
int hb_render_params(const psy_scene_desc* d, int dice_depth, const psy_render_params* prm, const float* albedo,
                     int n_albedo, const float* escale, int n_escale, float* sample_xyz, const float* d_film,
                     float* grad, int64_t* rays_traced, int threads) {
    HostBuilt B;
    if (int rc = build(d, dice_depth, B)) return rc;
    if (B.H.chromatic && (d_film || !g_spec.cells)) return PSY_ERR_UNSUPPORTED;
    if (albedo) {
        if (n_albedo != (int)B.H.shaders.size()) return PSY_ERR_INVALID;
        for (int i = 0; i < n_albedo; ++i)
            if (B.H.shaders[i].kind == PSY_SHADER_LAMBERT)
                for (int k = 0; k < 3; ++k) B.H.shaders[i].color[k] = albedo[i];
    }
    if (escale) {
        if (n_escale != (int)B.H.light_objects.size()) return PSY_ERR_INVALID;
        for (int j = 0; j < n_escale; ++j) B.H.escale[B.H.light_objects[j]] = escale[j];
    }
    const int64_t cap = (int64_t)prm->width * prm->height * prm->spp;
    std::vector<float> paths((size_t)(PATH_PLANES * cap));
    std::vector<float> rays[2] = {std::vector<float>((size_t)(20 * cap)), std::vector<float>((size_t)(20 * cap))};
    std::vector<uint32_t> act[2] = {std::vector<uint32_t>((size_t)cap), std::vector<uint32_t>((size_t)cap)};
    std::vector<float> ht((size_t)cap), hu((size_t)cap), hv((size_t)cap), haux((size_t)cap);
    std::vector<uint32_t> hinst((size_t)cap);
    const DevHits h{ht.data(), hinst.data(), hu.data(), hv.data(), haux.data()};
    RenderLaunch R;
    R.s = B.s;
    R.L = B.L;
    R.cam = B.H.cam;
    // Halton tables come from the product's own generator
    std::vector<HaltonDim> dims;
    std::vector<uint16_t> perm;
    make_halton_tables(dims, perm);
    R.dims = dims.data();
    R.perm = perm.data();
    R.sobol_m = nullptr;  // the Halton sampler
    R.sobol_cols = 0;
    R.width = prm->width;
    R.height = prm->height;
    R.spp = prm->spp;
    R.path_length = prm->path_length;
    R.seed = prm->seed;
    for (int k = 0; k < 3; ++k) R.bg[k] = B.H.background[k];
    R.spec = d_film ? DevSpectrum{} : g_spec;
    std::vector<float> qplanes(d_film ? (size_t)(ADJ_STATE_PLANES * cap) : 0);
    const int n_shaders = (int)B.H.shaders.size();
    HostAdj adj{PathBuf{qplanes.data(), cap}, n_shaders, B.H.light_param.data(), grad};
    if (d_film)
        for (size_t p = 0; p < (size_t)n_shaders + B.H.light_objects.size(); ++p) grad[p] = 0.0f;
    const PathBuf P{paths.data(), cap};
    for (int64_t i = 0; i < cap; ++i) render_init_one(R, 0, i, P, rays[0].data(), cap, act[0].data());
    if (d_film)
        for (int64_t i = 0; i < cap; ++i) adjoint_init_one(R, P, adj.Q, i, d_film);
    int64_t n_active = cap, traced = 0;
    int cur = 0;
    std::vector<WRay> nrs((size_t)cap);
    std::vector<char> alive((size_t)cap);
    for (int step = 0; step < 2 * prm->path_length && n_active > 0; ++step) {
        traced += n_active;
        const float* rin = rays[cur].data();
        parallel_for(n_active, threads, [&](int64_t j) { trace_ray<true>(B.s, rin, cap, j, 0u, h); });
        if (d_film) {
            for (int64_t j = 0; j < n_active; ++j)
                alive[j] = render_update_one(R, 0, P, act[cur][j], j, rin, h, cap, nullptr, 0, nrs[j], adj) ? 1 : 0;
        } else {
            parallel_for(n_active, threads, [&](int64_t j) {
                alive[j] = render_update_one(R, 0, P, act[cur][j], j, rin, h, cap, sample_xyz, 0, nrs[j]) ? 1 : 0;
            });
        }
        const int nxt = cur ^ 1;
        int64_t k = 0;
        for (int64_t j = 0; j < n_active; ++j) {
            if (!alive[j]) continue;
            act[nxt][k] = act[cur][j];
            const WRay& r = nrs[j];
            store_wray(rays[nxt].data(), cap, k, r.o, r.d, r.odx, r.ody, r.ddx, r.ddy, r.time, r.type);
            ++k;
        }
        n_active = k;
        cur = nxt;
    }
    if (rays_traced) *rays_traced = traced;
    return PSY_OK;
}

int hb_render(const psy_scene_desc* d, int dice_depth, const psy_render_params* prm, float* sample_xyz,
              int64_t* rays_traced, int threads) {
    return hb_render_params(d, dice_depth, prm, nullptr, 0, nullptr, 0, sample_xyz, nullptr, nullptr, rays_traced,
                            threads);
}

}  // extern "C"
