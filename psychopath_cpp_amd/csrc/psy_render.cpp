// psy_render.cpp -- host side of the path-tracing integrator: per-assembly
// light trees (LightTree::build, accel/light_tree.cpp:99-238) and the
// wavefront loop of PathTraceIntegrator (integrator/path_trace_integrator.cpp:
// 282-384) over the device kernels of psy_render_dev.hpp.
//
// The light trees keep the reference's structure exactly (midpoint splits
// with std::partition, time-sampled node bounds, energies scaled by the
// instance's inverse scale) because the tree decides which light each
// shadow sample picks: the same tree gives the same light per sample.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "psy_internal.hpp"

using namespace psy;

namespace {

// ---- BBox with the reference's operations (basics/bbox.hpp) ----------------
struct LBox {
    V3 lo = {INFINITY, INFINITY, INFINITY};
    V3 hi = {-INFINITY, -INFINITY, -INFINITY};
};

LBox box_union(const LBox& a, const LBox& b) {
    LBox r;
    r.lo = mk(a.lo.x < b.lo.x ? a.lo.x : b.lo.x, a.lo.y < b.lo.y ? a.lo.y : b.lo.y, a.lo.z < b.lo.z ? a.lo.z : b.lo.z);
    r.hi = mk(a.hi.x > b.hi.x ? a.hi.x : b.hi.x, a.hi.y > b.hi.y ? a.hi.y : b.hi.y, a.hi.z > b.hi.z ? a.hi.z : b.hi.z);
    return r;
}

LBox box_lerp_seq(float alpha, const std::vector<LBox>& seq) {
    float na = 0.0f;
    const int i = lerp_seq_index(alpha, (int)seq.size(), &na);
    if (i == -1) return seq[0];
    if (i == -2) return seq.back();
    LBox r;
    r.lo = lerpv(na, seq[i].lo, seq[i + 1].lo);
    r.hi = lerpv(na, seq[i].hi, seq[i + 1].hi);
    return r;
}

// BBox::inverse_transformed: the box corners through Transform::pos_from
LBox inverse_transformed(const LBox& b, const M44& xf) {
    const M44 inv = m_inverse(xf);
    LBox r;
    for (int c = 0; c < 8; ++c) {
        const V3 v = m_pos(inv, mk((c & 4) ? b.hi.x : b.lo.x, (c & 2) ? b.hi.y : b.lo.y, (c & 1) ? b.hi.z : b.lo.z));
        r.lo = mk(v.x < r.lo.x ? v.x : r.lo.x, v.y < r.lo.y ? v.y : r.lo.y, v.z < r.lo.z ? v.z : r.lo.z);
        r.hi = mk(v.x > r.hi.x ? v.x : r.hi.x, v.y > r.hi.y ? v.y : r.hi.y, v.z > r.hi.z ? v.z : r.hi.z);
    }
    return r;
}

// transform_from(bbs, xforms) (basics/bbox.hpp:644-668)
std::vector<LBox> transform_from(const std::vector<LBox>& in, const M44* xs, int tc) {
    std::vector<LBox> bbs = in;
    const int n = (int)bbs.size();
    if (tc == 0) return bbs;
    if (n == tc) {
        for (int i = 0; i < n; ++i) bbs[i] = inverse_transformed(bbs[i], xs[i]);
    } else if (n > tc) {
        const float s = (float)(n - 1);
        for (int i = 0; i < n; ++i) bbs[i] = inverse_transformed(bbs[i], m_lerp_seq((float)i / s, xs, tc));
    } else {
        const float s = (float)(tc - 1);
        std::vector<LBox> t;
        for (int i = 0; i < tc; ++i) t.push_back(inverse_transformed(box_lerp_seq((float)i / s, bbs), xs[i]));
        bbs = t;
    }
    return bbs;
}

// merge(vector<BBox>, vector<BBox>) (basics/bbox.hpp:620-641)
std::vector<LBox> merge(const std::vector<LBox>& a, const std::vector<LBox>& b) {
    std::vector<LBox> c;
    if (a.empty()) c = b;
    else if (b.empty()) c = a;
    else if (a.size() == b.size()) {
        for (size_t i = 0; i < a.size(); ++i) c.push_back(box_union(a[i], b[i]));
    } else if (a.size() > b.size()) {
        const float s = (float)(a.size() - 1);
        for (size_t i = 0; i < a.size(); ++i) c.push_back(box_union(a[i], box_lerp_seq((float)i / s, b)));
    } else {
        const float s = (float)(b.size() - 1);
        for (size_t i = 0; i < b.size(); ++i) c.push_back(box_union(b[i], box_lerp_seq((float)i / s, a)));
    }
    return c;
}

V3 box_center(const LBox& b) { return scale(add(b.lo, b.hi), 0.5f); }

// Color::energy of Color(e, e, e) / of an RGB colour (color/color.hpp:369-375)
float energy3(float r, float g, float b) { return ((r + g) + b) / 3.0f; }

bool grey(const float* c) { return c[0] == c[1] && c[1] == c[2]; }

// Light objects: bounds per time sample, emitted colour at sample 0.
std::vector<LBox> light_object_bounds(const psy_object_desc& od) {
    std::vector<LBox> bs;
    const int nf = floats_per_kind(od.kind);
    for (int t = 0; t < od.time_samples; ++t) {
        const float* p = od.data + (size_t)t * nf;
        LBox b;
        if (od.kind == PSY_OBJ_SPHERE_LIGHT) {
            b.lo = sub(mk(p[0], p[1], p[2]), splat(p[3]));
            b.hi = add(mk(p[0], p[1], p[2]), splat(p[3]));
        } else {
            const V3 h = mk(p[0] * 0.5f, p[1] * 0.5f, 0.0f);
            b.lo = sub(splat(0.0f), h);
            b.hi = add(splat(0.0f), h);
        }
        bs.push_back(b);
    }
    return bs;
}
float light_object_energy(const psy_object_desc& od) {
    if (od.kind == PSY_OBJ_SPHERE_LIGHT) return energy3(od.data[4], od.data[5], od.data[6]);
    if (od.kind == PSY_OBJ_RECTANGLE_LIGHT) return energy3(od.data[2], od.data[3], od.data[4]);
    return 0.0f;
}

struct HNode {
    std::vector<LBox> bounds;
    float energy = 0.0f;
    int index1 = 0, index2 = 0;
    bool leaf = false;
    int instance = 0;
};

struct HTree {
    bool built = false;
    std::vector<HNode> nodes;
    std::vector<LBox> bounds;  // LightAccel::bounds()
    float total_energy = 0.0f;
    int total_lights = 0;
};

struct BuildNode {
    int instance;
    V3 center;
    LBox bbox;
    float energy;
};

struct TreeBuilder {
    const psy_scene_desc* d;
    std::vector<HTree> trees;
    std::string err;

    std::vector<M44> inst_xforms(const psy_assembly_desc& ad, const psy_instance_desc& in) const {
        std::vector<M44> xs;
        for (int k = 0; k < in.transform_count; ++k) {
            M44 m;
            std::memcpy(m.m, ad.transforms + 16 * (in.transform_index + k), sizeof(float) * 16);
            xs.push_back(m);
        }
        return xs;
    }

    // split_lights (accel/light_tree.cpp:159-188)
    static std::vector<BuildNode>::iterator split(std::vector<BuildNode>::iterator start,
                                                  std::vector<BuildNode>::iterator end) {
        V3 mn = start->center, mx = start->center;
        for (auto it = start + 1; it < end; ++it) {
            for (int dd = 0; dd < 3; ++dd) {
                const float c = (&it->center.x)[dd];
                float& lo = (&mn.x)[dd];
                float& hi = (&mx.x)[dd];
                lo = lo < c ? lo : c;
                hi = hi > c ? hi : c;
            }
        }
        int axis = 0;
        if ((mx.y - mn.y) > (mx.x - mn.x)) axis = 1;
        if ((mx.z - mn.z) > (mx.y - mn.y)) axis = 2;
        const float pmid = .5f * ((&mn.x)[axis] + (&mx.x)[axis]);
        auto mid = std::partition(start, end, [axis, pmid](const BuildNode& bn) { return (&bn.center.x)[axis] < pmid; });
        if (mid == start) ++mid;
        return mid;
    }

    int recurse(int a, HTree& T, std::vector<BuildNode>::iterator start, std::vector<BuildNode>::iterator end) {
        const psy_assembly_desc& ad = d->assemblies[a];
        const int me = (int)T.nodes.size();
        T.nodes.emplace_back();
        if (start + 1 == end) {
            const psy_instance_desc& in = ad.instances[start->instance];
            const std::vector<M44> xs = inst_xforms(ad, in);
            std::vector<LBox> bs;
            if (in.type == PSY_INSTANCE_OBJECT) {
                bs = transform_from(light_object_bounds(ad.objects[in.data_index]), xs.data(), (int)xs.size());
            } else {
                const HTree& sub = trees[ad.assemblies[in.data_index]];
                bs = transform_from(sub.bounds, xs.data(), (int)xs.size());
            }
            T.nodes[me].leaf = true;
            T.nodes[me].instance = start->instance;
            T.nodes[me].bounds = bs;
            T.nodes[me].energy = start->energy;
        } else {
            auto mid = split(start, end);
            const int c1 = recurse(a, T, start, mid);
            const int c2 = recurse(a, T, mid, end);
            T.nodes[me].leaf = false;
            T.nodes[me].index1 = c1;
            T.nodes[me].index2 = c2;
            T.nodes[me].bounds = merge(T.nodes[c1].bounds, T.nodes[c2].bounds);
            T.nodes[me].energy = T.nodes[c1].energy + T.nodes[c2].energy;
        }
        return me;
    }

    static float surface_scale(const std::vector<M44>& xs) {
        const M44 x = xs.empty() ? m_identity() : m_lerp_seq(0.5f, xs.data(), (int)xs.size());
        // Transform::get_inv_scale (math/transform.hpp:104-111)
        float s[3];
        for (int r = 0; r < 3; ++r)
            s[r] = std::sqrt((x.m[4 * r + 0] * x.m[4 * r + 0]) + (x.m[4 * r + 1] * x.m[4 * r + 1]) +
                             (x.m[4 * r + 2] * x.m[4 * r + 2]));
        return ((s[0] * s[1]) + (s[0] * s[2]) + (s[1] * s[2])) * 0.33333333f;
    }

    bool build(int a, int depth) {
        if (depth > 64) { err = "assembly nesting deeper than 64 (cycle?)"; return false; }
        if (trees[a].built) return true;
        const psy_assembly_desc& ad = d->assemblies[a];
        for (int i = 0; i < ad.n_instances; ++i) {
            const psy_instance_desc& in = ad.instances[i];
            if (in.type == PSY_INSTANCE_ASSEMBLY && !build(ad.assemblies[in.data_index], depth + 1)) return false;
        }
        HTree& T = trees[a];
        std::vector<BuildNode> bn;
        for (int i = 0; i < ad.n_instances; ++i) {
            const psy_instance_desc& in = ad.instances[i];
            const std::vector<M44> xs = inst_xforms(ad, in);
            if (in.type == PSY_INSTANCE_OBJECT) {
                const psy_object_desc& od = ad.objects[in.data_index];
                const float e = light_object_energy(od);
                if (!(e > 0.0f)) continue;
                // Assembly::instance_bounds_at(0.5, i)
                LBox b = box_lerp_seq(0.5f, light_object_bounds(od));
                if (!xs.empty()) b = inverse_transformed(b, m_lerp_seq(0.5f, xs.data(), (int)xs.size()));
                bn.push_back(BuildNode{i, box_center(b), b, e / surface_scale(xs)});
                T.total_lights += 1;
            } else {
                const HTree& sub = trees[ad.assemblies[in.data_index]];
                const float e = energy3(sub.total_energy, sub.total_energy, sub.total_energy);
                if (!(sub.total_lights > 0 && e > 0.0f)) continue;
                LBox b = box_lerp_seq(0.5f, sub.bounds);
                if (!xs.empty()) b = inverse_transformed(b, m_lerp_seq(0.5f, xs.data(), (int)xs.size()));
                bn.push_back(BuildNode{i, box_center(b), b, e / surface_scale(xs)});
                T.total_lights += sub.total_lights;
            }
        }
        if (!bn.empty()) {
            recurse(a, T, bn.begin(), bn.end());
            T.bounds = T.nodes[0].bounds;
            T.total_energy = T.nodes[0].energy;
        } else {
            T.bounds.assign(1, LBox());
        }
        T.built = true;
        return true;
    }
};

}  // namespace

namespace psy {

int prepare_lights(const psy_scene_desc* d, const std::vector<int32_t>& obj_base, HostScene& H) {
    TreeBuilder tb;
    tb.d = d;
    tb.trees.resize(d->n_assemblies);
    // colours the device path can upsample: grey only
    for (int i = 0; i < d->n_shaders; ++i)
        if (!grey(d->shaders[i].color)) H.chromatic = true;
    if (!grey(d->background)) H.chromatic = true;
    for (int a = 0; a < d->n_assemblies; ++a) {
        const psy_assembly_desc& ad = d->assemblies[a];
        for (int i = 0; i < ad.n_objects; ++i) {
            const psy_object_desc& od = ad.objects[i];
            const int nf = floats_per_kind(od.kind);
            const int c0 = od.kind == PSY_OBJ_SPHERE_LIGHT ? 4 : (od.kind == PSY_OBJ_RECTANGLE_LIGHT ? 2 : -1);
            if (c0 < 0) continue;
            for (int t = 0; t < od.time_samples; ++t)
                if (!grey(od.data + (size_t)t * nf + c0)) H.chromatic = true;
        }
    }
    if (!tb.build(0, 0)) return fail(PSY_ERR_INVALID, "psy_scene_create: " + tb.err);

    std::vector<DevLNode>& nodes = H.l_nodes;
    std::vector<float4>& bounds = H.l_bounds;
    std::vector<DevAsm>& asms = H.l_asms;
    std::vector<DevAsmInst>& ainst = H.l_inst;
    std::vector<M44>& xf = H.l_xf;
    std::vector<int32_t>& subs = H.l_sub;
    asms.assign(d->n_assemblies, DevAsm{});
    for (int a = 0; a < d->n_assemblies; ++a) {
        const psy_assembly_desc& ad = d->assemblies[a];
        const HTree& T = tb.trees[a];
        DevAsm& A = asms[a];
        A = DevAsm{};
        A.node0 = (int32_t)nodes.size();
        A.n_nodes = T.built ? (int32_t)T.nodes.size() : 0;
        A.inst0 = (int32_t)ainst.size();
        A.xf0 = (int32_t)xf.size();
        A.sub0 = (int32_t)subs.size();
        A.obj0 = obj_base[a];
        A.total_lights = T.total_lights;
        if (T.built) {
            for (const HNode& h : T.nodes) {
                DevLNode n{};
                n.leaf = h.leaf ? 1 : 0;
                n.index1 = h.index1;
                n.index2 = h.index2;
                n.instance = h.instance;
                n.energy = h.energy;
                n.bounds = (int32_t)(bounds.size() / 2);
                n.n_bounds = (int32_t)h.bounds.size();
                for (const LBox& b : h.bounds) {
                    bounds.push_back(make_float4(b.lo.x, b.lo.y, b.lo.z, 0.0f));
                    bounds.push_back(make_float4(b.hi.x, b.hi.y, b.hi.z, 0.0f));
                }
                nodes.push_back(n);
            }
        }
        for (int i = 0; i < ad.n_instances; ++i) {
            const psy_instance_desc& in = ad.instances[i];
            ainst.push_back(DevAsmInst{in.type, in.data_index, in.transform_index, in.transform_count});
        }
        for (int k = 0; k < ad.n_transforms; ++k) {
            M44 m;
            std::memcpy(m.m, ad.transforms + 16 * k, sizeof(float) * 16);
            xf.push_back(m);
        }
        for (int k = 0; k < ad.n_assemblies; ++k) subs.push_back(ad.assemblies[k]);
    }
    return PSY_OK;
}

int upload_lights(const HostScene& H, psy_scene* sc) {
    int rc;
    if ((rc = sc->l_nodes.upload(H.l_nodes))) return rc;
    if ((rc = sc->l_bounds.upload(H.l_bounds))) return rc;
    if ((rc = sc->l_asms.upload(H.l_asms))) return rc;
    if ((rc = sc->l_inst.upload(H.l_inst))) return rc;
    if ((rc = sc->l_xf.upload(H.l_xf))) return rc;
    if ((rc = sc->l_sub.upload(H.l_sub))) return rc;
    DevLights& L = sc->lights;
    L.nodes = (const DevLNode*)sc->l_nodes.p;
    L.bounds = (const float4*)sc->l_bounds.p;
    L.asms = (const DevAsm*)sc->l_asms.p;
    L.inst = (const DevAsmInst*)sc->l_inst.p;
    L.xf = (const M44*)sc->l_xf.p;
    L.sub = (const int32_t*)sc->l_sub.p;
    L.n_asms = (int32_t)H.l_asms.size();
    return PSY_OK;
}

}  // namespace psy

// ---------------------------------------------------------------------------
// psy_render: PathTraceIntegrator::integrate over the whole image
// ---------------------------------------------------------------------------
namespace {

// Chunking and scratch layout of one render call: chunks of whole pixels,
// at most ctx->render_chunk paths; `extra` floats after the render scratch.
// With slots > 1 (the wavefront form on several streams) Bx[k - 1] and
// samples_x[k - 1] are chunk k's, in the context's scratch of slot k.
struct RenderSetup {
    int64_t spp, n_pix, total, chunk, cap;
    int slots = 1;
    RenderBuffers B;
    float* chunk_samples;
    RenderBuffers Bx[psy_ctx::MAX_RENDER_SLOTS - 1];
    float* samples_x[psy_ctx::MAX_RENDER_SLOTS - 1] = {};
    float* extra;
};

// Chunks in flight for a render of U's chunking on the context: one chunk
// needs no more, nor does a form other than the wavefront.
int render_slots(const psy_ctx* ctx, const RenderSetup& U) {
    const int64_t chunks = (U.total + U.chunk - 1) / U.chunk;
    return (int)std::min<int64_t>(ctx->render_streams, chunks);
}

int check_render(const psy_ctx* ctx, const psy_scene* s, const psy_render_params* prm, const char* who) {
    if (!ctx || !s || !prm) return fail(PSY_ERR_INVALID, std::string(who) + ": NULL argument");
    if (prm->width <= 0 || prm->height <= 0 || prm->spp <= 0)
        return fail(PSY_ERR_INVALID, std::string(who) + ": width, height and spp must be positive");
    if (prm->width > 32767 || prm->height > 32767)
        return fail(PSY_ERR_INVALID, std::string(who) + ": image sides are limited to 32767 pixels (short pixel coordinates)");
    if (prm->path_length < 1 || prm->path_length > 32)
        return fail(PSY_ERR_INVALID, std::string(who) + ": path_length must be in 1..32");
    if ((prm->flags & ~PSY_RENDER_SOBOL) != 0u)
        return fail(PSY_ERR_INVALID, std::string(who) + ": unknown render flags");
    if ((prm->flags & PSY_RENDER_SOBOL) && !ctx->sobol_m.p)
        return fail(PSY_ERR_INVALID, std::string(who) + ": PSY_RENDER_SOBOL needs the Sobol matrices (psy_sobol_set_matrices)");
    if (s->chromatic && !ctx->spec.cells)
        return fail(PSY_ERR_UNSUPPORTED,
                    std::string(who) + ": the scene has non-grey colours (or surfaces without a shader, which emit "
                    "magenta); set the spectrum grid (psy_spectrum_set_grid) to upsample them");
    return PSY_OK;
}

void plan_render(const psy_ctx* ctx, const psy_render_params* prm, RenderSetup& U) {
    U.spp = prm->spp;
    U.n_pix = (int64_t)prm->width * prm->height;
    U.total = U.n_pix * U.spp;
    int64_t chunk_pix = std::max<int64_t>(1, ctx->render_chunk / U.spp);
    chunk_pix = std::min(chunk_pix, U.n_pix);
    U.chunk = chunk_pix * U.spp;            // paths per chunk (whole pixels)
    U.cap = (U.chunk + 63) / 64 * 64;       // storage stride, keeps every array 256-byte aligned
}

// Floats of one chunk's scratch: paths, 2 ray streams, hits, active lists,
// counters, and the chunk samples when the render keeps its own.
int64_t slot_floats(int64_t cap, bool own_samples) {
    return (int64_t)PATH_PLANES * cap + 2 * 20 * cap + 5 * cap + 2 * cap + 64 + (own_samples ? 3 * cap : 0);
}

// Lays one chunk's scratch out from `base`; returns the float after it.
float* carve_slot(float* base, int64_t cap, bool own_samples, RenderBuffers& B, float** samples) {
    B.cap = cap;
    B.paths = base;
    B.rays[0] = base + (int64_t)PATH_PLANES * cap;
    B.rays[1] = B.rays[0] + 20 * cap;
    float* h = B.rays[1] + 20 * cap;
    B.hits = DevHits{h, (uint32_t*)(h + cap), h + 2 * cap, h + 3 * cap, h + 4 * cap};
    B.act[0] = (uint32_t*)(h + 5 * cap);
    B.act[1] = B.act[0] + cap;
    B.count = B.act[1] + cap;
    *samples = (float*)(B.count + 64);
    return *samples + (own_samples ? 3 * cap : 0);
}

// After plan_render: the scratch, with `extra_floats` after the render's
// own, and that of chunks 1..U.slots-1.
int alloc_render(psy_ctx* ctx, bool own_samples, int64_t extra_floats, RenderSetup& U) {
    const int64_t cap = U.cap;
    if (cap > (int64_t)INT32_MAX) return fail(PSY_ERR_UNSUPPORTED, "psy_render: more than 2^31 paths per chunk");
    const int64_t f_slot = slot_floats(cap, own_samples);
    const int64_t f_extra = (extra_floats + 63) / 64 * 64;
    if (int rc = ctx->render_scratch.reserve((size_t)(f_slot + f_extra) * sizeof(float))) return rc;
    if (int rc = ensure_queue(ctx, cap)) return rc;
    U.extra = carve_slot((float*)ctx->render_scratch.p, cap, own_samples, U.B, &U.chunk_samples);
    if (!ctx->host_counts)
        HIP_OK(hipHostMalloc((void**)&ctx->host_counts, 16 * psy_ctx::MAX_RENDER_SLOTS * sizeof(uint32_t),
                             hipHostMallocDefault));
    if (U.slots > 1) {
        if (!ctx->render_join) HIP_OK(hipEventCreateWithFlags(&ctx->render_join, hipEventDisableTiming));
        for (hipEvent_t& e : ctx->render_done)
            if (!e) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    for (int k = 0; k + 1 < U.slots; ++k) {
        if (int rc = ctx->slot_scratch[k].reserve((size_t)f_slot * sizeof(float))) return rc;
        if (int rc = ensure_trace_scratch(ctx->slot_queue[k], ctx->slot_counters[k], ctx->slot_queue_cap[k], cap))
            return rc;
        if (!ctx->slot_stream[k]) HIP_OK(hipStreamCreateWithFlags(&ctx->slot_stream[k], hipStreamNonBlocking));
        carve_slot((float*)ctx->slot_scratch[k].p, cap, own_samples, U.Bx[k], &U.samples_x[k]);
    }
    return PSY_OK;
}

RenderLaunch make_launch(const psy_ctx* ctx, const psy_scene* s, const psy_render_params* prm) {
    RenderLaunch R;
    R.s = s->dev;
    R.L = s->lights;
    R.cam = s->cam;
    R.dims = (const HaltonDim*)ctx->halton_dims.p;
    R.perm = (const uint16_t*)ctx->halton_perm.p;
    const bool sobol = (prm->flags & PSY_RENDER_SOBOL) != 0u;
    R.sobol_m = sobol ? (const uint64_t*)ctx->sobol_m.p : nullptr;
    R.sobol_cols = sobol ? ctx->sobol_dims * 52u : 0u;
    R.width = prm->width;
    R.height = prm->height;
    R.spp = prm->spp;
    R.path_length = prm->path_length;
    R.seed = prm->seed;
    for (int k = 0; k < 3; ++k) R.bg[k] = s->background[k];
    R.spec = ctx->spec;
    return R;
}

// One chunk in flight of the wavefront form: its scratch, trace scratch,
// stream and pinned count word, and where its paths are.
struct ChunkSlot {
    RenderBuffers B;
    float* samples;  // own chunk samples (unused when the caller keeps every sample)
    uint32_t* queue;
    uint32_t* counters;
    int64_t queue_cap;
    hipStream_t st;
    hipEvent_t done;  // recorded after each step's count read back (several slots)
    uint32_t* count_host;
    int64_t first = 0, n = 0, n_active = 0;  // n == 0: idle
    int cur = 0, step = 0;
};

// The wavefront loop (integrator/path_trace_integrator.cpp:321-349) over
// every chunk: init, then per step trace the active stream, shade it and
// compact the paths that go on.  With `adj` (one entry per slot), the
// shading step is the adjoint pass's (same paths, plus gradients) and no
// samples are written.
// With U.slots > 1 that many chunks are in flight, chunk k >= 1 on the
// context's stream of slot k: the host polls the slots' step events and,
// for each one that has ended, reads its next active count and queues its
// next step (or its resolve and the next chunk), so each chunk's short late
// steps run beside the other chunks' work.  Every path is computed as with
// one slot, so the samples are the same.
int run_wavefront(psy_ctx* ctx, const psy_scene* s, const psy_render_params* prm, const RenderLaunch& R,
                  const RenderSetup& U, float* film_xyz, float* sample_xyz, const AdjointBuffers* adj,
                  int64_t* traced_out, hipStream_t st) {
    const int n_slots = U.slots;
    ChunkSlot S[psy_ctx::MAX_RENDER_SLOTS];
    S[0].B = U.B;
    S[0].samples = U.chunk_samples;
    S[0].queue = (uint32_t*)ctx->queue.p;
    S[0].counters = (uint32_t*)ctx->counters.p;
    S[0].queue_cap = ctx->queue_cap;
    S[0].st = st;
    S[0].count_host = ctx->host_counts;
    S[0].done = ctx->render_done[0];
    if (n_slots > 1) HIP_OK(hipEventRecord(ctx->render_join, st));
    for (int k = 1; k < n_slots; ++k) {
        S[k].B = U.Bx[k - 1];
        S[k].samples = U.samples_x[k - 1];
        S[k].queue = (uint32_t*)ctx->slot_queue[k - 1].p;
        S[k].counters = (uint32_t*)ctx->slot_counters[k - 1].p;
        S[k].queue_cap = ctx->slot_queue_cap[k - 1];
        S[k].st = ctx->slot_stream[k - 1];
        S[k].count_host = ctx->host_counts + 16 * k;
        S[k].done = ctx->render_done[k];
        // the slot's stream starts after the caller's queued work
        HIP_OK(hipStreamWaitEvent(S[k].st, ctx->render_join, 0));
    }
    int64_t traced = 0, next_first = 0;
    auto step = [&](ChunkSlot& c) -> int {
        traced += c.n_active;
        TraceTuning t = ctx->tuning(true);
        t.cont = c.queue + 6 * c.queue_cap;
        t.cont_words = cont_area_words(c.queue_cap);
        int rc = launch_trace(s->dev, true, false, s->stack_need, c.B.rays[c.cur], c.n_active, c.B.cap, c.B.hits,
                              c.queue, c.counters, t, c.st);
        if (rc) return fail(rc, "psy_render: trace launch failed");
        if ((rc = launch_zero_u32(c.B.count + (c.cur ^ 1), 1, c.st))) return fail(rc, "psy_render: launch failed");
        float* samp = sample_xyz ? sample_xyz : c.samples;
        const int64_t sbase = sample_xyz ? 0 : c.first;
        rc = adj ? launch_render_adjoint(R, c.first, c.B, adj[&c - S], c.cur, c.n_active, c.step, c.st)
                 : launch_render_update(R, c.first, c.B, c.cur, c.n_active, c.step, ctx->render_shade_waves, samp,
                                        sbase, c.st);
        if (rc) return fail(rc, "psy_render: update launch failed");
        HIP_OK(hipMemcpyAsync(c.count_host, c.B.count + (c.cur ^ 1), sizeof(uint32_t), hipMemcpyDeviceToHost, c.st));
        if (n_slots > 1) HIP_OK(hipEventRecord(c.done, c.st));
        return PSY_OK;
    };
    // the slot's next chunk and its first step; no chunks left: idle
    auto begin = [&](ChunkSlot& c) -> int {
        c.n = std::min(U.chunk, U.total - next_first);
        if (c.n <= 0) { c.n = 0; return PSY_OK; }
        c.first = next_first;
        next_first += c.n;
        int rc = launch_render_init(R, c.first, c.n, c.B, c.st);
        if (!rc && adj) rc = launch_adjoint_init(R, c.B, adj[&c - S], c.n, c.st);
        if (rc) return fail(rc, "psy_render: init launch failed");
        c.n_active = c.n;
        c.cur = 0;
        c.step = 0;
        return step(c);
    };
    auto run = [&]() -> int {
        for (int k = 0; k < n_slots; ++k)
            if (int rc = begin(S[k])) return rc;
        auto busy = [&]() {
            for (int k = 0; k < n_slots; ++k)
                if (S[k].n > 0) return true;
            return false;
        };
        for (int k = 0; busy(); k = (k + 1) % n_slots) {
            ChunkSlot& c = S[k];
            if (c.n == 0) continue;
            if (n_slots > 1) {
                // whichever chunk's step ends first goes on first
                const hipError_t e = hipEventQuery(c.done);
                if (e == hipErrorNotReady) {
                    (void)hipGetLastError();  // not an error: the step is still running
                    continue;
                }
                HIP_OK(e);
            } else {
                HIP_OK(hipStreamSynchronize(c.st));
            }
            c.n_active = *(volatile const uint32_t*)c.count_host;  // written by the copy engine
            c.cur ^= 1;
            if (++c.step < 2 * prm->path_length && c.n_active > 0) {
                if (int rc = step(c)) return rc;
                continue;
            }
            if (film_xyz && !adj) {
                const float* samp = sample_xyz ? sample_xyz : c.samples;
                const int64_t sbase = sample_xyz ? 0 : c.first;
                if (int rc = launch_render_resolve(samp, sbase, c.first / U.spp, c.n / U.spp, (int)U.spp, film_xyz, c.st))
                    return fail(rc, "psy_render: resolve launch failed");
            }
            if (int rc = begin(c)) return rc;
        }
        return PSY_OK;
    };
    const int rc = run();
    // the caller's stream goes on after the other slots' work, also after a
    // failed launch (each wait takes the event as last recorded)
    for (int k = 1; k < n_slots; ++k) {
        HIP_OK(hipEventRecord(ctx->render_join, S[k].st));
        HIP_OK(hipStreamWaitEvent(st, ctx->render_join, 0));
    }
    if (rc) return rc;
    if (traced_out) *traced_out = traced;
    return PSY_OK;
}

// The fused persistent form (PSY_RENDER_FUSED).
int run_fused(psy_ctx* ctx, const psy_scene* s, const psy_render_params* prm, const RenderLaunch& R,
              const RenderSetup& U, float* film_xyz, float* sample_xyz, int64_t* rays_traced, hipStream_t st) {
    const RenderBuffers& B = U.B;
    const int64_t cap = U.cap, spp = U.spp;
    const int bpc = ctx->render_blocks_per_cu > 0 ? ctx->render_blocks_per_cu : std::max(1, ctx->render_occupancy);
    const int blocks = std::max(1, ctx->n_cu * bpc);
    uint32_t* counters = B.count + 8;  // u32 x 8 (8-byte aligned)
    unsigned long long* steps = (unsigned long long*)(B.count + 16);
    HIP_OK(hipMemsetAsync(steps, 0, sizeof(unsigned long long), st));
    for (int64_t first = 0; first < U.total; first += U.chunk) {
        const int64_t n = std::min(U.chunk, U.total - first);
        float* samp = sample_xyz ? sample_xyz : U.chunk_samples;
        const int64_t sbase = sample_xyz ? 0 : first;
        PersistBuffers PB;
        PB.paths = B.paths;
        PB.counters = counters;
        PB.defer_ids = B.act[0];
        PB.defer_rays = B.rays[0];
        PB.resume_ids = B.act[1];
        PB.resume_rays = B.rays[1];
        PB.sample_xyz = samp;
        PB.sbase = sbase;
        PB.cap = cap;
        uint32_t n_resume = 0;
        for (int round = 0;; ++round) {
            if (round > 4 * prm->path_length + 4) return fail(PSY_ERR_HIP, "psy_render: fused rounds did not converge");
            int rc = launch_set_counters(counters, n_resume, st);
            if (!rc) rc = launch_render_persist(R, first, round == 0 ? n : 0, PB, blocks, ctx->render_refill, st);
            if (rc) return fail(rc, "psy_render: fused launch failed");
            uint32_t n_defer = 0;
            HIP_OK(hipMemcpyAsync(&n_defer, counters + 2, sizeof n_defer, hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            if (n_defer == 0) break;
            // the parked rays through the two-pass tracer, then shading
            rc = launch_trace(s->dev, true, false, s->stack_need, PB.defer_rays, n_defer, cap, B.hits,
                              (uint32_t*)ctx->queue.p, (uint32_t*)ctx->counters.p, ctx->tuning(true), st);
            if (!rc) rc = launch_zero_u32(counters + 4, 1, st);
            if (!rc)
                rc = launch_render_update_io(R, first, B.paths, cap, PB.defer_ids, PB.defer_rays, B.hits, n_defer,
                                             PB.resume_ids, PB.resume_rays, counters + 4, samp, sbase, st);
            if (rc) return fail(rc, "psy_render: second-pass launch failed");
            HIP_OK(hipMemcpyAsync(&n_resume, counters + 4, sizeof n_resume, hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            if (n_resume == 0) break;
        }
        int rc = launch_render_count_steps(B.paths, cap, n, steps, st);
        if (!rc && film_xyz) rc = launch_render_resolve(samp, sbase, first / spp, n / spp, (int)spp, film_xyz, st);
        if (rc) return fail(rc, "psy_render: launch failed");
    }
    unsigned long long total_steps = 0;
    HIP_OK(hipMemcpyAsync(&total_steps, steps, sizeof total_steps, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (rays_traced) *rays_traced = (int64_t)total_steps;
    return PSY_OK;
}

// The stream form (PSY_RENDER_STREAM): a pool of U.cap path slots over
// windows of the image's samples, K steps per captured graph.
int capture_steps(psy_ctx* ctx, const psy_scene* s, const RenderLaunch& R, const RenderSetup& U, StreamState* S,
                  uint32_t* wait, float* samp, int64_t sbase, int iters, hipGraphExec_t* exec) {
    if (!ctx->capture_stream) HIP_OK(hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
    hipStream_t cs = ctx->capture_stream;
    const RenderBuffers& B = U.B;
    TraceTuning tune = ctx->tuning(true, false);
    const int blocks = std::max(1, 2 * ctx->n_cu);
    HIP_OK(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
    int rc = 0;
    for (int it = 0; it < iters && !rc; ++it) {
        const int cur = it & 1;
        tune.n_dev = &S->count[cur];
        rc = launch_trace(s->dev, true, false, s->stack_need, B.rays[cur], U.cap, U.cap, B.hits,
                          (uint32_t*)ctx->queue.p, (uint32_t*)ctx->counters.p, tune, cs);
        if (!rc) rc = launch_zero_u32(&S->count[cur ^ 1], 1, cs);
        if (!rc) rc = launch_stream_shade(R, B, cur, S, wait, blocks, samp, sbase, cs);
        if (!rc && cur) rc = launch_zero_u32(&S->n_wait, 1, cs);  // the odd step refilled every parked slot
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(rc ? rc : PSY_ERR_HIP, "psy_render: capturing the stream steps failed");
    }
    const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return fail(PSY_ERR_HIP, "psy_render: graph instantiation failed");
    return PSY_OK;
}

int run_stream(psy_ctx* ctx, const psy_scene* s, const psy_render_params* prm, const RenderLaunch& R,
               const RenderSetup& U, int64_t window, float* film_xyz, float* sample_xyz, int64_t* rays_traced,
               hipStream_t st) {
    StreamState* S = (StreamState*)U.extra;
    uint32_t* wait = (uint32_t*)(U.extra + 64);
    float* wsamp = sample_xyz ? sample_xyz : U.extra + 64 + U.cap;
    const int iters = ctx->render_graph_iters;
    for (int64_t w0 = 0; w0 < U.total; w0 += window) {
        const int64_t w1 = std::min(U.total, w0 + window);
        const int64_t sbase = sample_xyz ? 0 : w0;
        const int64_t n0 = std::min(U.cap, w1 - w0);
        if (int rc = launch_stream_init(R, U.B, w0, n0, w1, S, w0 == 0, st))
            return fail(rc, "psy_render: stream init launch failed");
        hipGraphExec_t exec = nullptr;
        if (int rc = capture_steps(ctx, s, R, U, S, wait, wsamp, sbase, iters, &exec)) return rc;
        // every iteration with a live stream traces at least one ray and a
        // path takes at most 2 * path_length of them
        const int64_t max_iters = 2 * (int64_t)prm->path_length * ((w1 - w0 + U.cap - 1) / U.cap + 1) * 4 + 64;
        int rc = PSY_OK;
        for (int64_t done = 0;; done += iters) {
            if (done > max_iters) {
                rc = fail(PSY_ERR_HIP, "psy_render: stream steps did not converge");
                break;
            }
            if (hipGraphLaunch(exec, st) != hipSuccess) {
                rc = fail(PSY_ERR_HIP, "psy_render: graph launch failed");
                break;
            }
            uint32_t n = 0;
            if (hipMemcpyAsync(&n, &S->count[0], sizeof n, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess) {
                rc = fail(PSY_ERR_HIP, "psy_render: stream synchronization failed");
                break;
            }
            if (n == 0) break;
        }
        (void)hipGraphExecDestroy(exec);
        if (rc) return rc;
        if (film_xyz) {
            rc = launch_render_resolve(wsamp, sbase, w0 / U.spp, (w1 - w0) / U.spp, (int)U.spp, film_xyz, st);
            if (rc) return fail(rc, "psy_render: resolve launch failed");
        }
    }
    unsigned long long total_rays = 0;
    HIP_OK(hipMemcpyAsync(&total_rays, &S->rays, sizeof total_rays, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (rays_traced) *rays_traced = (int64_t)total_rays;
    return PSY_OK;
}

int render_forms(psy_ctx* ctx, const psy_scene* s, const psy_render_params* prm, float* film_xyz, float* sample_xyz,
                 int64_t* rays_traced, hipStream_t st) {
    RenderSetup U;
    plan_render(ctx, prm, U);
    const RenderLaunch R = make_launch(ctx, s, prm);
    if (ctx->render_mode == PSY_RENDER_STREAM) {
        // sample windows of whole pixels: the whole image when the caller
        // keeps every sample, else up to max(4 pools, 2^26) samples at a time
        int64_t window = U.total;
        if (!sample_xyz) {
            window = std::min(U.total, std::max<int64_t>(4 * U.cap, (int64_t)1 << 26));
            window = std::max<int64_t>(U.spp, window / U.spp * U.spp);
        }
        if (int rc = alloc_render(ctx, false, 64 + U.cap + (sample_xyz ? 0 : 3 * window), U)) return rc;
        return run_stream(ctx, s, prm, R, U, window, film_xyz, sample_xyz, rays_traced, st);
    }
    if (ctx->render_mode == PSY_RENDER_WAVEFRONT) U.slots = render_slots(ctx, U);
    if (int rc = alloc_render(ctx, sample_xyz == nullptr, 0, U)) return rc;
    if (ctx->render_mode == PSY_RENDER_FUSED) return run_fused(ctx, s, prm, R, U, film_xyz, sample_xyz, rays_traced, st);
    return run_wavefront(ctx, s, prm, R, U, film_xyz, sample_xyz, nullptr, rays_traced, st);
}

}  // namespace

extern "C" {

int psy_render(psy_ctx* ctx, const psy_scene* s, const psy_render_params* prm, float* film_xyz,
               float* sample_xyz, int64_t* rays_traced, void* stream) {
    if (int rc = check_render(ctx, s, prm, "psy_render")) return rc;
    if (rays_traced) *rays_traced = 0;
    HIP_OK(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    if (int rc = scratch_order(ctx, st)) return rc;
    const int rc = render_forms(ctx, s, prm, film_xyz, sample_xyz, rays_traced, st);
    const int rr = scratch_release(ctx, st);
    return rc ? rc : rr;
}

// PathTraceIntegrator::integrate's bucket size
// (integrator/path_trace_integrator.cpp:28-34)
int32_t psy_bucket_size(int32_t width, int32_t height, int32_t spp, int32_t thread_count) {
    if (width <= 0 || height <= 0 || spp <= 0 || thread_count <= 0) return -1;
    const int max_bucket_size = (int)std::sqrt((width * height) / (thread_count * 4.0f));
    int bucket_size = (int)std::sqrt(static_cast<float>(65536) / spp);  // Config::samples_per_bucket
    bucket_size = std::min(max_bucket_size, bucket_size);
    bucket_size = std::max(1, bucket_size);
    return bucket_size;
}

int psy_film_adaptive(psy_ctx* ctx, const float* sample_xyz, int32_t width, int32_t height, int32_t spp,
                      int32_t spp_max, float variance_max, int32_t bucket_size, float* film_xyz, float* variance_xyz,
                      int32_t* passes, void* stream) {
    if (!ctx || !sample_xyz) return fail(PSY_ERR_INVALID, "psy_film_adaptive: NULL argument");
    if (width <= 0 || height <= 0 || spp <= 0 || bucket_size <= 0)
        return fail(PSY_ERR_INVALID, "psy_film_adaptive: width, height, spp and bucket_size must be positive");
    if ((int64_t)width * height > ((int64_t)1 << 31) / 10)
        return fail(PSY_ERR_INVALID, "psy_film_adaptive: image too large");
    HIP_OK(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    if (int rc = scratch_order(ctx, st)) return rc;  // film_state is the context's, like the trace scratch
    if (int rc = ctx->film_state.reserve(sizeof(float) * 10 * (size_t)width * height)) return rc;
    const int rc = launch_film_adaptive(sample_xyz, width, height, spp, bucket_size, spp_max, variance_max,
                                        (float*)ctx->film_state.p, film_xyz, variance_xyz, passes, st);
    if (rc) return fail(rc, "psy_film_adaptive: launch failed");
    return scratch_release(ctx, st);
}

int psy_render_adaptive(psy_ctx* ctx, const psy_scene* s, const psy_render_params* prm, int32_t spp_max,
                        float variance_max, int32_t thread_count, float* film_xyz, float* variance_xyz,
                        int32_t* passes, int64_t* rays_traced, void* stream) {
    if (int rc = check_render(ctx, s, prm, "psy_render_adaptive")) return rc;
    if (thread_count <= 0) return fail(PSY_ERR_INVALID, "psy_render_adaptive: thread_count must be positive");
    HIP_OK(hipSetDevice(ctx->device));
    const size_t bytes = sizeof(float) * 3 * (size_t)prm->width * prm->height * prm->spp;
    if (int rc = ctx->film_samples.reserve(bytes)) return rc;
    int rc = psy_render(ctx, s, prm, nullptr, (float*)ctx->film_samples.p, rays_traced, stream);
    if (rc) return rc;
    const int32_t bs = psy_bucket_size(prm->width, prm->height, prm->spp, thread_count);
    return psy_film_adaptive(ctx, (const float*)ctx->film_samples.p, prm->width, prm->height, prm->spp,
                             std::max(spp_max, prm->spp), variance_max, bs, film_xyz, variance_xyz, passes, stream);
}

// The render autotuner: coordinate descent over the render form and its
// paths in flight (wavefront chunk or stream pool), the main pass's launch
// form and LDS stack depth, and the stream form's steps per captured graph,
// timing psy_render on a sample budget of the image (every pixel at reduced
// spp).  Leaves the best setting on the context.
int psy_render_autotune(psy_ctx* ctx, psy_scene* s, const psy_render_params* prm, int64_t max_samples,
                        psy_render_tuning* out) {
    if (int rc = check_render(ctx, s, prm, "psy_render_autotune")) return rc;
    HIP_OK(hipSetDevice(ctx->device));
    psy_render_params p = *prm;
    const int64_t n_pix = (int64_t)p.width * p.height;
    if (max_samples <= 0) max_samples = (int64_t)1 << 23;
    p.spp = (int32_t)std::max<int64_t>(1, std::min<int64_t>(p.spp, max_samples / std::max<int64_t>(1, n_pix)));
    const int64_t total = n_pix * p.spp;
    // saved so that a failed trial leaves the context as it was
    const int s_mode = ctx->render_mode, s_tmode = ctx->mode, s_refill = ctx->refill_min, s_stack = ctx->trace_stack,
              s_iters = ctx->render_graph_iters, s_streams = ctx->render_streams, s_shade = ctx->render_shade_waves;
    const int64_t s_chunk = ctx->render_chunk;
    struct Cfg { int mode; int64_t pool; int tmode, refill, stack, iters, streams, shade; };
    auto apply = [&](const Cfg& c) {
        ctx->render_mode = c.mode;
        ctx->render_chunk = c.pool;
        ctx->mode = c.tmode;
        ctx->refill_min = c.refill;
        ctx->trace_stack = c.stack;
        ctx->render_graph_iters = c.iters;
        ctx->render_streams = c.streams;
        ctx->render_shade_waves = c.shade;
    };
    int trials = 0;
    const int64_t full = n_pix * prm->spp;
    auto timed = [&](const Cfg& c, double* mrays) -> int {
        apply(c);
        // k wavefront chunks in flight are timed on at least k chunks
        // (whole-image spp, within the full render's)
        psy_render_params q = p;
        if (c.mode == PSY_RENDER_WAVEFRONT && c.streams > 1)
            q.spp = (int32_t)std::min<int64_t>(
                prm->spp, std::max<int64_t>(p.spp, (c.streams * c.pool + n_pix - 1) / n_pix));
        int64_t rays = 0;
        int rc = psy_render(ctx, s, &q, nullptr, nullptr, &rays, nullptr);  // warm-up: buffers, code
        if (rc) return rc;
        HIP_OK(hipDeviceSynchronize());
        // the fastest of up to 3 timed renders, stopping once 30 ms are spent
        double fastest = 1e30, spent = 0.0;
        for (int rep = 0; rep < 3 && (rep == 0 || spent < 0.03); ++rep) {
            const auto t0 = std::chrono::steady_clock::now();
            rc = psy_render(ctx, s, &q, nullptr, nullptr, &rays, nullptr);
            if (rc) return rc;
            HIP_OK(hipDeviceSynchronize());
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            fastest = std::min(fastest, dt);
            spent += dt;
        }
        *mrays = (double)rays / std::max(fastest, 1e-9) / 1e6;
        ++trials;
        return PSY_OK;
    };
    Cfg best{PSY_RENDER_STREAM, std::min<int64_t>((int64_t)1 << 21, total), PSY_TRACE_MODE_PLAIN, 16, 16, 4, 2, 0};
    double best_m = 0.0;
    int rc = timed(best, &best_m);
    auto consider = [&](Cfg c) -> int {
        double m = 0.0;
        if (int r = timed(c, &m)) return r;
        if (m > best_m) { best_m = m; best = c; }
        return PSY_OK;
    };
    // form and paths in flight: powers of two from 2^19, up to the workload
    // (the sample for the stream form, the full render for the wavefront
    // form's two chunks in flight)
    for (int mode : {PSY_RENDER_STREAM, PSY_RENDER_WAVEFRONT}) {
        const int64_t upto = mode == PSY_RENDER_WAVEFRONT ? full : total;
        for (int64_t pool = (int64_t)1 << 19; !rc && pool <= ((int64_t)1 << 23); pool <<= 1) {
            const int64_t c = std::min(pool, upto);
            if (mode != best.mode || c != best.pool) { Cfg t = best; t.mode = mode; t.pool = c; rc = consider(t); }
            if (c == upto) break;
        }
    }
    if (!rc) { Cfg t = best; t.tmode = PSY_TRACE_MODE_PERSISTENT; t.refill = 32; rc = consider(t); }
    if (!rc) { Cfg t = best; t.stack = 32; rc = consider(t); }
    if (!rc && best.tmode == PSY_TRACE_MODE_PLAIN) { Cfg t = best; t.tmode = PSY_TRACE_MODE_PLAIN_COHERENT; rc = consider(t); }
    if (best.mode == PSY_RENDER_STREAM)
        for (int it : {2, 8, 16}) if (!rc && it != best.iters) { Cfg t = best; t.iters = it; rc = consider(t); }
    // wavefront chunks in flight: 1, 3 and 4 against the search's 2, then
    // the best count at half the chunk, then the whole image as k chunks all
    // in flight (k = 2..4, chunks within the search's 2^23 paths; such a
    // trial renders the whole image)
    if (best.mode == PSY_RENDER_WAVEFRONT) {
        for (int k : {1, 3, 4})
            if (!rc && k != best.streams) { Cfg t = best; t.streams = k; rc = consider(t); }
        if (!rc && best.streams > 1 && best.pool / 2 >= ((int64_t)1 << 19)) {
            Cfg t = best; t.pool = best.pool / 2; rc = consider(t);
        }
        for (int k = 2; k <= psy_ctx::MAX_RENDER_SLOTS && !rc; ++k) {
            const int64_t pool = (n_pix + k - 1) / k * prm->spp;
            if (pool > ((int64_t)1 << 23) || pool >= full) continue;
            if (pool == best.pool && k == best.streams) continue;
            Cfg t = best; t.pool = pool; t.streams = k; rc = consider(t);
        }
        // the hit step's shading held to 4 waves per SIMD
        if (!rc) { Cfg t = best; t.shade = 4; rc = consider(t); }
        // the other stack depth again, with the chunks now in flight (more
        // chunks in flight favour the 16-entry stack's higher occupancy)
        if (!rc) { Cfg t = best; t.stack = best.stack == 16 ? 32 : 16; rc = consider(t); }
        // and the other one-thread-per-ray launch form
        if (!rc && (best.tmode == PSY_TRACE_MODE_PLAIN || best.tmode == PSY_TRACE_MODE_PLAIN_COHERENT)) {
            Cfg t = best;
            t.tmode = best.tmode == PSY_TRACE_MODE_PLAIN ? PSY_TRACE_MODE_PLAIN_COHERENT : PSY_TRACE_MODE_PLAIN;
            rc = consider(t);
        }
    }
    // dicing depth: the scene's dice cache re-built 4 and 2 levels shallower
    // and 2 deeper (within 16 GiB and the 2^32-node cap); rays that need
    // deeper splits than the cache holds go to the second pass
    const int s_levels = s->info.dice_depth;
    int best_levels = s_levels;
    if (!rc && !s->patch_ids.empty()) {
        const double cap = 16.0 * 1024.0 * 1024.0 * 1024.0;
        apply(best);
        for (int d : {s_levels - 4, s_levels - 2, s_levels + 2}) {
            if (d < 0 || d > 24) continue;
            if ((double)((int64_t)1 << (d + 1)) * 32.0 * (double)s->patch_ids.size() > cap) continue;
            if ((rc = rebuild_dice(s, d))) break;
            double m = 0.0;
            if ((rc = timed(best, &m))) break;
            if (m > best_m) { best_m = m; best_levels = d; }
        }
        if (rc || s->info.dice_depth != best_levels) {
            const int rrc = rebuild_dice(s, rc ? s_levels : best_levels);
            if (!rc) rc = rrc;
        }
    }
    if (rc) {
        ctx->render_mode = s_mode; ctx->mode = s_tmode; ctx->refill_min = s_refill; ctx->trace_stack = s_stack;
        ctx->render_graph_iters = s_iters; ctx->render_chunk = s_chunk; ctx->render_streams = s_streams;
        ctx->render_shade_waves = s_shade;
        return rc;
    }
    apply(best);
    if (out) {
        out->mode = best.mode;
        out->trace_mode = best.tmode;
        out->refill_min = best.refill;
        out->stack = best.stack;
        out->steps_per_graph = best.iters;
        out->trials = trials;
        out->pool = best.pool;
        out->dice_depth = best_levels;
        out->streams = best.streams;
        out->shade_waves = best.shade;
        out->mrays_per_s = best_m;
    }
    return PSY_OK;
}

int psy_adjoint(psy_ctx* ctx, const psy_scene* s, const psy_render_params* prm, const float* d_film_xyz,
                float* d_albedo, float* d_emission, int64_t* rays_traced, void* stream) {
    if (int rc = check_render(ctx, s, prm, "psy_adjoint")) return rc;
    if (s->chromatic)
        return fail(PSY_ERR_UNSUPPORTED, "psy_adjoint: the scene has non-grey colours; the parameters are grey "
                    "albedos and emission scales");
    if (!d_film_xyz) return fail(PSY_ERR_INVALID, "psy_adjoint: NULL film gradient");
    for (const DevShader& sh : s->host_shaders)
        if (sh.kind == PSY_SHADER_GTR)
            return fail(PSY_ERR_UNSUPPORTED, "psy_adjoint: the scene has a GTR shader; the adjoint differentiates "
                        "Lambert albedos and emission scales only");
    if (prm->path_length > ADJ_MAX_VERTS)
        return fail(PSY_ERR_UNSUPPORTED, "psy_adjoint: path_length above 8 (scattering vertices kept per path)");
    const int64_t n_albedo = (int64_t)s->host_shaders.size();
    const int64_t n_light = (int64_t)s->light_objects.size();
    if ((n_albedo > 0 && !d_albedo) || (n_light > 0 && !d_emission))
        return fail(PSY_ERR_INVALID, "psy_adjoint: NULL gradient output");
    const int64_t n_params = n_albedo + n_light;
    if (n_params > (int64_t)INT32_MAX / 2) return fail(PSY_ERR_UNSUPPORTED, "psy_adjoint: too many parameters");
    if (rays_traced) *rays_traced = 0;
    HIP_OK(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    AdjointBuffers J{};
    J.n_albedo = (int32_t)n_albedo;
    J.n_params = (int32_t)n_params;
    J.lds = n_params <= ADJ_LDS_MAX_PARAMS;
    J.rows = std::max(1, 2 * ctx->n_cu);
    J.d_film = d_film_xyz;
    J.lparam = (const int32_t*)s->lparam_dev.p;
    // extra scratch: adjoint path state per chunk in flight, gradient
    // vector, partial rows per chunk in flight (a block of one chunk's
    // launch adds into the row of its index, so the two chunks' launches,
    // which may overlap, keep rows of their own)
    RenderSetup U;
    plan_render(ctx, prm, U);
    U.slots = render_slots(ctx, U);
    const int64_t f_q = (int64_t)ADJ_STATE_PLANES * U.cap;
    const int64_t f_grad = (n_params + 63) / 64 * 64;
    const int64_t f_part = J.lds ? (int64_t)J.rows * n_params : 0;
    if (int rc = alloc_render(ctx, true, U.slots * (f_q + f_part) + f_grad, U)) return rc;
    J.q = U.extra;
    J.grad = J.q + U.slots * f_q;
    J.partials = J.grad + f_grad;
    AdjointBuffers JS[psy_ctx::MAX_RENDER_SLOTS];
    for (int k = 0; k < U.slots; ++k) {
        JS[k] = J;
        JS[k].q = J.q + k * f_q;
        JS[k].partials = J.partials + k * f_part;
    }
    J.rows *= U.slots;  // the reduction sums the rows of every chunk in flight
    if (int rc = scratch_order(ctx, st)) return rc;
    int rc = launch_fill_f32(J.grad, f_grad + U.slots * f_part, 0.0f, st);
    if (rc) return fail(rc, "psy_adjoint: launch failed");
    // the replay takes the flat grey spectrum: paths do not depend on
    // colours, and every contribution is then the same at the 4 wavelengths
    RenderLaunch R = make_launch(ctx, s, prm);
    R.spec = DevSpectrum{};
    if ((rc = run_wavefront(ctx, s, prm, R, U, nullptr, nullptr, JS, rays_traced, st))) return rc;
    if ((rc = launch_adjoint_reduce(J, d_albedo, d_emission, st))) return fail(rc, "psy_adjoint: launch failed");
    return scratch_release(ctx, st);
}

}  // extern "C"
