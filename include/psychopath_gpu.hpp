/*
 * psychopath_gpu.hpp -- the C++ host interface of the MI355X ray-stream
 * library: the reference's scene / tracer / integrator classes, kept by
 * name, argument meaning and ownership, over the C ABI of psy_gpu.h.
 * Header only; link libpsy_gpu.so.  No HIP header is needed: device
 * memory goes through psy_malloc / psy_copy_*.
 *
 *   Assembly            scene/assembly.hpp:74-330
 *   Scene               scene/scene.hpp:33-53
 *   Camera              basics/camera.hpp:14-105
 *   Sphere, Bilinear, Bicubic, SphereLight, RectangleLight
 *                       object/sphere.hpp, object/bilinear.hpp,
 *                       object/bicubic.hpp, lights/sphere_light.hpp,
 *                       lights/rectangle_light.hpp
 *   LambertShader, EmitShader
 *                       shading/surface_shader.hpp:28-53
 *   WorldRay            basics/ray.hpp:162-185
 *   Intersection        basics/intersection.hpp:22-49, with the closures
 *                       LambertClosure / EmitClosure / SurfaceClosureUnion
 *                       (shading/surface_closure.hpp, shading/closure_union.hpp)
 *   Tracer::trace       tracer/tracer.hpp:57-76, tracer/tracer.cpp:38-121
 *   Film                film/film.hpp:78-153 (per-pixel CIE XYZ means)
 *   PathTraceIntegrator integrator/path_trace_integrator.hpp:103-120
 *
 * Differences a caller of the reference should know:
 *   - Scene::finalize() builds the scene on the GPU (flattened instance
 *     BVH, dice cache, light trees) and needs a device; failures throw
 *     std::runtime_error with psy_last_error()'s text (the reference has
 *     no error path);
 *   - Intersection carries the closest hit (t, InstanceID, differential
 *     geometry, offset, space, backfacing, light pdf) and the closure the
 *     shader binds, re-materialised on the host from the device's record;
 *     the closures keep linear RGB colours and give RGB factors (the
 *     reference upsamples them to spectral samples; PathTraceIntegrator
 *     does that on the device);
 *   - colours are grey (r == g == b) for rendering; the tracer takes any.
 */
#ifndef PSYCHOPATH_GPU_HPP
#define PSYCHOPATH_GPU_HPP

#include <array>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <functional>
#include <sstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "psy_gpu.h"

namespace psy {

inline void check(int rc, const char* what) {
    if (rc != PSY_OK) throw std::runtime_error(std::string(what) + ": " + psy_last_error());
}

struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};

struct Color {
    float r = 0.0f, g = 0.0f, b = 0.0f;
    Color() = default;
    explicit Color(float v) : r(v), g(v), b(v) {}
    Color(float r_, float g_, float b_) : r(r_), g(g_), b(b_) {}
};

// Row-major 4x4 in the reference's row-vector convention (p' = p * M),
// the layout of Matrix44::data (math/matrix.hpp:18-22).
struct Transform {
    std::array<float, 16> m{{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}};
    Transform() = default;
    explicit Transform(const float* d) { std::memcpy(m.data(), d, sizeof(float) * 16); }
};

// ---- objects ----------------------------------------------------------------
class Object {
public:
    virtual ~Object() = default;
    virtual int kind() const = 0;
    virtual int time_samples() const = 0;
    virtual std::vector<float> data() const = 0;  // time_samples * floats-per-kind
};

class Sphere : public Object {
public:
    Sphere(Vec3 center, float radius) : c_{center}, r_{radius} {}
    Sphere(std::vector<Vec3> centers, std::vector<float> radii) : c_(std::move(centers)), r_(std::move(radii)) {
        if (c_.empty() || c_.size() != r_.size()) throw std::invalid_argument("Sphere: one radius per center");
    }
    int kind() const override { return PSY_OBJ_SPHERE; }
    int time_samples() const override { return (int)c_.size(); }
    std::vector<float> data() const override {
        std::vector<float> d;
        for (size_t i = 0; i < c_.size(); ++i) d.insert(d.end(), {c_[i].x, c_[i].y, c_[i].z, r_[i]});
        return d;
    }

private:
    std::vector<Vec3> c_;
    std::vector<float> r_;
};

class Bilinear : public Object {
public:
    Bilinear(Vec3 v1, Vec3 v2, Vec3 v3, Vec3 v4) { add_time_sample(v1, v2, v3, v4); }
    void add_time_sample(Vec3 v1, Vec3 v2, Vec3 v3, Vec3 v4) { v_.push_back({v1, v2, v3, v4}); }
    int kind() const override { return PSY_OBJ_BILINEAR; }
    int time_samples() const override { return (int)v_.size(); }
    std::vector<float> data() const override {
        std::vector<float> d;
        for (const auto& s : v_)
            for (const Vec3& p : s) d.insert(d.end(), {p.x, p.y, p.z});
        return d;
    }

private:
    std::vector<std::array<Vec3, 4>> v_;
};

class Bicubic : public Object {
public:
    explicit Bicubic(const std::array<Vec3, 16>& v) { add_time_sample(v); }
    void add_time_sample(const std::array<Vec3, 16>& v) { v_.push_back(v); }
    int kind() const override { return PSY_OBJ_BICUBIC; }
    int time_samples() const override { return (int)v_.size(); }
    std::vector<float> data() const override {
        std::vector<float> d;
        for (const auto& s : v_)
            for (const Vec3& p : s) d.insert(d.end(), {p.x, p.y, p.z});
        return d;
    }

private:
    std::vector<std::array<Vec3, 16>> v_;
};

// SubdivisionSurface (object/subdivision_surface.hpp:15-71): a Catmull-Clark
// control mesh, verts_per_motion_sample points per time sample.  The scene
// turns it into the limit surface's bicubic patches (patches() returns them,
// as SubdivisionSurface::finalize fills `patches`); every hit on them
// carries the instance's id.  The device path takes one time sample.
class SubdivisionSurface : public Object {
public:
    SubdivisionSurface() = default;
    void set_verts(std::vector<Vec3>&& verts, int verts_per_motion_sample) {
        if (verts_per_motion_sample < 0 || (verts_per_motion_sample == 0 && !verts.empty()) ||
            (verts_per_motion_sample > 0 && verts.size() % (size_t)verts_per_motion_sample != 0))
            throw std::invalid_argument("SubdivisionSurface: vertex count is not a multiple of the points per sample");
        verts_ = std::move(verts);
        per_ = verts_per_motion_sample;
    }
    void set_face_vert_counts(std::vector<int>&& counts) { counts_.assign(counts.begin(), counts.end()); }
    void set_face_vert_indices(std::vector<int>&& indices) { indices_.assign(indices.begin(), indices.end()); }
    int kind() const override { return PSY_OBJ_SUBDIVISION; }
    int time_samples() const override { return per_ > 0 ? (int)(verts_.size() / (size_t)per_) : 1; }
    std::vector<float> data() const override {
        std::vector<float> d;
        for (const Vec3& p : verts_) d.insert(d.end(), {p.x, p.y, p.z});
        return d;
    }
    int verts_per_motion_sample() const { return per_; }
    const std::vector<int32_t>& face_vert_counts() const { return counts_; }
    const std::vector<int32_t>& face_vert_indices() const { return indices_; }
    // The first time sample's limit surface as bicubic Bezier patches
    // (psy_subdivision_patches; host only).
    std::vector<std::array<Vec3, 16>> patches() const {
        const std::vector<float> d = data();
        int64_t n = 0;
        check(psy_subdivision_patches(d.data(), per_, counts_.data(), (int32_t)counts_.size(), indices_.data(),
                                      nullptr, 0, &n));
        std::vector<float> f((size_t)n * 48);
        check(psy_subdivision_patches(d.data(), per_, counts_.data(), (int32_t)counts_.size(), indices_.data(),
                                      f.data(), n, &n));
        std::vector<std::array<Vec3, 16>> out((size_t)n);
        for (int64_t k = 0; k < n; ++k)
            for (int i = 0; i < 16; ++i)
                out[(size_t)k][i] = Vec3{f[48 * k + 3 * i], f[48 * k + 3 * i + 1], f[48 * k + 3 * i + 2]};
        return out;
    }

private:
    static void check(int rc) {
        if (rc != PSY_OK) throw std::runtime_error(std::string("psy_subdivision_patches: ") + psy_last_error());
    }
    std::vector<Vec3> verts_;
    int per_ = 0;
    std::vector<int32_t> counts_, indices_;
};

class SphereLight : public Object {
public:
    SphereLight(std::vector<Vec3> positions, std::vector<float> radii, std::vector<Color> colors)
        : p_(std::move(positions)), r_(std::move(radii)), c_(std::move(colors)) {
        if (p_.empty() || p_.size() != r_.size() || p_.size() != c_.size())
            throw std::invalid_argument("SphereLight: one radius and colour per position");
    }
    int kind() const override { return PSY_OBJ_SPHERE_LIGHT; }
    int time_samples() const override { return (int)p_.size(); }
    std::vector<float> data() const override {
        std::vector<float> d;
        for (size_t i = 0; i < p_.size(); ++i)
            d.insert(d.end(), {p_[i].x, p_[i].y, p_[i].z, r_[i], c_[i].r, c_[i].g, c_[i].b});
        return d;
    }

private:
    std::vector<Vec3> p_;
    std::vector<float> r_;
    std::vector<Color> c_;
};

class RectangleLight : public Object {
public:
    RectangleLight(std::vector<std::pair<float, float>> dimensions, std::vector<Color> colors)
        : dim_(std::move(dimensions)), c_(std::move(colors)) {
        if (dim_.empty() || dim_.size() != c_.size())
            throw std::invalid_argument("RectangleLight: one colour per dimension sample");
    }
    int kind() const override { return PSY_OBJ_RECTANGLE_LIGHT; }
    int time_samples() const override { return (int)dim_.size(); }
    std::vector<float> data() const override {
        std::vector<float> d;
        for (size_t i = 0; i < dim_.size(); ++i)
            d.insert(d.end(), {dim_[i].first, dim_[i].second, c_[i].r, c_[i].g, c_[i].b});
        return d;
    }

private:
    std::vector<std::pair<float, float>> dim_;
    std::vector<Color> c_;
};

// ---- shaders ----------------------------------------------------------------
class SurfaceShader {
public:
    virtual ~SurfaceShader() = default;
    virtual int kind() const = 0;
    Color col;

protected:
    explicit SurfaceShader(Color c) : col(c) {}
};

class LambertShader : public SurfaceShader {
public:
    explicit LambertShader(Color c) : SurfaceShader(c) {}
    int kind() const override { return PSY_SHADER_LAMBERT; }
};

class EmitShader : public SurfaceShader {
public:
    explicit EmitShader(Color c) : SurfaceShader(c) {}
    int kind() const override { return PSY_SHADER_EMIT; }
};

// GTRShader (shading/surface_shader.hpp:55-71): the GTR microfacet closure;
// roughness 0 is a perfect mirror.
class GTRShader : public SurfaceShader {
public:
    GTRShader(Color c, float roughness_, float tail_shape_, float fresnel_)
        : SurfaceShader(c), roughness(roughness_), tail_shape(tail_shape_), fresnel(fresnel_) {}
    int kind() const override { return PSY_SHADER_GTR; }
    float roughness, tail_shape, fresnel;
};

// ---- assembly ---------------------------------------------------------------
class Assembly {
public:
    bool add_surface_shader(const std::string& name, std::unique_ptr<SurfaceShader>&& shader) {
        if (shader_map_.count(name)) return false;
        shader_map_[name] = shaders_.size();
        shaders_.push_back(std::move(shader));
        return true;
    }
    const SurfaceShader* get_surface_shader(const std::string& name) const {
        auto it = shader_map_.find(name);
        return it == shader_map_.end() ? nullptr : shaders_[it->second].get();
    }
    bool add_object(const std::string& name, std::unique_ptr<Object>&& object) {
        if (object_map_.count(name)) return false;
        object_map_[name] = objects_.size();
        objects_.push_back(std::move(object));
        return true;
    }
    bool add_assembly(const std::string& name, std::unique_ptr<Assembly>&& assembly) {
        if (assembly_map_.count(name)) return false;
        assembly_map_[name] = assemblies_.size();
        assemblies_.push_back(std::move(assembly));
        return true;
    }
    bool create_object_instance(const std::string& name, const std::vector<Transform>& transforms,
                                const SurfaceShader* surface_shader = nullptr) {
        auto it = object_map_.find(name);
        if (it == object_map_.end()) return false;
        instances_.push_back({PSY_INSTANCE_OBJECT, (int)it->second, transforms, surface_shader});
        return true;
    }
    bool create_assembly_instance(const std::string& name, const std::vector<Transform>& transforms,
                                  const SurfaceShader* surface_shader = nullptr) {
        auto it = assembly_map_.find(name);
        if (it == assembly_map_.end()) return false;
        instances_.push_back({PSY_INSTANCE_ASSEMBLY, (int)it->second, transforms, surface_shader});
        return true;
    }
    // The device build happens in Scene::finalize.
    bool finalize() { return true; }

private:
    friend struct Scene;
    struct Instance {
        int type, index;
        std::vector<Transform> transforms;
        const SurfaceShader* shader;
    };
    std::vector<std::unique_ptr<SurfaceShader>> shaders_;
    std::vector<std::unique_ptr<Object>> objects_;
    std::vector<std::unique_ptr<Assembly>> assemblies_;
    std::vector<Instance> instances_;
    std::unordered_map<std::string, size_t> shader_map_, object_map_, assembly_map_;
};

// ---- camera, scene ----------------------------------------------------------
struct Camera {
    std::vector<Transform> transforms;
    std::vector<float> fovs;  // radians
    std::vector<float> aperture_radii;
    std::vector<float> focus_distances;
    Camera(std::vector<Transform>& transforms_, std::vector<float>& fovs_, std::vector<float>& aperture_radii_,
           std::vector<float>& focus_distances_)
        : transforms(transforms_), fovs(fovs_), aperture_radii(aperture_radii_), focus_distances(focus_distances_) {}
};

// The device context every Scene, Tracer and integrator of the process
// shares (one per process, on device 0 unless set_device is called first).
class Device {
public:
    static Device& get() {
        static Device d;
        return d;
    }
    static void set_device(int device) { get().device_ = device; }
    psy_ctx* ctx() {
        if (!ctx_) check(psy_ctx_create(device_, &ctx_), "psy_ctx_create");
        return ctx_;
    }
    // Colour upsampling for non-grey colours: the reference's spectrum grid
    // (color/spectra_xyz_5nm_380_780_0.97.h), e.g.
    //   Device::get().set_spectrum_grid((const int32_t*)spectrum_grid, spectrum_grid_width,
    //       spectrum_grid_height, (const float*)spectrum_data_points,
    //       sizeof(spectrum_data_points) / sizeof(spectrum_data_points[0]), spectrum_num_samples,
    //       spectrum_sample_min, spectrum_sample_max, spectrum_mat_xy_to_uv);
    void set_spectrum_grid(const int32_t* cells, int grid_width, int grid_height, const float* points,
                           int n_points, int n_samples, float sample_min, float sample_max,
                           const float* mat_xy_to_uv) {
        check(psy_spectrum_set_grid(ctx(), cells, grid_width, grid_height, points, n_points, n_samples, sample_min,
                                    sample_max, mat_xy_to_uv),
              "psy_spectrum_set_grid");
    }
    ~Device() {
        if (ctx_) psy_ctx_destroy(ctx_);
    }

private:
    int device_ = 0;
    psy_ctx* ctx_ = nullptr;
};

struct Scene {
    std::string name;
    Color background_color;
    std::unique_ptr<Camera> camera;
    std::unique_ptr<Assembly> root;
    float dice_rate = 0.25f;        // Config::dice_rate (config.cpp:7)
    float min_upoly_size = 1e-5f;   // Config::min_upoly_size (config.cpp:8)
    int dice_depth = -1;            // dice cache depth; < 0: from the scene
    // rectangle lights test y against their height (PSY_SCENE_RECT_LIGHT_FIX;
    // the reference tests it against the width)
    bool rect_light_fix = false;

    Scene() : root(new Assembly()) {}
    ~Scene() {
        if (handle) psy_scene_destroy(handle);
    }

    // The flattened description psy_scene_create reads; valid while *this
    // is alive and unchanged.
    struct Description {
        psy_scene_desc desc{};
        std::vector<psy_assembly_desc> assemblies;
        std::vector<psy_shader_desc> shaders;
        std::vector<std::vector<psy_object_desc>> objects;
        std::vector<std::vector<psy_instance_desc>> instances;
        std::vector<std::vector<float>> transforms, object_data;
        std::vector<std::vector<int32_t>> children;
        std::vector<float> cam_xf;
    };
    std::unique_ptr<Description> describe() const {
        if (!camera) throw std::runtime_error("Scene: no camera");
        std::unique_ptr<Description> D(new Description());
        std::vector<const Assembly*> order;
        std::unordered_map<const SurfaceShader*, int> shader_index;
        // assemblies in depth-first order, root first
        std::function<int(const Assembly*)> visit = [&](const Assembly* a) -> int {
            const int me = (int)order.size();
            order.push_back(a);
            for (const auto& sh : a->shaders_) {
                shader_index[sh.get()] = (int)D->shaders.size();
                psy_shader_desc sd{};
                sd.kind = sh->kind();
                sd.color[0] = sh->col.r;
                sd.color[1] = sh->col.g;
                sd.color[2] = sh->col.b;
                if (const GTRShader* g = dynamic_cast<const GTRShader*>(sh.get())) {
                    sd.roughness = g->roughness;
                    sd.tail_shape = g->tail_shape;
                    sd.fresnel = g->fresnel;
                }
                D->shaders.push_back(sd);
            }
            D->children.emplace_back();
            for (const auto& c : a->assemblies_) {
                const int ci = visit(c.get());
                D->children[me].push_back(ci);
            }
            return me;
        };
        visit(root.get());
        const size_t na = order.size();
        D->objects.resize(na);
        D->instances.resize(na);
        D->transforms.resize(na);
        D->object_data.resize(na);
        for (size_t a = 0; a < na; ++a) {
            const Assembly* A = order[a];
            std::vector<size_t> data_off;
            for (const auto& o : A->objects_) {
                data_off.push_back(D->object_data[a].size());
                const std::vector<float> d = o->data();
                D->object_data[a].insert(D->object_data[a].end(), d.begin(), d.end());
            }
            for (size_t i = 0; i < A->objects_.size(); ++i) {
                psy_object_desc od{};
                od.kind = A->objects_[i]->kind();
                od.time_samples = A->objects_[i]->time_samples();
                od.data = D->object_data[a].data() + data_off[i];
                if (const SubdivisionSurface* sd = dynamic_cast<const SubdivisionSurface*>(A->objects_[i].get())) {
                    od.n_vertices = sd->verts_per_motion_sample();
                    od.n_faces = (int32_t)sd->face_vert_counts().size();
                    od.face_vert_counts = sd->face_vert_counts().data();
                    od.face_vert_indices = sd->face_vert_indices().data();
                }
                D->objects[a].push_back(od);
            }
            for (const auto& in : A->instances_) {
                psy_instance_desc id{};
                id.type = in.type;
                id.data_index = in.index;  // object index, or index into this assembly's children
                id.transform_index = (int32_t)(D->transforms[a].size() / 16);
                id.transform_count = (int32_t)in.transforms.size();
                for (const Transform& t : in.transforms)
                    D->transforms[a].insert(D->transforms[a].end(), t.m.begin(), t.m.end());
                id.shader_index = -1;
                if (in.shader) {
                    auto it = shader_index.find(in.shader);
                    if (it == shader_index.end()) throw std::runtime_error("Scene: instance shader not in the scene");
                    id.shader_index = it->second;
                }
                D->instances[a].push_back(id);
            }
        }
        D->assemblies.resize(na);
        for (size_t a = 0; a < na; ++a) {
            psy_assembly_desc& ad = D->assemblies[a];
            ad.objects = D->objects[a].data();
            ad.n_objects = (int32_t)D->objects[a].size();
            ad.instances = D->instances[a].data();
            ad.n_instances = (int32_t)D->instances[a].size();
            ad.transforms = D->transforms[a].data();
            ad.n_transforms = (int32_t)(D->transforms[a].size() / 16);
            ad.assemblies = D->children[a].data();
            ad.n_assemblies = (int32_t)D->children[a].size();
        }
        for (const Transform& t : camera->transforms) D->cam_xf.insert(D->cam_xf.end(), t.m.begin(), t.m.end());
        psy_scene_desc& d = D->desc;
        d.assemblies = D->assemblies.data();
        d.n_assemblies = (int32_t)na;
        d.shaders = D->shaders.data();
        d.n_shaders = (int32_t)D->shaders.size();
        d.camera.transforms = D->cam_xf.data();
        d.camera.n_transforms = (int32_t)camera->transforms.size();
        d.camera.fovs = camera->fovs.data();
        d.camera.n_fovs = (int32_t)camera->fovs.size();
        d.camera.aperture_radii = camera->aperture_radii.data();
        d.camera.n_aperture_radii = (int32_t)camera->aperture_radii.size();
        d.camera.focus_distances = camera->focus_distances.data();
        d.camera.n_focus_distances = (int32_t)camera->focus_distances.size();
        d.background[0] = background_color.r;
        d.background[1] = background_color.g;
        d.background[2] = background_color.b;
        d.dice_rate = dice_rate;
        d.min_upoly_size = min_upoly_size;
        d.flags = rect_light_fix ? PSY_SCENE_RECT_LIGHT_FIX : 0u;
        return D;
    }

    // Scene::finalize: builds the scene on the device.
    void finalize() {
        root->finalize();
        std::unique_ptr<Description> D = describe();
        if (handle) {
            psy_scene_destroy(handle);
            handle = nullptr;
        }
        check(psy_scene_create(Device::get().ctx(), &D->desc, dice_depth, &handle), "psy_scene_create");
        shader_table = D->shaders;
    }

    psy_scene* handle = nullptr;
    std::vector<psy_shader_desc> shader_table;  // the finalized scene's shaders, by index
};

// ---- tracing ----------------------------------------------------------------
struct WorldRay {
    enum Type : uint16_t {
        NONE = 0,
        CAMERA = 1 << 0,
        R_DIFFUSE = 1 << 1,
        R_SPECULAR = 1 << 2,
        T_DIFFUSE = 1 << 3,
        T_SPECULAR = 1 << 4,
        OCCLUSION = 1 << 5
    };
    Vec3 o, d;
    Vec3 odx, ody;
    Vec3 ddx, ddy;
    float time = 0.0f;
    Type type = NONE;
};

// DifferentialGeometry (basics/differential_geometry.hpp:8-71): a hit in
// the space of the object hit.
struct DifferentialGeometry {
    Vec3 p, n;
    float u = 0.0f, v = 0.0f;
    Vec3 dpdu, dpdv, dndu, dndv;
};

namespace detail {
inline Vec3 v3(float x, float y, float z) { return Vec3{x, y, z}; }
inline float dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 scale(const Vec3& a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline Vec3 add(const Vec3& a, const Vec3& b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec3 sub(const Vec3& a, const Vec3& b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 cross(const Vec3& a, const Vec3& b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline Vec3 normalized(const Vec3& a) {
    const float l = std::sqrt(dot(a, a));
    return l > 0.0f ? scale(a, 1.0f / l) : a;
}
constexpr double PI = 3.1415926535897932384626433;
}  // namespace detail

// The closures a hit's shader binds (shading/surface_closure.hpp:23-331).
// The reference evaluates them to SpectralSamples at a hero wavelength;
// here they keep their linear RGB colour and give the factor that
// multiplies it, for the caller's own colour handling.
class SurfaceClosure {
public:
    virtual ~SurfaceClosure() = default;
    virtual bool is_delta() const = 0;
    virtual bool is_emitter() const = 0;
    const Color& color() const { return col_; }

protected:
    explicit SurfaceClosure(Color c) : col_(c) {}
    Color col_;
};

// LambertClosure (shading/surface_closure.hpp:233-331).
class LambertClosure final : public SurfaceClosure {
public:
    explicit LambertClosure(Color c = Color(1.0f)) : SurfaceClosure(c) {}
    bool is_delta() const override { return false; }
    bool is_emitter() const override { return false; }
    // evaluate(): col * max(0, n.out) / pi, n turned to the side of `in`
    Color evaluate(const Vec3& in, const Vec3& out, const DifferentialGeometry& geo) const {
        const float f = factor(in, out, geo);
        return Color(col_.r * f, col_.g * f, col_.b * f);
    }
    float sample_pdf(const Vec3& in, const Vec3& out, const DifferentialGeometry& geo) const {
        Vec3 nn = detail::normalized(geo.n);
        if (detail::dot(nn, detail::normalized(in)) > 0.0f) nn = detail::scale(nn, -1.0f);
        const float c = detail::dot(nn, detail::normalized(out)) * (float)(1.0 / detail::PI);
        return c > 0.0f ? c : 0.0f;
    }
    // sample(): a cosine-distributed direction about the normal (concentric
    // disc mapping of si, sj), its filter colour and pdf
    void sample(const Vec3& in, const DifferentialGeometry& geo, float si, float sj, Vec3* out, Color* filter,
                float* pdf) const {
        Vec3 nn = detail::normalized(geo.n);
        if (detail::dot(nn, in) > 0.0f) nn = detail::scale(nn, -1.0f);
        float x = (si * 2) - 1, y = (sj * 2) - 1;
        square_to_circle(&x, &y);
        const double zz = 1.0 - (double)((x * x) + (y * y));
        const float z = (float)std::sqrt(zz > 0.0 ? zz : 0.0);
        *pdf = z * (float)(1.0 / detail::PI);
        Vec3 tx, ty;
        if (std::fabs(nn.x) > std::fabs(nn.y)) {
            const float il = 1.f / std::sqrt(nn.x * nn.x + nn.z * nn.z);
            tx = detail::v3(-nn.z * il, 0.f, nn.x * il);
        } else {
            const float il = 1.f / std::sqrt(nn.y * nn.y + nn.z * nn.z);
            tx = detail::v3(0.f, nn.z * il, -nn.y * il);
        }
        ty = detail::cross(nn, tx);
        *out = detail::add(detail::add(detail::scale(tx, x), detail::scale(ty, y)), detail::scale(nn, z));
        *filter = evaluate(in, *out, geo);
    }

private:
    float factor(const Vec3& in, const Vec3& out, const DifferentialGeometry& geo) const {
        Vec3 nn = detail::normalized(geo.n);
        if (detail::dot(nn, in) > 0.0f) nn = detail::scale(nn, -1.0f);
        const float c = detail::dot(nn, detail::normalized(out));
        return (c > 0.0f ? c : 0.0f) * (float)(1.0 / detail::PI);
    }
    // square_to_circle (utils/monte_carlo.hpp:29-53)
    static void square_to_circle(float* x, float* y) {
        if (*x == 0.0f && *y == 0.0f) return;
        float radius, angle;
        const double q = detail::PI / 4;
        if (*x > std::fabs(*y)) { radius = *x; angle = (float)(q * (double)(*y / *x)); }
        else if (*y > std::fabs(*x)) { radius = *y; angle = (float)(q * (double)(2 - (*x / *y))); }
        else if (*x < -std::fabs(*y)) { radius = -*x; angle = (float)(q * (double)(4 + (*y / *x))); }
        else { radius = -*y; angle = (float)(q * (double)(6 - (*x / *y))); }
        *x = radius * std::cos(angle);
        *y = radius * std::sin(angle);
    }
};

// EmitClosure (shading/surface_closure.hpp:186-228): emits its colour.
class EmitClosure final : public SurfaceClosure {
public:
    explicit EmitClosure(Color c = Color(1.0f)) : SurfaceClosure(c) {}
    bool is_delta() const override { return false; }
    bool is_emitter() const override { return true; }
    const Color& emitted_color() const { return col_; }
};

// GTRClosure (shading/surface_closure.hpp:339-608), validated as the
// reference's constructor does; evaluate() applies Schlick's Fresnel per RGB
// channel (the reference applies it per wavelength of the spectral colour).
class GTRClosure final : public SurfaceClosure {
public:
    GTRClosure(Color c = Color(1.0f), float roughness = 0.05f, float tail_shape = 2.0f, float fresnel = 1.0f)
        : SurfaceClosure(c), rough_(roughness), tail_(tail_shape), fresnel_(fresnel) {
        rough_ = std::max(0.0f, std::min(0.9999f, rough_));
        tail_ = std::max(0.0001f, tail_);
        if (rough_ < 0.000244140625f) rough_ = 0.0f;
        if (std::abs(tail_ - 1.0f) < 0.0001f) tail_ = 1.0f + 0.0001f;
        norm_ = normalization(rough_, tail_);
    }
    bool is_delta() const override { return rough_ == 0.0f; }
    bool is_emitter() const override { return false; }
    float roughness() const { return rough_; }
    float tail_shape() const { return tail_; }
    float fresnel() const { return fresnel_; }

    Color evaluate(const Vec3& in, const Vec3& out, const DifferentialGeometry& geo) const {
        Vec3 nn = detail::normalized(geo.n);
        const Vec3 aa = detail::scale(detail::normalized(in), -1.0f);
        const Vec3 bb = detail::normalized(out);
        const Vec3 hh = detail::normalized(detail::add(aa, bb));
        if (detail::dot(nn, hh) < 0.0f) nn = detail::scale(nn, -1.0f);
        const float na = clamp(detail::dot(nn, aa)), nb = clamp(detail::dot(nn, bb));
        const float ha = clamp(detail::dot(hh, aa)), hb = clamp(detail::dot(hh, bb));
        const float nh = clamp(detail::dot(nn, hh));
        const float c[3] = {fres(col_.r, hb), fres(col_.g, hb), fres(col_.b, hb)};
        if (rough_ == 0.0f) return Color(c[0], c[1], c[2]);
        const float r2 = rough_ * rough_;
        float D = 1.0f;
        if (nh > 0.0f) D = norm_ / std::pow(1 + ((r2 - 1) * (nh * nh)), tail_);
        const float f = (D * g1(na, ha, r2) * g1(nb, hb, r2)) * (float)(1.0 / detail::PI);
        return Color(c[0] * f, c[1] * f, c[2] * f);
    }
    float sample_pdf(const Vec3& in, const Vec3& out, const DifferentialGeometry& geo) const {
        Vec3 nn = detail::normalized(geo.n);
        const Vec3 aa = detail::scale(detail::normalized(in), -1.0f);
        const Vec3 bb = detail::normalized(out);
        const Vec3 hh = detail::normalized(detail::scale(detail::add(aa, bb), 0.5f));
        if (detail::dot(nn, hh) < 0.0f) nn = detail::scale(nn, -1.0f);
        const float nh = clamp(detail::dot(nn, hh));
        const float r2 = rough_ * rough_;
        const float D = nh > 0.0f ? norm_ / std::pow(1 + ((r2 - 1) * (nh * nh)), tail_) : 0.0f;
        return D * (float)(1.0 / detail::PI);
    }
    // sample(): a half vector from the distribution, `in` mirrored about it
    void sample(const Vec3& in, const DifferentialGeometry& geo, float si, float sj, Vec3* out, Color* filter,
                float* pdf) const {
        Vec3 nn = detail::normalized(geo.n);
        if (detail::dot(nn, in) > 0.0f) nn = detail::scale(nn, -1.0f);
        const float r2 = rough_ * rough_;
        const float top = 1.0f - std::pow((std::pow(r2, 1.0f - tail_) * (1.0f - si)) + si, 1.0f / (1.0f - tail_));
        const float ct = std::sqrt(top / (1.0f - r2));
        const float st = std::sqrt(1.0f - (ct * ct));
        const float angle = (float)((double)sj * detail::PI * 2);
        Vec3 tx, ty;
        if (std::fabs(nn.x) > std::fabs(nn.y)) {
            const float il = 1.f / std::sqrt(nn.x * nn.x + nn.z * nn.z);
            tx = detail::v3(-nn.z * il, 0.f, nn.x * il);
        } else {
            const float il = 1.f / std::sqrt(nn.y * nn.y + nn.z * nn.z);
            tx = detail::v3(0.f, nn.z * il, -nn.y * il);
        }
        ty = detail::cross(nn, tx);
        const Vec3 h = detail::normalized(detail::add(
            detail::add(detail::scale(tx, std::cos(angle) * st), detail::scale(ty, std::sin(angle) * st)),
            detail::scale(nn, ct)));
        *out = detail::sub(in, detail::scale(detail::scale(h, 2.0f), detail::dot(in, h)));
        *pdf = sample_pdf(in, *out, geo);
        *filter = evaluate(in, *out, geo);
    }

private:
    static float clamp(float v) { return std::min(1.0f, std::max(-1.0f, v)); }
    static float normalization(float r, float t) {
        const float r2 = r * r;
        const float top = (t - 1.0f) * (r2 - 1);
        return (float)((double)top / (detail::PI * (double)(1.0f - std::pow(r2, 1.0f - t))));
    }
    // lerp(1 - fresnel, schlick_fresnel_from_fac(c, hb), c)
    float fres(float c, float hb) const {
        const float c1 = 1.0f - hb, c2 = c1 * c1;
        const float s = c + ((1.0f - c) * c1 * c2 * c2);
        const float a = 1.0f - fresnel_;
        return (float)((double)s * (1.0 - (double)a) + (double)(c * a));
    }
    static float g1(float nx, float hx, float r2) {
        const float nx2 = nx * nx;
        const float a = r2 * std::sqrt((1.0f - nx2) / nx2);
        const float b = (std::sqrt(1.0f + a * a) - 1.0f) * 0.5f;
        return ((hx * nx) > 0.0f ? 1.0f : 0.0f) / (1.0f + b);
    }
    float rough_, tail_, fresnel_, norm_;
};

// SurfaceClosureUnion (shading/closure_union.hpp:16-43): storage for one
// closure of any kind; get() is null until init().
class SurfaceClosureUnion {
public:
    void init(const LambertClosure& c) { lambert_ = c; kind_ = 1; }
    void init(const EmitClosure& c) { emit_ = c; kind_ = 2; }
    void init(const GTRClosure& c) { gtr_ = c; kind_ = 3; }
    SurfaceClosure* get() { return const_cast<SurfaceClosure*>(static_cast<const SurfaceClosureUnion*>(this)->get()); }
    const SurfaceClosure* get() const {
        switch (kind_) {
            case 1: return &lambert_;
            case 2: return &emit_;
            case 3: return &gtr_;
            default: return nullptr;
        }
    }

private:
    LambertClosure lambert_;
    EmitClosure emit_;
    GTRClosure gtr_;
    int kind_ = 0;
};

// Intersection (basics/intersection.hpp:22-49), as the reference's tracer
// fills it at the accepted hit (tracer/tracer.cpp:241-250).
struct Intersection {
    bool hit = false;
    uint64_t id = 0;                  // InstanceID::id of the leaf hit
    float t = 1.0f / 0.0f;            // +inf on a miss
    float u = 0.0f, v = 0.0f;         // surface parameters (patches, spheres; geo.u, geo.v)
    DifferentialGeometry geo;         // in the space of the object hit
    Vec3 offset;                      // ray origin offset off the surface (object space)
    float light_pdf = 9999.0f;        // light hits: pdf of sampling this direction
    bool backfacing = false;
    Transform space;                  // world -> object at the ray's time
    SurfaceClosureUnion surface_closure;
    float closure_prob = 1.0f;        // one closure per shader: always selected
    int32_t shader = -1;              // the scene's shader index, -1 for lights / none
    float offset_dim = 0.0f;          // patch leaf extent (offset = n * 1.74 * offset_dim)
    uint32_t instance = 0xffffffffu;  // flattened leaf instance
};

class Tracer {
public:
    explicit Tracer(Scene* scene) : scene_(scene) {
        if (!scene_ || !scene_->handle) throw std::runtime_error("Tracer: the scene is not finalized");
        psy_scene_info info{};
        check(psy_scene_get_info(scene_->handle, &info), "psy_scene_get_info");
        ids_.resize((size_t)info.n_leaf_instances);
        if (!ids_.empty())
            check(psy_scene_instance_ids(scene_->handle, ids_.data(), (int64_t)ids_.size()),
                  "psy_scene_instance_ids");
    }
    ~Tracer() { release(); }
    Tracer(const Tracer&) = delete;
    Tracer& operator=(const Tracer&) = delete;

    // Tracer::trace: closest hit (any hit for OCCLUSION rays) of every ray.
    uint32_t trace(const WorldRay* w_rays_begin, const WorldRay* w_rays_end, Intersection* intersections_begin,
                   Intersection* intersections_end) {
        const int64_t n = w_rays_end - w_rays_begin;
        if (n != intersections_end - intersections_begin)
            throw std::invalid_argument("Tracer::trace: one intersection per ray");
        if (n == 0) return 0;
        reserve(n);
        // WorldRay records to the 20-plane stream layout
        host_.assign((size_t)(PSY_WRAY_PLANES * n), 0.0f);
        for (int64_t i = 0; i < n; ++i) {
            const WorldRay& r = w_rays_begin[i];
            const float v[18] = {r.o.x,   r.o.y,   r.o.z,   r.d.x,   r.d.y,   r.d.z,   r.odx.x, r.odx.y, r.odx.z,
                                 r.ody.x, r.ody.y, r.ody.z, r.ddx.x, r.ddx.y, r.ddx.z, r.ddy.x, r.ddy.y, r.ddy.z};
            for (int p = 0; p < 18; ++p) host_[(size_t)(p * n + i)] = v[p];
            host_[(size_t)(18 * n + i)] = r.time;
            const uint32_t ty = (uint32_t)r.type;
            std::memcpy(&host_[(size_t)(19 * n + i)], &ty, 4);
        }
        psy_ctx* ctx = Device::get().ctx();
        check(psy_copy_to_device(ctx, rays_, host_.data(), sizeof(float) * host_.size()), "Tracer::trace");
        float* h = (float*)hits_;
        psy_hits hits{h, (uint32_t*)(h + n), h + 2 * n, h + 3 * n, h + 4 * n};
        check(psy_trace_world(ctx, scene_->handle, (const float*)rays_, n, &hits, 0u, nullptr), "psy_trace_world");
        check(psy_intersections(ctx, scene_->handle, (const float*)rays_, n, &hits, (float*)isect_, nullptr),
              "psy_intersections");
        host_.resize((size_t)(5 * n));
        check(psy_copy_to_host(ctx, host_.data(), hits_, sizeof(float) * 5 * (size_t)n), "Tracer::trace");
        rec_.resize((size_t)(PSY_ISECT_PLANES * n));
        check(psy_copy_to_host(ctx, rec_.data(), isect_, sizeof(float) * rec_.size()), "Tracer::trace");
        auto at = [&](int plane, int64_t i) { return rec_[(size_t)(plane * n + i)]; };
        auto at_u = [&](int plane, int64_t i) {
            uint32_t u;
            std::memcpy(&u, &rec_[(size_t)(plane * n + i)], 4);
            return u;
        };
        for (int64_t i = 0; i < n; ++i) {
            Intersection& x = intersections_begin[i];
            x = Intersection{};
            uint32_t inst;
            std::memcpy(&inst, &host_[(size_t)(n + i)], 4);
            x.hit = inst != 0xffffffffu;
            x.instance = inst;
            x.t = host_[(size_t)i];
            x.u = host_[(size_t)(2 * n + i)];
            x.v = host_[(size_t)(3 * n + i)];
            x.offset_dim = host_[(size_t)(4 * n + i)];
            if (!x.hit) continue;
            x.id = ids_[inst];
            Vec3* vs[7] = {&x.geo.p, &x.geo.n, &x.geo.dpdu, &x.geo.dpdv, &x.geo.dndu, &x.geo.dndv, &x.offset};
            for (int k = 0; k < 7; ++k) *vs[k] = Vec3{at(3 * k, i), at(3 * k + 1, i), at(3 * k + 2, i)};
            x.geo.u = x.u;
            x.geo.v = x.v;
            for (int k = 0; k < 16; ++k) x.space.m[(size_t)k] = at(21 + k, i);
            x.light_pdf = at(37, i);
            x.backfacing = at_u(38, i) != 0;
            const Color col(at(40, i), at(41, i), at(42, i));
            const uint32_t kind = at_u(39, i);
            const int32_t sh = (int32_t)at_u(43, i);
            if (kind == PSY_SHADER_GTR && sh >= 0 && (size_t)sh < scene_->shader_table.size()) {
                const psy_shader_desc& sd = scene_->shader_table[(size_t)sh];
                x.surface_closure.init(GTRClosure(col, sd.roughness, sd.tail_shape, sd.fresnel));
            } else if (kind == PSY_SHADER_LAMBERT) x.surface_closure.init(LambertClosure(col));
            else x.surface_closure.init(EmitClosure(col));
            x.shader = (int32_t)at_u(43, i);
        }
        return (uint32_t)n;
    }

private:
    void reserve(int64_t n) {
        if (n <= cap_) return;
        release();
        psy_ctx* ctx = Device::get().ctx();
        check(psy_malloc(ctx, sizeof(float) * PSY_WRAY_PLANES * (size_t)n, &rays_), "psy_malloc");
        check(psy_malloc(ctx, sizeof(float) * 5 * (size_t)n, &hits_), "psy_malloc");
        check(psy_malloc(ctx, sizeof(float) * PSY_ISECT_PLANES * (size_t)n, &isect_), "psy_malloc");
        cap_ = n;
    }
    void release() {
        if (rays_) psy_free(Device::get().ctx(), rays_);
        if (hits_) psy_free(Device::get().ctx(), hits_);
        if (isect_) psy_free(Device::get().ctx(), isect_);
        rays_ = hits_ = isect_ = nullptr;
        cap_ = 0;
    }
    Scene* scene_;
    std::vector<uint64_t> ids_;
    std::vector<float> host_, rec_;
    void* rays_ = nullptr;
    void* hits_ = nullptr;
    void* isect_ = nullptr;
    int64_t cap_ = 0;
};

// ---- film, integrator -------------------------------------------------------
// Film: per-pixel CIE XYZ (the mean of the pixel's samples), row-major.
struct Film {
    uint16_t width, height;
    std::vector<float> xyz;  // width * height * 3
    Film(uint16_t w, uint16_t h) : width(w), height(h), xyz((size_t)w * h * 3, 0.0f) {}
};

class PathTraceIntegrator {
public:
    // spp_max and variance_max drive the reference's adaptive loop over
    // buckets (psy_render_adaptive); thread_count sets the bucket size, as
    // in the reference, and nothing else.
    PathTraceIntegrator(Scene* scene_, Film* image_, int spp_, int spp_max_, float variance_max_, uint32_t seed_,
                        int thread_count_ = 1, std::function<void()> callback_ = std::function<void()>())
        : scene(scene_), image(image_), spp(spp_), spp_max(spp_max_), variance_max(variance_max_), seed(seed_),
          thread_count(thread_count_), callback(std::move(callback_)) {}
    int path_length = 4;  // the reference integrator's bounces

    void integrate() {
        if (!scene || !scene->handle) throw std::runtime_error("PathTraceIntegrator: the scene is not finalized");
        psy_ctx* ctx = Device::get().ctx();
        const size_t bytes = sizeof(float) * image->xyz.size();
        void* film = nullptr;
        check(psy_malloc(ctx, bytes, &film), "psy_malloc");
        psy_render_params p{};
        p.width = image->width;
        p.height = image->height;
        p.spp = spp;
        p.seed = seed;
        p.path_length = path_length;
        int64_t rays = 0;
        // with spp_max above spp, the adaptive loop (whose single pass is
        // psy_render's film)
        int rc = spp_max > spp ? psy_render_adaptive(ctx, scene->handle, &p, spp_max, variance_max, thread_count,
                                                     (float*)film, nullptr, nullptr, &rays, nullptr)
                               : psy_render(ctx, scene->handle, &p, (float*)film, nullptr, &rays, nullptr);
        if (rc == PSY_OK) rc = psy_copy_to_host(ctx, image->xyz.data(), film, bytes);
        psy_free(ctx, film);
        check(rc, "PathTraceIntegrator::integrate");
        rays_shot = rays;
        if (callback) callback();
    }

    Scene* scene;
    Film* image;
    int spp, spp_max;
    float variance_max;
    uint32_t seed;
    int thread_count;
    std::function<void()> callback;
    int64_t rays_shot = 0;  // Global::Stats::rays_shot of the last integrate()
};

// write_png_from_film (renderer/renderer.cpp:24-42): the film as dithered
// 8-bit sRGB (Film::scanline_image_8bbc) in a PNG file.
inline void write_png(const Film& film, const std::string& path, uint32_t dither_seed = 1) {
    std::vector<uint8_t> rgb((size_t)film.width * film.height * 3);
    check(psy_film_to_srgb8(film.xyz.data(), nullptr, film.width, film.height, dither_seed, rgb.data()),
          "psy_film_to_srgb8");
    check(psy_write_png(path.c_str(), rgb.data(), film.width, film.height), "psy_write_png");
}

// One parsed frame (renderer/renderer.hpp): its scene, already built on the
// device, and its render settings.  render() fills a Film and, with an
// output path, writes it as a PNG there.
struct Renderer {
    std::unique_ptr<Scene> scene;
    int res_x = 1, res_y = 1, spp = 1;
    uint32_t seed = 1;
    std::string output_path;
    std::string warnings;  // what the parser skipped, one line each

    void render(Film& film, int path_length = 4, bool write_output = false) {
        PathTraceIntegrator integ(scene.get(), &film, spp, spp, 0.0f, seed);
        integ.path_length = path_length;
        integ.integrate();
        if (write_output && !output_path.empty()) write_png(film, output_path);
    }
};

// Parser (parser/parser.hpp:16-91) over a .psy file: parse_next_frame()
// returns the next Scene node as a Renderer, nullptr after the last.
class Parser {
public:
    explicit Parser(const std::string& input_path) {
        std::ifstream in(input_path, std::ios::in | std::ios::binary);
        if (!in) throw std::runtime_error("Parser: cannot read " + input_path);
        std::ostringstream ss;
        ss << in.rdbuf();
        text_ = ss.str();
        int32_t n = 0;
        check(psy_parse_scene_count(text_.data(), (int64_t)text_.size(), &n), "psy_parse_scene_count");
        count_ = n;
    }
    ~Parser() { release(); }
    Parser(const Parser&) = delete;
    Parser& operator=(const Parser&) = delete;

    int frame_count() const { return count_; }

    // The next frame's description and settings, without building it on
    // the device (valid until the next call or the Parser's end).
    const psy_scene_desc* parse_next_description(psy_render_settings* settings) {
        release();
        if (next_ >= count_) return nullptr;
        check(psy_parse_scene(text_.data(), (int64_t)text_.size(), next_++, &parsed_), "psy_parse_scene");
        if (settings) check(psy_parsed_scene_settings(parsed_, settings), "psy_parsed_scene_settings");
        return psy_parsed_scene_desc(parsed_);
    }

    std::unique_ptr<Renderer> parse_next_frame(int dice_depth = -1) {
        psy_render_settings st{};
        const psy_scene_desc* d = parse_next_description(&st);
        if (!d) return nullptr;
        std::unique_ptr<Renderer> r(new Renderer());
        r->scene.reset(new Scene());
        r->scene->name = st.name;
        r->scene->dice_rate = d->dice_rate;
        r->scene->background_color = Color(d->background[0], d->background[1], d->background[2]);
        check(psy_scene_create(Device::get().ctx(), d, dice_depth, &r->scene->handle), "psy_scene_create");
        r->res_x = st.width;
        r->res_y = st.height;
        r->spp = st.spp;
        r->seed = (uint32_t)st.seed;
        r->output_path = st.output_path;
        r->warnings = psy_parsed_scene_warnings(parsed_);
        return r;
    }

private:
    void release() {
        if (parsed_) psy_parsed_scene_destroy(parsed_);
        parsed_ = nullptr;
    }
    std::string text_;
    int count_ = 0, next_ = 0;
    psy_parsed_scene* parsed_ = nullptr;
};

}  // namespace psy

#endif  // PSYCHOPATH_GPU_HPP
