// TEST PROGRAM for include/psychopath_gpu.hpp: builds a scene through the
// reference-style C++ classes (Assembly, Scene, Camera, shaders, objects,
// an instanced sub-assembly) and
//   describe           prints the flattened psy_scene_desc, one fact a line
//   nodevice           finalizes without a GPU and expects the library's error
//   gpu <out>          finalizes, traces a grid of camera rays with
//                      Tracer::trace and renders with PathTraceIntegrator,
//                      writing the results to <out> (raw little-endian)
//   psyparse <file>    reads a .psy file with Parser and prints every
//                      frame's settings and description (no GPU)
//   psyrender <file> <out>
//                      renders the file's first frame with Parser /
//                      Renderer and writes its film to <out>
// tests/test_cpp_api.py and tests/test_gpu_cpp_api.py build the same scene
// in Python and compare.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>

#include "psychopath_gpu.hpp"

using namespace psy;

// World -> local transform of an instance placed at (x, y, z) (instance
// transforms map the parent's space to the instance's).
static Transform translate(float x, float y, float z) {
    Transform t;
    t.m[12] = -x;
    t.m[13] = -y;
    t.m[14] = -z;
    return t;
}

static void print_desc(const psy_scene_desc& d) {
    std::printf("assemblies %d\nshaders %d\n", d.n_assemblies, d.n_shaders);
    for (int i = 0; i < d.n_shaders; ++i)
        std::printf("shader %d %d %.7g %.7g %.7g\n", i, d.shaders[i].kind, d.shaders[i].color[0],
                    d.shaders[i].color[1], d.shaders[i].color[2]);
    for (int a = 0; a < d.n_assemblies; ++a) {
        const psy_assembly_desc& A = d.assemblies[a];
        double xs = 0.0;
        for (int k = 0; k < 16 * A.n_transforms; ++k) xs += A.transforms[k] * (double)(k % 16 + 1);
        std::printf("assembly %d objects %d instances %d transforms %d children %d xfsum %.6f\n", a, A.n_objects,
                    A.n_instances, A.n_transforms, A.n_assemblies, xs);
        for (int i = 0; i < A.n_objects; ++i) {
            const psy_object_desc& o = A.objects[i];
            const int nf = o.kind == PSY_OBJ_SPHERE ? 4 : o.kind == PSY_OBJ_BILINEAR ? 12
                           : o.kind == PSY_OBJ_BICUBIC ? 48 : o.kind == PSY_OBJ_SPHERE_LIGHT ? 7
                           : o.kind == PSY_OBJ_SUBDIVISION ? 3 * o.n_vertices : 5;
            double sum = 0.0;
            for (int k = 0; k < nf * o.time_samples; ++k) sum += o.data[k];
            std::printf("object %d %d kind %d ts %d sum %.6f\n", a, i, o.kind, o.time_samples, sum);
            if (o.kind == PSY_OBJ_SUBDIVISION) {
                long long sc = 0, si = 0;
                for (int f = 0; f < o.n_faces; ++f) sc += o.face_vert_counts[f];
                for (long long k = 0; k < sc; ++k) si += o.face_vert_indices[k] * (k + 1);
                std::printf("mesh %d %d verts %d faces %d counts %lld indices %lld\n", a, i, o.n_vertices, o.n_faces,
                            sc, si);
            }
        }
        for (int i = 0; i < A.n_instances; ++i)
            std::printf("instance %d %d type %d data %d xf %d+%d shader %d\n", a, i, A.instances[i].type,
                        A.instances[i].data_index, A.instances[i].transform_index,
                        A.instances[i].transform_count, A.instances[i].shader_index);
    }
}

// A Catmull-Clark cube of half-size 0.6: its limit surface is a rounded blob.
static std::unique_ptr<Object> make_blob() {
    std::unique_ptr<SubdivisionSurface> s(new SubdivisionSurface());
    std::vector<Vec3> v;
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y)
            for (int z = 0; z < 2; ++z) v.push_back(Vec3{x ? 0.6f : -0.6f, y ? 0.6f : -0.6f, z ? 0.6f : -0.6f});
    s->set_verts(std::move(v), 8);
    s->set_face_vert_counts({4, 4, 4, 4, 4, 4});
    s->set_face_vert_indices({0, 2, 6, 4, 1, 5, 7, 3, 0, 4, 5, 1, 2, 3, 7, 6, 0, 1, 3, 2, 4, 6, 7, 5});
    return std::unique_ptr<Object>(s.release());
}

static std::unique_ptr<Scene> make_scene() {
    std::unique_ptr<Scene> sc(new Scene());
    const float cam[16] = {1.0f, -0.0f, 0.0f, 0.0f, 0.0f, 0.4472135901451111f, 0.8944271802902222f, 0.0f,
                           0.0f, 0.8944271802902222f, -0.4472135901451111f, 0.0f, 0.0f, -9.0f, 5.0f, 1.0f};
    std::vector<Transform> cx{Transform(cam)};
    std::vector<float> fov{0.8726646f};
    std::vector<float> ap{0.0f};
    std::vector<float> fd{10.0f};
    sc->camera.reset(new Camera(cx, fov, ap, fd));
    sc->background_color = Color(0.05f);
    Assembly& root = *sc->root;
    root.add_surface_shader("grey", std::unique_ptr<SurfaceShader>(new LambertShader(Color(0.6f))));
    root.add_surface_shader("dark", std::unique_ptr<SurfaceShader>(new LambertShader(Color(0.3f))));
    root.add_object("ground", std::unique_ptr<Object>(new Bilinear({-8, -8, 0}, {8, -8, 0}, {-8, 8, 0}, {8, 8, 0})));
    root.create_object_instance("ground", {}, root.get_surface_shader("grey"));
    root.add_object("ball", std::unique_ptr<Object>(new Sphere({0.0f, 0.0f, 1.0f}, 1.0f)));
    root.create_object_instance("ball", {}, root.get_surface_shader("dark"));
    std::array<Vec3, 16> cp;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i)
            cp[j * 4 + i] = Vec3{-3.0f + (float)i, 2.0f + (float)j, (i == 1 || i == 2) && (j == 1 || j == 2) ? 1.5f : 0.2f};
    root.add_object("hill", std::unique_ptr<Object>(new Bicubic(cp)));
    root.create_object_instance("hill", {translate(0.5f, 0.0f, 0.0f)}, root.get_surface_shader("grey"));
    root.add_object("blob", make_blob());
    root.create_object_instance("blob", {translate(2.5f, 2.0f, 0.8f)}, root.get_surface_shader("grey"));
    std::unique_ptr<Assembly> lamp(new Assembly());
    lamp->add_object("bulb", std::unique_ptr<Object>(new SphereLight({{0.0f, 0.0f, 0.0f}}, {0.4f}, {Color(30.0f)})));
    lamp->create_object_instance("bulb", {}, nullptr);
    lamp->add_object("panel", std::unique_ptr<Object>(new RectangleLight({{1.0f, 1.0f}}, {Color(10.0f)})));
    lamp->create_object_instance("panel", {translate(0.0f, 0.0f, 1.0f)}, nullptr);
    root.add_assembly("lamp", std::move(lamp));
    root.create_assembly_instance("lamp", {translate(-2.5f, 1.0f, 4.0f)}, nullptr);
    root.create_assembly_instance("lamp", {translate(2.0f, -1.0f, 3.0f)}, nullptr);
    return sc;
}

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "describe";
    std::unique_ptr<Scene> sc = make_scene();
    if (mode == "describe") {
        auto D = sc->describe();
        print_desc(D->desc);
        return 0;
    }
    if (mode == "psyparse" && argc > 2) {
        Parser parser(argv[2]);
        psy_render_settings st{};
        for (const psy_scene_desc* d; (d = parser.parse_next_description(&st));) {
            std::printf("frame %s %dx%d spp %d seed %d dice %.7g output %s\n", st.name, st.width, st.height, st.spp,
                        st.seed, st.dice_rate, st.output_path);
            print_desc(*d);
        }
        return 0;
    }
    if (mode == "psyrender" && argc > 3) {
        Parser parser(argv[2]);
        std::unique_ptr<Renderer> r = parser.parse_next_frame();
        if (!r) return 3;
        Film film((uint16_t)r->res_x, (uint16_t)r->res_y);
        r->render(film);
        std::ofstream f(argv[3], std::ios::binary);
        const int32_t hdr[3] = {r->res_x, r->res_y, r->spp};
        f.write((const char*)hdr, sizeof hdr);
        f.write((const char*)film.xyz.data(), sizeof(float) * film.xyz.size());
        std::printf("rendered %s %dx%d spp %d\n", r->scene->name.c_str(), r->res_x, r->res_y, r->spp);
        return 0;
    }
    if (mode == "subdiv") {
        // SubdivisionSurface::patches on the host: count and checksum
        std::unique_ptr<Object> b = make_blob();
        const auto P = static_cast<SubdivisionSurface*>(b.get())->patches();
        double sum = 0.0;
        for (size_t k = 0; k < P.size(); ++k)
            for (int i = 0; i < 16; ++i) sum += (P[k][i].x + 2.0 * P[k][i].y + 3.0 * P[k][i].z) * (double)(k % 7 + 1);
        std::printf("patches %zu sum %.17g\n", P.size(), sum);
        return 0;
    }
    if (mode == "nodevice") {
        try {
            sc->finalize();
        } catch (const std::runtime_error& e) {
            std::printf("error %s\n", e.what());
            return 0;
        }
        std::printf("finalized\n");
        return 0;
    }
    if (mode == "gpu" && argc > 2) {
        sc->finalize();
        // a 32 x 24 grid of pinhole camera rays through pixel centres
        const int W = 32, H = 24;
        std::vector<WorldRay> rays((size_t)W * H);
        const float tf = std::tan(0.8726646f * 0.5f);
        const float* m = sc->camera->transforms[0].m.data();
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const float sx = (((float)x + 0.5f) / W * 2.0f - 1.0f) * tf;
                const float sy = -(((float)y + 0.5f) / H * 2.0f - 1.0f) * tf * H / W;
                WorldRay& r = rays[(size_t)y * W + x];
                // camera space (sx, sy, 1) to world: row vector times the matrix
                r.d = Vec3{sx * m[0] + sy * m[4] + m[8], sx * m[1] + sy * m[5] + m[9], sx * m[2] + sy * m[6] + m[10]};
                r.o = Vec3{m[12], m[13], m[14]};
                r.ddx = Vec3{0.002f, 0.0f, 0.0f};
                r.ddy = Vec3{0.0f, 0.002f, 0.0f};
                r.type = WorldRay::CAMERA;
            }
        Tracer tracer(sc.get());
        std::vector<Intersection> hits(rays.size());
        const uint32_t n = tracer.trace(rays.data(), rays.data() + rays.size(), hits.data(), hits.data() + hits.size());
        Film film(W, H);
        PathTraceIntegrator integ(sc.get(), &film, 4, 4, 0.0f, 1u, 8);
        integ.integrate();
        std::ofstream f(argv[2], std::ios::binary);
        const int32_t hdr[3] = {(int32_t)n, W, H};
        f.write((const char*)hdr, sizeof hdr);
        for (const WorldRay& r : rays) {
            const float v[20] = {r.o.x,   r.o.y,   r.o.z,   r.d.x,   r.d.y,   r.d.z,   r.odx.x, r.odx.y, r.odx.z, r.ody.x,
                                 r.ody.y, r.ody.z, r.ddx.x, r.ddx.y, r.ddx.z, r.ddy.x, r.ddy.y, r.ddy.z, r.time, 0.0f};
            f.write((const char*)v, sizeof v);
        }
        for (const Intersection& h : hits) {
            const float tv[3] = {h.t, h.u, h.v};
            f.write((const char*)tv, sizeof tv);
            f.write((const char*)&h.id, sizeof h.id);
            f.write((const char*)&h.instance, sizeof h.instance);
        }
        f.write((const char*)film.xyz.data(), sizeof(float) * film.xyz.size());
        f.write((const char*)&integ.rays_shot, sizeof integ.rays_shot);
        // the rest of each Intersection, in the plane order of psy_intersections
        for (const Intersection& h : hits) {
            float rec[PSY_ISECT_PLANES] = {};
            if (h.hit) {
                const Vec3* vs[7] = {&h.geo.p, &h.geo.n, &h.geo.dpdu, &h.geo.dpdv, &h.geo.dndu, &h.geo.dndv, &h.offset};
                for (int k = 0; k < 7; ++k) {
                    rec[3 * k] = vs[k]->x; rec[3 * k + 1] = vs[k]->y; rec[3 * k + 2] = vs[k]->z;
                }
                for (int k = 0; k < 16; ++k) rec[21 + k] = h.space.m[(size_t)k];
                rec[37] = h.light_pdf;
                const uint32_t bf = h.backfacing ? 1u : 0u;
                const SurfaceClosure* c = h.surface_closure.get();
                const uint32_t kind = c ? (c->is_emitter() ? (uint32_t)PSY_SHADER_EMIT : (uint32_t)PSY_SHADER_LAMBERT) : 0u;
                std::memcpy(&rec[38], &bf, 4);
                std::memcpy(&rec[39], &kind, 4);
                if (c) { rec[40] = c->color().r; rec[41] = c->color().g; rec[42] = c->color().b; }
                std::memcpy(&rec[43], &h.shader, 4);
            }
            f.write((const char*)rec, sizeof rec);
        }
        std::printf("traced %u rays, rendered %lld rays\n", n, (long long)integ.rays_shot);
        return 0;
    }
    std::fprintf(stderr, "usage: test_cpp_api describe | nodevice | gpu <out> | psyparse <file> | "
                         "psyrender <file> <out>\n");
    return 2;
}
