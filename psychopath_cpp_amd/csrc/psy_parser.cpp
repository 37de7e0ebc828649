// psy_parser.cpp -- .psy scene files: the DataTree text format and the scene
// builder on top of it, producing the psy_scene_desc that psy_scene_create
// takes.  Host code only.
//
// Restates:
//   DataTree (lexer, parse_node, print_tree)   parser/data_tree.cpp:1-400
//   Parser::parse_next_frame, parse_matrix, parse_camera, parse_assembly,
//   parse_bilinear_patch, parse_bicubic_patch, parse_surface_shader,
//   parse_sphere_light, parse_rectangle_light, parse_sphere,
//   parse_subdivision_surface                  parser/parser.cpp:37-662
//   Assembly name lookups (object/assembly maps per assembly, shaders
//   searched up the parent chain)              scene/assembly.hpp:105-190
//
// The text is handled byte-wise: bytes >= 0x80 (UTF-8 lead and continuation
// bytes) are never whitespace, comment or reserved characters, so they group
// exactly as the reference's code-point iteration groups them.
//
// Numbers are found as the reference finds them, by scanning leaf text for
// the patterns -?[0-9]+[.]?[0-9]* (floats, read with strtof) and -?[0-9]+
// (integers): "1e-5" reads as 1 and -5, ".5" as 5.
#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "psy_internal.hpp"

namespace {

// ---------------------------------------------------------------------------
// DataTree
// ---------------------------------------------------------------------------
struct DtNode {
    std::string type, name;
    std::vector<DtNode> children;
    std::string leaf;
    bool is_leaf = false;
    int line = 0;
};

enum TokKind { T_OPEN_INNER, T_CLOSE_INNER, T_OPEN_LEAF, T_CLOSE_LEAF, T_TYPE, T_NAME, T_END, T_UNKNOWN };

struct Lexer {
    const char* p;
    const char* end;
    int line = 1;
    std::string err;

    static bool ws(char c) { return c == ' ' || c == '\t'; }
    static bool nl(char c) { return c == '\n' || c == '\r'; }
    static bool reserved(char c) { return c == '{' || c == '}' || c == '[' || c == ']' || c == '\\' || c == '$'; }
    static bool ident(char c) { return !ws(c) && !nl(c) && !reserved(c) && c != '#'; }

    bool at_end() const { return p >= end; }
    char get() {
        const char c = *p++;
        if (c == '\n') ++line;
        return c;
    }
    // whitespace and comments (a comment runs to the end of its line)
    void skip_space() {
        while (!at_end()) {
            if (ws(*p) || nl(*p)) {
                get();
            } else if (*p == '#') {
                while (!at_end() && !nl(*p)) get();
            } else {
                break;
            }
        }
    }
    TokKind next(std::string& text) {
        text.clear();
        skip_space();
        if (at_end()) return T_END;
        const char c = *p;
        if (c == '$') {
            // a name: '$' then identifier characters, with escapes
            text.push_back(get());
            while (!at_end()) {
                if (*p == '\\') {
                    get();
                    if (at_end()) break;
                    text.push_back(get());
                    continue;
                }
                if (!ident(*p)) break;
                text.push_back(get());
            }
            return T_NAME;
        }
        if (reserved(c)) {
            text.push_back(get());
            switch (c) {
                case '{': return T_OPEN_INNER;
                case '}': return T_CLOSE_INNER;
                case '[': return T_OPEN_LEAF;
                case ']': return T_CLOSE_LEAF;
                default: return T_UNKNOWN;  // a stray backslash
            }
        }
        while (!at_end() && ident(*p)) text.push_back(get());
        return T_TYPE;
    }
    // Leaf contents after '[': runs of whitespace and comments become one
    // space, escaped characters are taken literally, leading and trailing
    // space is dropped; stops before the closing ']'.
    std::string leaf() {
        std::string out;
        skip_space();
        while (!at_end()) {
            const char c = *p;
            if (ws(c) || nl(c) || c == '#') {
                skip_space();
                out.push_back(' ');
                continue;
            }
            if (c == '\\') {
                get();
                if (at_end()) break;
                out.push_back(get());
                continue;
            }
            if (c == ']') break;
            out.push_back(get());
        }
        if (!out.empty() && out.back() == ' ') out.pop_back();
        return out;
    }
};

bool parse_node(Lexer& lx, const std::string& type, DtNode& node, int depth) {
    node.type = type;
    node.line = lx.line;
    if (depth > 256) {
        lx.err = "nesting deeper than 256";
        return false;
    }
    std::string tok;
    TokKind k = lx.next(tok);
    if (k == T_NAME) {
        node.name = tok;
        k = lx.next(tok);
    }
    if (k == T_OPEN_INNER) {
        for (;;) {
            k = lx.next(tok);
            if (k == T_TYPE) {
                node.children.emplace_back();
                if (!parse_node(lx, tok, node.children.back(), depth + 1)) return false;
            } else if (k == T_CLOSE_INNER) {
                return true;
            } else if (k == T_END) {
                lx.err = "expected a node type or '}' in " + node.type + ", reached the end of the text";
                return false;
            } else {
                lx.err = "expected a node type or '}' in " + node.type + ", got '" + tok + "'";
                return false;
            }
        }
    }
    if (k == T_OPEN_LEAF) {
        node.is_leaf = true;
        node.leaf = lx.leaf();
        k = lx.next(tok);
        if (k != T_CLOSE_LEAF) {
            lx.err = "expected ']' to close " + node.type;
            return false;
        }
        return true;
    }
    lx.err = "expected '{' or '[' after " + node.type;
    return false;
}

// DataTree::build_from_file on text: the root's children are the top-level
// nodes; parsing stops at the first token that is not a node type.
bool build_tree(const char* text, int64_t len, DtNode& root, std::string& err, int& line) {
    Lexer lx{text, text + len};
    root = DtNode{};
    root.type = "ROOT";
    std::string tok;
    for (;;) {
        const TokKind k = lx.next(tok);
        if (k != T_TYPE) break;
        root.children.emplace_back();
        if (!parse_node(lx, tok, root.children.back(), 0)) {
            err = lx.err;
            line = lx.line;
            return false;
        }
    }
    return true;
}

// DataTree::print_tree
void print_tree(const DtNode& n, const std::string& indent, std::string& out) {
    if (!n.children.empty()) {
        out += indent + n.type + " ";
        if (!n.name.empty()) out += n.name + " ";
        out += "{\n";
        for (const DtNode& c : n.children) print_tree(c, indent + "    ", out);
        out += indent + "}\n";
    } else {
        out += indent + n.type + " [" + n.leaf + "]\n";
    }
}

// ---------------------------------------------------------------------------
// Numbers in leaf text
// ---------------------------------------------------------------------------
bool isdig(char c) { return c >= '0' && c <= '9'; }

// every match of -?[0-9]+[.]?[0-9]* (floats) or -?[0-9]+ (ints), leftmost
// first, as strings
std::vector<std::string> scan_numbers(const std::string& s, bool allow_point) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        size_t j = i;
        if (s[j] == '-') ++j;
        if (j < s.size() && isdig(s[j])) {
            while (j < s.size() && isdig(s[j])) ++j;
            if (allow_point && j < s.size() && s[j] == '.') {
                ++j;
                while (j < s.size() && isdig(s[j])) ++j;
            }
            out.push_back(s.substr(i, j - i));
            i = j;
        } else {
            ++i;
        }
    }
    return out;
}
std::vector<float> floats(const std::string& s) {
    std::vector<float> v;
    for (const std::string& t : scan_numbers(s, true)) v.push_back(std::strtof(t.c_str(), nullptr));
    return v;
}
bool ints(const std::string& s, std::vector<int>& v) {
    v.clear();
    for (const std::string& t : scan_numbers(s, false)) {
        errno = 0;
        const long x = std::strtol(t.c_str(), nullptr, 10);
        if (errno == ERANGE || x > 2147483647L || x < -2147483647L - 1) return false;
        v.push_back((int)x);
    }
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// The parsed scene: a psy_scene_desc and the storage it points into
// ---------------------------------------------------------------------------
struct psy_parsed_scene {
    struct Asm {
        std::vector<psy_object_desc> objects;
        std::vector<std::vector<float>> object_data;
        // a subdivision surface's face vertex counts and indices (empty for
        // the other kinds)
        std::vector<std::vector<int32_t>> object_counts, object_indices;
        std::vector<psy_instance_desc> instances;
        std::vector<float> transforms;
        std::vector<int32_t> children;  // scene indices of the child assemblies
        // name lookups (the first definition of a name wins, as with the
        // reference's map::emplace)
        std::map<std::string, int> object_map, assembly_map, shader_map;
        int parent = -1;
    };
    std::vector<Asm> asms;
    std::vector<psy_shader_desc> shaders;  // renderable shaders, scene-wide
    std::vector<psy_assembly_desc> asm_desc;
    std::vector<float> cam_xf, cam_fov, cam_ap, cam_fd;
    psy_scene_desc desc{};
    psy_render_settings settings{};
    std::string warnings;
};

namespace {

using PS = psy_parsed_scene;

// a shader binding: index into PS::shaders, -1 for none

struct Builder {
    PS& ps;
    std::string err;
    int err_line = 0;

    void warn(const DtNode& n, const std::string& msg) {
        ps.warnings += "line " + std::to_string(n.line) + ": " + msg + "\n";
    }
    bool fail(const DtNode& n, const std::string& msg) {
        err = msg;
        err_line = n.line;
        return false;
    }

    // Parser::parse_matrix: up to 16 numbers, row-major, identity for the rest
    static void matrix(const std::string& leaf, float m[16]) {
        const float id[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        std::memcpy(m, id, sizeof id);
        const std::vector<float> v = floats(leaf);
        for (size_t i = 0; i < v.size() && i < 16; ++i) m[i] = v[i];
    }
    static void color3(const std::string& leaf, float c[3]) {
        const std::vector<float> v = floats(leaf);
        for (size_t i = 0; i < v.size() && i < 3; ++i) c[i] = v[i];
    }

    // Assembly::get_surface_shader: this assembly, then its parents
    int find_shader(int a, const std::string& name) const {
        for (int x = a; x >= 0; x = ps.asms[x].parent) {
            auto it = ps.asms[x].shader_map.find(name);
            if (it != ps.asms[x].shader_map.end()) return it->second;
        }
        return -1;
    }

    // Parser::parse_surface_shader
    bool surface_shader(int a, const DtNode& n) {
        const DtNode* type = nullptr;
        for (const DtNode& c : n.children)
            if (c.type == "Type") { type = &c; break; }
        int index = -1;
        if (!type) {
            warn(n, "surface shader " + n.name + " without a type");
        } else if (type->leaf == "Emit" || type->leaf == "Lambert") {
            psy_shader_desc d{};
            d.kind = type->leaf == "Emit" ? PSY_SHADER_EMIT : PSY_SHADER_LAMBERT;
            d.color[0] = d.color[1] = d.color[2] = 0.9f;
            for (const DtNode& c : n.children)
                if (c.type == "Color") color3(c.leaf, d.color);
            index = (int)ps.shaders.size();
            ps.shaders.push_back(d);
        } else if (type->leaf == "GTR") {
            // the reference's defaults and fields (parser/parser.cpp:511-546):
            // the first number of Roughness, TailShape and Fresnel
            psy_shader_desc d{};
            d.kind = PSY_SHADER_GTR;
            d.color[0] = d.color[1] = d.color[2] = 0.9f;
            d.roughness = 0.1f;
            d.tail_shape = 2.0f;
            d.fresnel = 0.25f;
            for (const DtNode& c : n.children) {
                if (c.type == "Color") {
                    color3(c.leaf, d.color);
                } else if (c.type == "Roughness" || c.type == "TailShape" || c.type == "Fresnel") {
                    const std::vector<float> v = floats(c.leaf);
                    if (v.empty()) continue;
                    (c.type == "Roughness" ? d.roughness : c.type == "TailShape" ? d.tail_shape : d.fresnel) = v[0];
                }
            }
            index = (int)ps.shaders.size();
            ps.shaders.push_back(d);
        } else {
            warn(n, "unknown surface shader type '" + type->leaf + "'");
        }
        ps.asms[a].shader_map.emplace(n.name, index);
        return true;
    }

    int add_object(int a, const DtNode& n, int kind, int ts, std::vector<float>&& data) {
        PS::Asm& A = ps.asms[a];
        const int idx = (int)A.objects.size();
        A.object_data.push_back(std::move(data));
        A.object_counts.emplace_back();
        A.object_indices.emplace_back();
        psy_object_desc o{};
        o.kind = kind;
        o.time_samples = ts;
        A.objects.push_back(o);
        A.object_map.emplace(n.name, idx);
        return idx;
    }

    // Parser::parse_bilinear_patch / parse_bicubic_patch: one time sample
    // per Vertices leaf
    bool patch(int a, const DtNode& n, int nf, int kind) {
        std::vector<float> data;
        int ts = 0;
        for (const DtNode& c : n.children) {
            if (c.type != "Vertices") continue;
            const std::vector<float> v = floats(c.leaf);
            if ((int)v.size() < nf)
                return fail(c, n.type + " " + n.name + ": Vertices holds " + std::to_string(v.size()) + " numbers, needs " +
                                   std::to_string(nf));
            data.insert(data.end(), v.begin(), v.begin() + nf);
            ++ts;
        }
        if (ts == 0) return fail(n, n.type + " " + n.name + " has no Vertices");
        add_object(a, n, kind, ts, std::move(data));
        return true;
    }

    // Parser::parse_subdivision_surface (parser/parser.cpp:416-468): each
    // Vertices leaf adds its whole xyz triples, the first one fixing the
    // points per time sample and the total cut to a multiple of it; the
    // last FaceVertCounts and FaceVertIndices leaves win
    bool subdivision(int a, const DtNode& n) {
        std::vector<float> verts;
        int per = 0;
        std::vector<int> counts, indices;
        for (const DtNode& c : n.children) {
            if (c.type == "Vertices") {
                const std::vector<float> v = floats(c.leaf);
                const size_t whole = v.size() / 3;
                verts.insert(verts.end(), v.begin(), v.begin() + 3 * whole);
                if (per == 0) per = (int)whole;
                if (per > 0) verts.resize(verts.size() - verts.size() % (3 * (size_t)per));
            } else if (c.type == "FaceVertCounts") {
                if (!ints(c.leaf, counts)) return fail(c, "FaceVertCounts: integer out of range");
            } else if (c.type == "FaceVertIndices") {
                if (!ints(c.leaf, indices)) return fail(c, "FaceVertIndices: integer out of range");
            }
        }
        int64_t total = 0;
        for (int k : counts) total += k;
        if (total != (int64_t)indices.size())
            return fail(n, "SubdivisionSurface " + n.name + ": FaceVertIndices holds " + std::to_string(indices.size()) +
                               " entries, FaceVertCounts sum to " + std::to_string(total));
        const int ts = per > 0 ? (int)(verts.size() / (3 * (size_t)per)) : 1;
        const int idx = add_object(a, n, PSY_OBJ_SUBDIVISION, ts, std::move(verts));
        PS::Asm& A = ps.asms[a];
        A.object_counts[idx].assign(counts.begin(), counts.end());
        A.object_indices[idx].assign(indices.begin(), indices.end());
        A.objects[idx].n_vertices = per;
        A.objects[idx].n_faces = (int32_t)counts.size();
        return true;
    }

    // Parser::parse_sphere: one time sample, defaults (0,0,0) and 0.5
    bool sphere(int a, const DtNode& n) {
        std::vector<float> d = {0.0f, 0.0f, 0.0f, 0.5f};
        for (const DtNode& c : n.children) {
            const std::vector<float> v = floats(c.leaf);
            if (c.type == "Radius" && !v.empty()) d[3] = v[0];
            else if (c.type == "Location")
                for (size_t i = 0; i < v.size() && i < 3; ++i) d[i] = v[i];
        }
        add_object(a, n, PSY_OBJ_SPHERE, 1, std::move(d));
        return true;
    }

    // Parser::parse_sphere_light
    bool sphere_light(int a, const DtNode& n) {
        std::vector<std::array<float, 3>> loc, col;
        std::vector<float> rad;
        for (const DtNode& c : n.children) {
            const std::vector<float> v = floats(c.leaf);
            if (c.type == "Color") {
                std::array<float, 3> x = {0.0f, 0.0f, 0.0f};
                for (size_t i = 0; i < v.size() && i < 3; ++i) x[i] = v[i];
                col.push_back(x);
            } else if (c.type == "Radius") {
                if (!v.empty()) rad.push_back(v[0]);
            } else if (c.type == "Location") {
                std::array<float, 3> x = {0.0f, 0.0f, 0.0f};
                for (size_t i = 0; i < v.size() && i < 3; ++i) x[i] = v[i];
                loc.push_back(x);
            }
        }
        if (loc.empty() || rad.empty() || col.empty())
            return fail(n, "SphereLight " + n.name + " needs Location, Radius and Color");
        const size_t ts = std::max(loc.size(), std::max(rad.size(), col.size()));
        if ((loc.size() != 1 && loc.size() != ts) || (rad.size() != 1 && rad.size() != ts) ||
            (col.size() != 1 && col.size() != ts))
            return fail(n, "SphereLight " + n.name +
                               ": Location, Radius and Color have different motion sample counts (only equal counts, "
                               "or one sample for a constant, are supported)");
        std::vector<float> d;
        for (size_t t = 0; t < ts; ++t) {
            const auto& L = loc[loc.size() == 1 ? 0 : t];
            const auto& C = col[col.size() == 1 ? 0 : t];
            d.insert(d.end(), {L[0], L[1], L[2], rad[rad.size() == 1 ? 0 : t], C[0], C[1], C[2]});
        }
        add_object(a, n, PSY_OBJ_SPHERE_LIGHT, (int)ts, std::move(d));
        return true;
    }

    // Parser::parse_rectangle_light: Dimensions default to 1 x 1 per leaf
    bool rect_light(int a, const DtNode& n) {
        std::vector<std::array<float, 2>> dim;
        std::vector<std::array<float, 3>> col;
        for (const DtNode& c : n.children) {
            const std::vector<float> v = floats(c.leaf);
            if (c.type == "Color") {
                std::array<float, 3> x = {0.0f, 0.0f, 0.0f};
                for (size_t i = 0; i < v.size() && i < 3; ++i) x[i] = v[i];
                col.push_back(x);
            } else if (c.type == "Dimensions") {
                dim.push_back({v.size() > 0 ? v[0] : 1.0f, v.size() > 1 ? v[1] : 1.0f});
            }
        }
        if (dim.empty() || col.empty()) return fail(n, "RectangleLight " + n.name + " needs Dimensions and Color");
        const size_t ts = std::max(dim.size(), col.size());
        if ((dim.size() != 1 && dim.size() != ts) || (col.size() != 1 && col.size() != ts))
            return fail(n, "RectangleLight " + n.name +
                               ": Dimensions and Color have different motion sample counts (only equal counts, or "
                               "one sample for a constant, are supported)");
        std::vector<float> d;
        for (size_t t = 0; t < ts; ++t) {
            const auto& D = dim[dim.size() == 1 ? 0 : t];
            const auto& C = col[col.size() == 1 ? 0 : t];
            d.insert(d.end(), {D[0], D[1], C[0], C[1], C[2]});
        }
        add_object(a, n, PSY_OBJ_RECTANGLE_LIGHT, (int)ts, std::move(d));
        return true;
    }

    // Parser::parse_assembly; returns the new assembly's scene index or -1
    int assembly(const DtNode& n, int parent) {
        const int a = (int)ps.asms.size();
        ps.asms.emplace_back();
        ps.asms[a].parent = parent;
        for (const DtNode& c : n.children) {
            bool ok = true;
            if (c.type == "Assembly") {
                const int sub = assembly(c, a);
                if (sub < 0) return -1;
                PS::Asm& A = ps.asms[a];
                A.assembly_map.emplace(c.name, (int)A.children.size());
                A.children.push_back(sub);
            } else if (c.type == "BilinearPatch") {
                ok = patch(a, c, 12, PSY_OBJ_BILINEAR);
            } else if (c.type == "BicubicPatch") {
                ok = patch(a, c, 48, PSY_OBJ_BICUBIC);
            } else if (c.type == "SubdivisionSurface") {
                ok = subdivision(a, c);
            } else if (c.type == "Sphere") {
                ok = sphere(a, c);
            } else if (c.type == "SurfaceShader") {
                ok = surface_shader(a, c);
            } else if (c.type == "SphereLight") {
                ok = sphere_light(a, c);
            } else if (c.type == "RectangleLight") {
                ok = rect_light(a, c);
            } else if (c.type == "Instance") {
                ok = instance(a, c);
            }
            if (!ok) return -1;
        }
        return a;
    }

    bool instance(int a, const DtNode& n) {
        std::string name;
        std::vector<float> xf;
        int shader = -1;
        for (const DtNode& c : n.children) {
            if (c.type == "Transform") {
                float m[16];
                matrix(c.leaf, m);
                xf.insert(xf.end(), m, m + 16);
            } else if (c.type == "Data") {
                name = c.leaf;
            } else if (c.type == "SurfaceShaderBind") {
                shader = find_shader(a, c.leaf);
                if (shader == -1) warn(c, "bound surface shader " + c.leaf + " does not exist");
            }
        }
        PS::Asm& A = ps.asms[a];
        psy_instance_desc d{};
        d.transform_index = (int32_t)(A.transforms.size() / 16);
        d.transform_count = (int32_t)(xf.size() / 16);
        d.shader_index = shader;
        auto o = A.object_map.find(name);
        auto s = A.assembly_map.find(name);
        if (o != A.object_map.end()) {
            d.type = PSY_INSTANCE_OBJECT;
            d.data_index = o->second;
        } else if (s != A.assembly_map.end()) {
            d.type = PSY_INSTANCE_ASSEMBLY;
            d.data_index = s->second;
        } else {
            warn(n, "instance of " + name + ", which does not exist in this assembly; skipped");
            return true;
        }
        A.transforms.insert(A.transforms.end(), xf.begin(), xf.end());
        A.instances.push_back(d);
        return true;
    }

    // Parser::parse_camera: Fov in degrees
    void camera(const DtNode& n) {
        for (const DtNode& c : n.children) {
            const std::vector<float> v = floats(c.leaf);
            if (c.type == "Fov" && !v.empty()) ps.cam_fov.push_back((3.1415926536f / 180.0f) * v[0]);
            else if (c.type == "FocalDistance" && !v.empty()) ps.cam_fd.push_back(v[0]);
            else if (c.type == "ApertureRadius" && !v.empty()) ps.cam_ap.push_back(v[0]);
            else if (c.type == "Transform") {
                float m[16];
                matrix(c.leaf, m);
                ps.cam_xf.insert(ps.cam_xf.end(), m, m + 16);
            }
        }
    }

    // Parser::parse_next_frame for one Scene node
    bool scene(const DtNode& n) {
        psy_render_settings& S = ps.settings;
        S.width = S.height = S.spp = S.seed = 1;
        S.dice_rate = 0.25f;
        std::snprintf(S.name, sizeof S.name, "%s", n.name.empty() ? "" : n.name.c_str() + 1);
        float bg[3] = {0.0f, 0.0f, 0.0f};
        bool have_root = false;
        std::vector<int> iv;
        for (const DtNode& c : n.children) {
            if (c.type == "Output") {
                for (const DtNode& o : c.children) {
                    if (o.type != "Path") continue;
                    // the text between the first and the last double quote
                    const size_t q0 = o.leaf.find('"'), q1 = o.leaf.rfind('"');
                    if (q0 != std::string::npos && q1 > q0)
                        std::snprintf(S.output_path, sizeof S.output_path, "%s", o.leaf.substr(q0 + 1, q1 - q0 - 1).c_str());
                }
            } else if (c.type == "RenderSettings") {
                for (const DtNode& r : c.children) {
                    if (r.type == "Resolution" || r.type == "SamplesPerPixel" || r.type == "Seed") {
                        if (!ints(r.leaf, iv)) return fail(r, r.type + ": number out of range");
                        if (r.type == "Resolution") {
                            if (iv.size() > 0) S.width = iv[0];
                            if (iv.size() > 1) S.height = iv[1];
                        } else if (!iv.empty()) {
                            (r.type == "Seed" ? S.seed : S.spp) = iv[0];
                        }
                    } else if (r.type == "DicingRate") {
                        const std::vector<float> v = floats(r.leaf);
                        if (!v.empty()) S.dice_rate = v[0];
                    }
                }
            } else if (c.type == "Camera") {
                camera(c);
            } else if (c.type == "World") {
                for (const DtNode& w : c.children) {
                    if (w.type != "BackgroundShader") continue;
                    const DtNode* type = nullptr;
                    for (const DtNode& x : w.children)
                        if (x.type == "Type") { type = &x; break; }
                    if (!type) return fail(w, "background shader without a type");
                    if (type->leaf == "Color") {
                        float col[3] = {0.0f, 0.0f, 0.0f};
                        for (const DtNode& x : w.children)
                            if (x.type == "Color") color3(x.leaf, col);
                        std::memcpy(bg, col, sizeof bg);
                    }
                }
            } else if (c.type == "Assembly") {
                // a later root assembly replaces an earlier one
                ps.asms.clear();
                ps.shaders.clear();
                if (assembly(c, -1) != 0) return false;
                have_root = true;
            }
        }
        if (!have_root) {
            ps.asms.clear();
            ps.asms.emplace_back();
        }
        for (int k = 0; k < 3; ++k) ps.desc.background[k] = bg[k];
        ps.desc.dice_rate = S.dice_rate;
        ps.desc.min_upoly_size = 1e-5f;
        return true;
    }
};

void finish_desc(PS& ps) {
    ps.asm_desc.resize(ps.asms.size());
    for (size_t a = 0; a < ps.asms.size(); ++a) {
        PS::Asm& A = ps.asms[a];
        for (size_t i = 0; i < A.objects.size(); ++i) {
            A.objects[i].data = A.object_data[i].data();
            A.objects[i].face_vert_counts = A.object_counts[i].data();
            A.objects[i].face_vert_indices = A.object_indices[i].data();
        }
        psy_assembly_desc& d = ps.asm_desc[a];
        d.objects = A.objects.data();
        d.n_objects = (int32_t)A.objects.size();
        d.instances = A.instances.data();
        d.n_instances = (int32_t)A.instances.size();
        d.transforms = A.transforms.data();
        d.n_transforms = (int32_t)(A.transforms.size() / 16);
        d.assemblies = A.children.data();
        d.n_assemblies = (int32_t)A.children.size();
    }
    psy_scene_desc& D = ps.desc;
    D.assemblies = ps.asm_desc.data();
    D.n_assemblies = (int32_t)ps.asm_desc.size();
    D.shaders = ps.shaders.data();
    D.n_shaders = (int32_t)ps.shaders.size();
    psy_camera_desc& C = D.camera;
    C.transforms = ps.cam_xf.data();
    C.n_transforms = (int32_t)(ps.cam_xf.size() / 16);
    C.fovs = ps.cam_fov.data();
    C.n_fovs = (int32_t)ps.cam_fov.size();
    C.aperture_radii = ps.cam_ap.data();
    C.n_aperture_radii = (int32_t)ps.cam_ap.size();
    C.focus_distances = ps.cam_fd.data();
    C.n_focus_distances = (int32_t)ps.cam_fd.size();
}

int tree_or_fail(const char* text, int64_t len, DtNode& root, const char* who) {
    if (!text && len > 0) return psy::fail(PSY_ERR_INVALID, std::string(who) + ": NULL text");
    if (len < 0) return psy::fail(PSY_ERR_INVALID, std::string(who) + ": negative length");
    std::string err;
    int line = 0;
    if (!build_tree(text ? text : "", len, root, err, line))
        return psy::fail(PSY_ERR_INVALID, std::string(who) + ": line " + std::to_string(line) + ": " + err);
    return PSY_OK;
}

}  // namespace

extern "C" {

int psy_parse_scene_count(const char* text, int64_t len, int32_t* n_scenes) {
    if (!n_scenes) return psy::fail(PSY_ERR_INVALID, "psy_parse_scene_count: NULL argument");
    DtNode root;
    if (int rc = tree_or_fail(text, len, root, "psy_parse_scene_count")) return rc;
    int32_t n = 0;
    for (const DtNode& c : root.children)
        if (c.type == "Scene") ++n;
    *n_scenes = n;
    return PSY_OK;
}

int psy_parse_scene(const char* text, int64_t len, int32_t scene_index, psy_parsed_scene** out) {
    if (!out) return psy::fail(PSY_ERR_INVALID, "psy_parse_scene: NULL argument");
    *out = nullptr;
    DtNode root;
    if (int rc = tree_or_fail(text, len, root, "psy_parse_scene")) return rc;
    const DtNode* sc = nullptr;
    int32_t k = 0;
    for (const DtNode& c : root.children)
        if (c.type == "Scene" && k++ == scene_index) { sc = &c; break; }
    if (!sc || scene_index < 0) return psy::fail(PSY_ERR_INVALID, "psy_parse_scene: no Scene node with that index");
    std::unique_ptr<psy_parsed_scene> ps(new psy_parsed_scene());
    Builder b{*ps};
    if (!b.scene(*sc))
        return psy::fail(PSY_ERR_UNSUPPORTED, "psy_parse_scene: line " + std::to_string(b.err_line) + ": " + b.err);
    finish_desc(*ps);
    *out = ps.release();
    return PSY_OK;
}

const psy_scene_desc* psy_parsed_scene_desc(const psy_parsed_scene* p) { return p ? &p->desc : nullptr; }

int psy_parsed_scene_settings(const psy_parsed_scene* p, psy_render_settings* out) {
    if (!p || !out) return psy::fail(PSY_ERR_INVALID, "psy_parsed_scene_settings: NULL argument");
    *out = p->settings;
    return PSY_OK;
}

const char* psy_parsed_scene_warnings(const psy_parsed_scene* p) { return p ? p->warnings.c_str() : ""; }

void psy_parsed_scene_destroy(psy_parsed_scene* p) { delete p; }

int psy_parse_dump_tree(const char* text, int64_t len, char* out, int64_t cap, int64_t* needed) {
    DtNode root;
    if (int rc = tree_or_fail(text, len, root, "psy_parse_dump_tree")) return rc;
    std::string s;
    for (const DtNode& c : root.children) print_tree(c, "", s);
    if (needed) *needed = (int64_t)s.size() + 1;
    if (out && cap > 0) {
        const size_t n = std::min<size_t>(s.size(), (size_t)cap - 1);
        std::memcpy(out, s.data(), n);
        out[n] = '\0';
    }
    return PSY_OK;
}

}  // extern "C"
