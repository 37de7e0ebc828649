// psy_subdiv.cpp -- Catmull-Clark control meshes to bicubic Bezier patches.
// Host code only.
//
// Stands in for the patch extraction of SubdivisionSurface::finalize
// (object/subdivision_surface.cpp:85-233), which hands the control mesh to
// OpenSubdiv: Catmull-Clark (Sdc::SCHEME_CATMARK) with boundary edges
// infinitely sharp and boundary vertices smooth (VTX_BOUNDARY_EDGE_ONLY,
// :93), refined adaptively around irregular features down to level 5
// (maxIsolation, :87), each regular face a bicubic B-spline patch whose
// missing boundary row is mirrored (2*boundary - inner, :167-226) and then
// converted to Bezier form (bspline_to_bezier_patch,
// object/patch_utils.hpp:204-228).
//
// The extraction here is feature-adaptive in the same sense:
//   - level 0 is the control mesh and every face of it is "active"; at each
//     level an active face is regular when it is a quad whose four corners
//     are regular vertices (interior with four quads around it, or on the
//     boundary with two) and it has no two opposite boundary edges; a
//     regular active face becomes the B-spline patch of the 4x4 grid around
//     it, an irregular one is split by a Catmull-Clark step and its children
//     are active at the next level;
//   - active faces still irregular at level 5 become end caps: Bezier
//     patches whose corners are the limit positions of the face's corners,
//     whose edge points follow the limit tangents along the face's edges,
//     and whose four inner points are those of the face's B-spline patch
//     (for a face with regular corners this is exactly its B-spline patch,
//     and two faces sharing an edge share its boundary curve);
//   - only the faces within two vertex rings of the irregular active faces
//     are refined at each level; each refined vertex records whether its
//     whole stencil was present ("valid") and whether every face around it
//     was refined ("complete"), and patches are built from valid positions
//     around complete corners only.
//
// Positions are refined in double precision and each Bezier control point
// is rounded to float once.  The orders of every sum are fixed (faces in
// level order, a face's vertices in its own order, a vertex's fan in
// rotation order from its first half-edge), so the oracle's restatement
// gives the same floats.
//
// Refinement rules (Catmull-Clark, boundary edges as creases):
//   face point     the average of the face's vertices;
//   edge point     (v0 + v1 + F0 + F1) / 4 inside, (v0 + v1) / 2 on the boundary;
//   vertex point   (n-2)/n V + (sum E + sum F) / n^2 inside (E: the n neighbours,
//                  F: the face points around), (E_a + 6 V + E_b) / 8 on the
//                  boundary (E_a, E_b: the two boundary neighbours).
// Limit masks (Halstead, Kass and DeRose 1993) for the end caps:
//   position       (n^2 V + 4 sum E + sum D) / (n (n+5)) inside (D: the
//                  diagonal vertex of each quad around), (E_a + 4 V + E_b) / 6
//                  on the boundary;
//   tangent        sum_k A_n cos(2 pi k/n) E_k + (cos(2 pi k/n) + cos(2 pi (k+1)/n)) D_k,
//                  A_n = 1 + cos(2 pi/n) + cos(pi/n) sqrt(2 (9 + cos(2 pi/n))), over 3n
//                  (the B-spline derivative per unit parameter at n = 4); on the
//                  boundary (E_edge - E_other) / 2 along a boundary edge, and across
//                  (D_a - E_a + 4 (W - V) + D_b - E_b) / 6 along an interior edge to W
//                  (E_a, E_b its neighbours in the fan, D_a, D_b the diagonals
//                  between), exact with two faces.
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "psy_internal.hpp"

namespace psy {
namespace {

constexpr int MAX_LEVEL = 5;

struct D3 {
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
inline D3 operator/(D3 a, double s) { return D3{a.x / s, a.y / s, a.z / s}; }

struct Level {
    std::vector<D3> p;
    std::vector<uint8_t> valid, complete;
    std::vector<int32_t> off, fv;  // face f: vertices fv[off[f] .. off[f+1])
    std::vector<uint8_t> active;
    // topology: per half-edge (index into fv) its face and twin (-1: none
    // present); per vertex the first outgoing half-edge of its fan (on the
    // boundary the one whose edge has no twin), the fan's face count, and
    // whether the fan ends at a boundary
    std::vector<int32_t> hface, twin, vfirst, vfan;
    std::vector<uint8_t> vbound, vquads;

    int nf() const { return (int)off.size() - 1; }
    int next(int h) const {
        const int f = hface[h];
        return h + 1 < off[f + 1] ? h + 1 : off[f];
    }
    int prev(int h) const {
        const int f = hface[h];
        return h > off[f] ? h - 1 : off[f + 1] - 1;
    }
    int org(int h) const { return fv[h]; }
    int dst(int h) const { return fv[next(h)]; }
    int size(int f) const { return off[f + 1] - off[f]; }
    // the next outgoing half-edge of org(h)'s fan, -1 at a boundary
    int rot(int h) const { return twin[prev(h)]; }
};

bool build_topology(Level& L, bool full, std::string& err) {
    const int nf = L.nf();
    const size_t nh = L.fv.size();
    const size_t nv = L.p.size();
    L.hface.resize(nh);
    for (int f = 0; f < nf; ++f)
        for (int h = L.off[f]; h < L.off[f + 1]; ++h) L.hface[h] = f;
    // outgoing half-edges of every vertex, in half-edge order (CSR): a
    // twin is found among the few half-edges leaving its edge's far end
    std::vector<int32_t> outc(nv, 0), lowest(nv, -1), out_off(nv + 1, 0), out_h(nh);
    for (size_t h = 0; h < nh; ++h) {
        const int v = L.org((int)h);
        if (lowest[v] < 0) lowest[v] = (int32_t)h;
        ++outc[v];
    }
    for (size_t v = 0; v < nv; ++v) out_off[v + 1] = out_off[v] + outc[v];
    {
        std::vector<int32_t> fill(out_off.begin(), out_off.end() - 1);
        for (size_t h = 0; h < nh; ++h) out_h[fill[L.org((int)h)]++] = (int32_t)h;
    }
    L.twin.assign(nh, -1);
    for (size_t h = 0; h < nh; ++h) {
        const int a = L.org((int)h), b = L.dst((int)h);
        for (int k = out_off[a]; k < out_off[a + 1]; ++k) {
            const int g = out_h[k];
            if (g != (int)h && L.dst(g) == b) {
                err = "an edge is used twice in the same direction (a non-manifold edge or inconsistently oriented faces)";
                return false;
            }
        }
        for (int k = out_off[b]; k < out_off[b + 1]; ++k)
            if (L.dst(out_h[k]) == a) {
                L.twin[h] = out_h[k];
                break;
            }
    }
    L.vfirst.assign(nv, -1);
    L.vfan.assign(nv, 0);
    L.vbound.assign(nv, 0);
    L.vquads.assign(nv, 0);
    for (size_t v = 0; v < nv; ++v) {
        if (lowest[v] < 0) continue;
        // back up to the start of the fan holding the lowest half-edge
        int h = lowest[v];
        bool bnd = false;
        for (int steps = 0;; ++steps) {
            if (steps > outc[v]) { err = "broken vertex fan"; return false; }
            const int t = L.twin[h];
            if (t < 0) { bnd = true; break; }
            const int g = L.next(t);
            if (g == lowest[v]) break;
            h = g;
        }
        const int first = bnd ? h : lowest[v];
        int count = 1;
        bool quads = L.size(L.hface[first]) == 4;
        for (int g = first;;) {
            const int t = L.rot(g);
            if (t < 0 || t == first) break;
            g = t;
            ++count;
            quads = quads && L.size(L.hface[g]) == 4;
            if (count > outc[v]) { err = "broken vertex fan"; return false; }
        }
        if (full && count != outc[v]) {
            err = "vertex " + std::to_string(v) + " is non-manifold (its faces form more than one fan)";
            return false;
        }
        L.vfirst[v] = first;
        L.vfan[v] = count;
        L.vbound[v] = bnd ? 1 : 0;
        L.vquads[v] = quads ? 1 : 0;
    }
    return true;
}

bool regular_vertex(const Level& L, int v) {
    if (!L.complete[v] || !L.vquads[v]) return false;
    return L.vbound[v] ? L.vfan[v] == 2 : L.vfan[v] == 4;
}

bool regular_face(const Level& L, int f) {
    if (L.size(f) != 4) return false;
    const int h0 = L.off[f];
    for (int i = 0; i < 4; ++i)
        if (!regular_vertex(L, L.fv[h0 + i])) return false;
    const bool b[4] = {L.twin[h0] < 0, L.twin[h0 + 1] < 0, L.twin[h0 + 2] < 0, L.twin[h0 + 3] < 0};
    return !((b[0] && b[2]) || (b[1] && b[3]));
}

// Bezier form of the bicubic B-spline patch over a 4x4 grid (row-major,
// rows along v): B = M G M^T with the rows of M (1,4,1,0)/6, (0,4,2,0)/6,
// (0,2,4,0)/6, (0,1,4,1)/6.
void bspline_to_bezier(const D3 g[16], D3 b[16]) {
    static const double M[4][4] = {{1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0, 0.0},
                                   {0.0, 4.0 / 6.0, 2.0 / 6.0, 0.0},
                                   {0.0, 2.0 / 6.0, 4.0 / 6.0, 0.0},
                                   {0.0, 1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0}};
    D3 t[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            D3 s{0.0, 0.0, 0.0};
            for (int k = 0; k < 4; ++k) s = s + g[r * 4 + k] * M[c][k];
            t[r * 4 + c] = s;
        }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            D3 s{0.0, 0.0, 0.0};
            for (int k = 0; k < 4; ++k) s = s + t[k * 4 + c] * M[r][k];
            b[r * 4 + c] = s;
        }
}

void emit(std::vector<float>& out, const D3 b[16]) {
    for (int i = 0; i < 16; ++i) {
        out.push_back((float)b[i].x);
        out.push_back((float)b[i].y);
        out.push_back((float)b[i].z);
    }
}

// The B-spline patch of a regular face.  Grid cells (row, col): the face's
// corners c0..c3 sit at (1,1), (1,2), (2,2), (2,1); the face across side i
// (from c_i to c_{i+1}) gives two grid points and the face diagonal to c_i
// one more; a boundary side's row or column is mirrored.
bool regular_patch(const Level& L, int f, std::vector<float>& out, std::string& err) {
    static const int corner_cell[4] = {5, 6, 10, 9};
    // per side: cells of the point next to c_i, next to c_{i+1}, and diagonal to c_i
    static const int side_cell[4][3] = {{1, 2, 0}, {7, 11, 3}, {14, 13, 15}, {8, 4, 12}};
    int idx[16];
    for (int i = 0; i < 16; ++i) idx[i] = -1;
    const int h0 = L.off[f];
    for (int i = 0; i < 4; ++i) idx[corner_cell[i]] = L.fv[h0 + i];
    int boundary_side = -1;
    for (int i = 0; i < 4; ++i) {
        const int t = L.twin[h0 + i];
        if (t < 0) {
            boundary_side = i;
            continue;
        }
        const int n1 = L.next(t);
        idx[side_cell[i][0]] = L.dst(n1);
        idx[side_cell[i][1]] = L.dst(L.next(n1));
        const int u = L.twin[n1];
        if (u >= 0) idx[side_cell[i][2]] = L.dst(L.next(L.next(u)));
    }
    D3 g[16];
    for (int i = 0; i < 16; ++i)
        if (idx[i] >= 0) {
            if (!L.valid[idx[i]]) { err = "internal: a patch point outside the refined region"; return false; }
            g[i] = L.p[idx[i]];
        }
    // mirrored boundary row or column (object/subdivision_surface.cpp:167-196)
    auto mirror = [&](int dst, int edge, int inner) {
        if (idx[dst] >= 0) return;
        g[dst] = g[edge] * 2.0 - g[inner];
        idx[dst] = -2;
    };
    switch (boundary_side) {
        case 0: for (int c = 0; c < 4; ++c) mirror(c, 4 + c, 8 + c); break;
        case 1: for (int r = 0; r < 4; ++r) mirror(4 * r + 3, 4 * r + 2, 4 * r + 1); break;
        case 2: for (int c = 0; c < 4; ++c) mirror(12 + c, 8 + c, 4 + c); break;
        case 3: for (int r = 0; r < 4; ++r) mirror(4 * r, 4 * r + 1, 4 * r + 2); break;
        default: break;
    }
    for (int i = 0; i < 16; ++i)
        if (idx[i] == -1) { err = "internal: incomplete patch grid"; return false; }
    D3 b[16];
    bspline_to_bezier(g, b);
    emit(out, b);
    return true;
}

// Limit position of a complete vertex.
D3 limit_position(const Level& L, int v) {
    const D3 V = L.p[v];
    if (L.vbound[v]) {
        const int first = L.vfirst[v];
        int last = first;
        for (int t = L.rot(last); t >= 0 && t != first; t = L.rot(last)) last = t;
        const D3 ea = L.p[L.dst(first)];
        const D3 eb = L.p[L.org(L.prev(last))];
        return ((ea + V * 4.0) + eb) / 6.0;
    }
    const int n = L.vfan[v];
    D3 se{0.0, 0.0, 0.0}, sd{0.0, 0.0, 0.0};
    int g = L.vfirst[v];
    for (int k = 0; k < n; ++k) {
        se = se + L.p[L.dst(g)];
        sd = sd + L.p[L.dst(L.next(g))];
        g = L.rot(g);
    }
    return ((V * (double)(n * n) + se * 4.0) + sd) / (double)(n * (n + 5));
}

// Limit tangent of a complete vertex v along its edge to w, per unit
// parameter of a face at this level.
bool limit_tangent(const Level& L, int v, int w, D3& out, std::string& err) {
    const int first = L.vfirst[v];
    if (L.vbound[v]) {
        // the fan's outgoing half-edges g_0 .. g_{k-1}; neighbours
        // e_0 .. e_k with e_0 = dst(g_0), e_k the origin of the last
        // face's incoming boundary edge
        std::vector<int> gs;
        gs.push_back(first);
        for (int t = L.rot(first); t >= 0 && t != first; t = L.rot(gs.back())) gs.push_back(t);
        const int k = (int)gs.size();
        std::vector<int> e(k + 1);
        for (int j = 0; j < k; ++j) e[j] = L.dst(gs[j]);
        e[k] = L.org(L.prev(gs[k - 1]));
        if (w == e[0]) { out = (L.p[e[0]] - L.p[e[k]]) * 0.5; return true; }
        if (w == e[k]) { out = (L.p[e[k]] - L.p[e[0]]) * 0.5; return true; }
        for (int j = 1; j < k; ++j)
            if (e[j] == w) {
                const D3 da = L.p[L.dst(L.next(gs[j - 1]))];
                const D3 db = L.p[L.dst(L.next(gs[j]))];
                out = (((da - L.p[e[j - 1]]) + (L.p[w] - L.p[v]) * 4.0) + (db - L.p[e[j + 1]])) / 6.0;
                return true;
            }
        err = "internal: end-cap edge not in the vertex fan";
        return false;
    }
    const int n = L.vfan[v];
    int h = first;
    int k0 = -1;
    for (int k = 0; k < n; ++k) {
        if (L.dst(h) == w) { k0 = k; break; }
        h = L.rot(h);
    }
    if (k0 < 0) { err = "internal: end-cap edge not in the vertex fan"; return false; }
    const double pi = 3.14159265358979323846;
    const double c1 = std::cos(2.0 * pi / n);
    const double an = 1.0 + c1 + std::cos(pi / n) * std::sqrt(2.0 * (9.0 + c1));
    D3 s{0.0, 0.0, 0.0};
    for (int k = 0; k < n; ++k) {
        const double ck = std::cos(2.0 * pi * k / n), ck1 = std::cos(2.0 * pi * (k + 1) / n);
        s = s + L.p[L.dst(h)] * (an * ck) + L.p[L.dst(L.next(h))] * (ck + ck1);
        h = L.rot(h);
    }
    out = s / (3.0 * n);
    return true;
}

bool end_cap(const Level& L, int f, std::vector<float>& out, std::string& err) {
    const int h0 = L.off[f];
    int c[4];
    for (int i = 0; i < 4; ++i) {
        c[i] = L.fv[h0 + i];
        if (!L.complete[c[i]] || !L.valid[c[i]]) { err = "internal: end-cap corner outside the refined region"; return false; }
    }
    // every position a corner's masks read must be known
    for (int i = 0; i < 4; ++i) {
        const int v = c[i];
        const int n = L.vfan[v];
        int g = L.vfirst[v];
        for (int k = 0; k < n; ++k) {
            if (L.size(L.hface[g]) != 4) { err = "internal: end cap next to a non-quad"; return false; }
            const int a = L.dst(g), d = L.dst(L.next(g)), b = L.org(L.prev(g));
            if (!L.valid[a] || !L.valid[d] || !L.valid[b]) { err = "internal: end-cap stencil outside the refined region"; return false; }
            g = L.rot(g);
            if (g < 0) break;
        }
    }
    D3 lim[4];
    for (int i = 0; i < 4; ++i) lim[i] = limit_position(L, c[i]);
    D3 b[16];
    b[0] = lim[0];
    b[3] = lim[1];
    b[15] = lim[2];
    b[12] = lim[3];
    // edge points: (cell, corner, towards)
    static const int edge_pt[8][3] = {{1, 0, 1}, {4, 0, 3}, {2, 1, 0}, {7, 1, 2}, {11, 2, 1}, {14, 2, 3}, {13, 3, 2}, {8, 3, 0}};
    for (int e = 0; e < 8; ++e) {
        D3 t;
        if (!limit_tangent(L, c[edge_pt[e][1]], c[edge_pt[e][2]], t, err)) return false;
        b[edge_pt[e][0]] = lim[edge_pt[e][1]] + t / 3.0;
    }
    const D3 p0 = L.p[c[0]], p1 = L.p[c[1]], p2 = L.p[c[2]], p3 = L.p[c[3]];
    b[5] = (((p0 * 4.0 + p1 * 2.0) + p3 * 2.0) + p2) / 9.0;
    b[6] = (((p1 * 4.0 + p0 * 2.0) + p2 * 2.0) + p3) / 9.0;
    b[10] = (((p2 * 4.0 + p1 * 2.0) + p3 * 2.0) + p0) / 9.0;
    b[9] = (((p3 * 4.0 + p0 * 2.0) + p2 * 2.0) + p1) / 9.0;
    emit(out, b);
    return true;
}

// Faces within two vertex rings of the marked faces.
std::vector<uint8_t> support(const Level& L, const std::vector<uint8_t>& mark) {
    const int nf = L.nf();
    const size_t nv = L.p.size();
    std::vector<int32_t> vf_off(nv + 1, 0), vf;
    for (int32_t v : L.fv) ++vf_off[v + 1];
    for (size_t v = 0; v < nv; ++v) vf_off[v + 1] += vf_off[v];
    vf.resize(L.fv.size());
    std::vector<int32_t> fill(vf_off.begin(), vf_off.end() - 1);
    for (int f = 0; f < nf; ++f)
        for (int h = L.off[f]; h < L.off[f + 1]; ++h) vf[fill[L.fv[h]]++] = f;
    std::vector<uint8_t> cur = mark;
    for (int ring = 0; ring < 2; ++ring) {
        std::vector<uint8_t> vm(nv, 0);
        for (int f = 0; f < nf; ++f)
            if (cur[f])
                for (int h = L.off[f]; h < L.off[f + 1]; ++h) vm[L.fv[h]] = 1;
        std::vector<uint8_t> nx(nf, 0);
        for (size_t v = 0; v < nv; ++v)
            if (vm[v])
                for (int k = vf_off[v]; k < vf_off[v + 1]; ++k) nx[vf[k]] = 1;
        cur.swap(nx);
    }
    return cur;
}

// One Catmull-Clark step over the faces in S.  Child vertices: the face
// points of S's faces in order, then the edge points, then the vertex
// points, each in the order S's faces first reach them; face f's child i is
// (V(v_i), E(v_i v_i+1), F(f), E(v_i-1 v_i)).
Level refine(const Level& L, const std::vector<uint8_t>& S, const std::vector<uint8_t>& split) {
    const int nf = L.nf();
    const size_t nh = L.fv.size();
    const size_t nv = L.p.size();
    std::vector<int32_t> fpi(nf, -1), epi(nh, -1), vpi(nv, -1);
    int nc = 0;
    for (int f = 0; f < nf; ++f)
        if (S[f]) fpi[f] = nc++;
    const int e0 = nc;
    std::vector<int32_t> ep_half;  // the half-edge each edge point was first reached by
    for (int f = 0; f < nf; ++f)
        if (S[f])
            for (int h = L.off[f]; h < L.off[f + 1]; ++h)
                if (epi[h] < 0) {
                    epi[h] = nc;
                    if (L.twin[h] >= 0) epi[L.twin[h]] = nc;
                    ep_half.push_back(h);
                    ++nc;
                }
    const int v0 = nc;
    std::vector<int32_t> vp_vert;
    for (int f = 0; f < nf; ++f)
        if (S[f])
            for (int h = L.off[f]; h < L.off[f + 1]; ++h)
                if (vpi[L.fv[h]] < 0) {
                    vpi[L.fv[h]] = nc++;
                    vp_vert.push_back(L.fv[h]);
                }
    Level C;
    C.p.assign(nc, D3{0.0, 0.0, 0.0});
    C.valid.assign(nc, 0);
    C.complete.assign(nc, 0);
    // face points
    for (int f = 0; f < nf; ++f) {
        if (!S[f]) continue;
        D3 s{0.0, 0.0, 0.0};
        bool ok = true;
        for (int h = L.off[f]; h < L.off[f + 1]; ++h) {
            s = s + L.p[L.fv[h]];
            ok = ok && L.valid[L.fv[h]];
        }
        C.p[fpi[f]] = s / (double)L.size(f);
        C.valid[fpi[f]] = ok;
        C.complete[fpi[f]] = 1;
    }
    // edge points
    for (size_t k = 0; k < ep_half.size(); ++k) {
        const int h = ep_half[k];
        const int id = e0 + (int)k;
        const int a = L.org(h), b = L.dst(h);
        const int t = L.twin[h];
        if (t < 0) {
            // a boundary of the whole mesh only when an end's fan is complete
            if (L.complete[a] || L.complete[b]) {
                C.p[id] = (L.p[a] + L.p[b]) * 0.5;
                C.valid[id] = L.valid[a] && L.valid[b];
                C.complete[id] = 1;
            }
        } else if (S[L.hface[t]]) {
            const int fa = fpi[L.hface[h]], fb = fpi[L.hface[t]];
            C.p[id] = (((L.p[a] + L.p[b]) + C.p[fa]) + C.p[fb]) * 0.25;
            C.valid[id] = L.valid[a] && L.valid[b] && C.valid[fa] && C.valid[fb];
            C.complete[id] = 1;
        }
    }
    // vertex points
    for (size_t k = 0; k < vp_vert.size(); ++k) {
        const int v = vp_vert[k];
        const int id = v0 + (int)k;
        if (!L.complete[v]) continue;
        const int first = L.vfirst[v];
        bool all_in = true;
        int last = first;
        for (int g = first;;) {
            all_in = all_in && S[L.hface[g]];
            last = g;
            const int t = L.rot(g);
            if (t < 0 || t == first) break;
            g = t;
        }
        C.complete[id] = all_in;
        if (L.vbound[v]) {
            const int ea = L.dst(first), eb = L.org(L.prev(last));
            C.p[id] = ((L.p[ea] + L.p[v] * 6.0) + L.p[eb]) * 0.125;
            C.valid[id] = L.valid[v] && L.valid[ea] && L.valid[eb];
        } else if (all_in) {
            const int n = L.vfan[v];
            D3 se{0.0, 0.0, 0.0}, sf{0.0, 0.0, 0.0};
            bool ok = L.valid[v];
            int g = first;
            for (int j = 0; j < n; ++j) {
                se = se + L.p[L.dst(g)];
                sf = sf + C.p[fpi[L.hface[g]]];
                ok = ok && L.valid[L.dst(g)] && C.valid[fpi[L.hface[g]]];
                g = L.rot(g);
            }
            C.p[id] = L.p[v] * ((double)(n - 2) / (double)n) + (se + sf) / (double)(n * n);
            C.valid[id] = ok;
        }
    }
    // child faces
    C.off.push_back(0);
    for (int f = 0; f < nf; ++f) {
        if (!S[f]) continue;
        const int n = L.size(f);
        for (int i = 0; i < n; ++i) {
            const int h = L.off[f] + i;
            const int hp = L.off[f] + (i + n - 1) % n;
            C.fv.push_back(vpi[L.fv[h]]);
            C.fv.push_back(epi[h]);
            C.fv.push_back(fpi[f]);
            C.fv.push_back(epi[hp]);
            C.off.push_back((int32_t)C.fv.size());
            C.active.push_back(split[f]);
        }
    }
    return C;
}

}  // namespace

int subdivide_patches(const float* verts, int32_t n_verts, const int32_t* counts, int32_t n_faces,
                      const int32_t* indices, std::vector<float>& out, std::string& err) {
    out.clear();
    if (n_verts < 0 || n_faces < 0) { err = "negative vertex or face count"; return PSY_ERR_INVALID; }
    if (n_faces == 0) return PSY_OK;
    if (!verts || !counts || !indices) { err = "NULL mesh array"; return PSY_ERR_INVALID; }
    Level L;
    L.p.resize(n_verts);
    for (int i = 0; i < n_verts; ++i) {
        L.p[i] = D3{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
        if (!std::isfinite(verts[3 * i]) || !std::isfinite(verts[3 * i + 1]) || !std::isfinite(verts[3 * i + 2])) {
            err = "vertex " + std::to_string(i) + " is not finite";
            return PSY_ERR_INVALID;
        }
    }
    L.valid.assign(n_verts, 1);
    L.complete.assign(n_verts, 1);
    L.off.push_back(0);
    int64_t total = 0;
    for (int f = 0; f < n_faces; ++f) {
        const int n = counts[f];
        if (n < 3) { err = "face " + std::to_string(f) + " has fewer than 3 vertices"; return PSY_ERR_INVALID; }
        if (total + n > (int64_t)INT32_MAX / 8) { err = "too many face vertices"; return PSY_ERR_UNSUPPORTED; }
        for (int i = 0; i < n; ++i) {
            const int v = indices[total + i];
            if (v < 0 || v >= n_verts) { err = "face " + std::to_string(f) + " has a vertex index out of range"; return PSY_ERR_INVALID; }
            for (int j = 0; j < i; ++j)
                if (indices[total + j] == v) { err = "face " + std::to_string(f) + " repeats a vertex"; return PSY_ERR_INVALID; }
            L.fv.push_back(v);
        }
        total += n;
        L.off.push_back((int32_t)L.fv.size());
    }
    L.active.assign(n_faces, 1);
    for (int level = 0; level <= MAX_LEVEL; ++level) {
        if (!build_topology(L, level == 0, err)) return level == 0 ? PSY_ERR_INVALID : PSY_ERR_UNSUPPORTED;
        const int nf = L.nf();
        std::vector<uint8_t> split(nf, 0);
        bool any = false;
        for (int f = 0; f < nf; ++f) {
            if (!L.active[f]) continue;
            if (regular_face(L, f)) {
                if (!regular_patch(L, f, out, err)) return PSY_ERR_UNSUPPORTED;
            } else if (level == MAX_LEVEL) {
                if (!end_cap(L, f, out, err)) return PSY_ERR_UNSUPPORTED;
            } else {
                split[f] = 1;
                any = true;
            }
        }
        if (!any) break;
        const std::vector<uint8_t> S = support(L, split);
        Level C = refine(L, S, split);
        L = std::move(C);
        if ((int64_t)L.fv.size() > (int64_t)INT32_MAX / 8) { err = "refinement too large"; return PSY_ERR_UNSUPPORTED; }
    }
    return PSY_OK;
}

}  // namespace psy

extern "C" int psy_subdivision_patches(const float* verts, int32_t n_verts, const int32_t* face_vert_counts,
                                       int32_t n_faces, const int32_t* face_vert_indices, float* patches,
                                       int64_t max_patches, int64_t* n_patches) {
    if (!n_patches) return psy::fail(PSY_ERR_INVALID, "psy_subdivision_patches: NULL n_patches");
    std::vector<float> out;
    std::string err;
    const int rc = psy::subdivide_patches(verts, n_verts, face_vert_counts, n_faces, face_vert_indices, out, err);
    if (rc) return psy::fail(rc, "psy_subdivision_patches: " + err);
    const int64_t n = (int64_t)(out.size() / 48);
    *n_patches = n;
    if (patches) {
        if (max_patches < n) return psy::fail(PSY_ERR_INVALID, "psy_subdivision_patches: max_patches below the patch count");
        std::copy(out.begin(), out.end(), patches);
    }
    return PSY_OK;
}
