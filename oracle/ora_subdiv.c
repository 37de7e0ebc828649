/*
 * ora_subdiv.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * SubdivisionSurface (object/subdivision_surface.cpp, .hpp) restated:
 *   - the patch extraction of finalize (:85-233).  The reference leaves it
 *     to OpenSubdiv (Far::TopologyRefiner::RefineAdaptive, maxIsolation 5,
 *     Sdc::SCHEME_CATMARK with VTX_BOUNDARY_EDGE_ONLY; Far::PatchTable with
 *     ENDCAP_BSPLINE_BASIS), a library that is not in the reference tree and
 *     whose version the tree does not pin.  What is restated is the
 *     published Catmull-Clark scheme (Catmull and Clark 1978; boundary
 *     edges as infinitely sharp creases) with feature-adaptive patching
 *     (Niessner, Loop, Meyer and DeRose 2012): regular faces become the
 *     B-spline patch of their 4x4 grid with a missing boundary row mirrored
 *     (:167-196) and converted to Bezier form (object/patch_utils.hpp:
 *     204-228); faces around irregular vertices are refined down to level 5.
 *     OpenSubdiv's B-spline end caps (Gregory-basis conversions) are not
 *     restated: faces still irregular at level 5 get end caps from the
 *     Catmull-Clark limit masks (Halstead, Kass and DeRose 1993), so inside
 *     the level-5 faces around extraordinary vertices and mesh corners
 *     (1/32 of a control face across) parity with the reference is
 *     unpinned; elsewhere the patches are the limit surface, as the
 *     reference's are.
 *   - the patch BVH (build_bvh / build_bvh_recursive, :237-354) and
 *     intersect_rays (:41-83) are in ora_trace.c and ora_scene.c.
 *
 * Extraction steps (the product's psy_subdiv.cpp follows the same steps
 * and the same order of every sum, so the two agree to the bit):
 *   level 0 is the control mesh, every face active; per level, an active
 *   face that is a quad with four regular corners (interior: four quads
 *   around; boundary: two) and no two opposite boundary edges is emitted as
 *   its B-spline patch; at the last level any other active face is emitted
 *   as an end cap; the rest are split.  Each split refines the faces within
 *   two vertex rings of the faces to split (or every face, prune == 0);
 *   child vertices are numbered face points, then edge points, then vertex
 *   points, in the order the refined faces reach them, and child i of face
 *   f is (V(v_i), E(v_i v_i+1), F(f), E(v_i-1 v_i)).  A refined position
 *   is valid when its whole stencil was present, a refined vertex complete
 *   when every face around it was refined.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "psy_oracle.h"

typedef struct { double x, y, z; } d3;
static d3 dadd(d3 a, d3 b) { d3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static d3 dsub(d3 a, d3 b) { d3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static d3 dmul(d3 a, double s) { d3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static d3 ddiv(d3 a, double s) { d3 r = {a.x / s, a.y / s, a.z / s}; return r; }
static const d3 DZERO = {0.0, 0.0, 0.0};

typedef struct {
    int nv, nf, nh;
    d3* p;
    unsigned char *valid, *complete, *active;
    int *off, *fv;
    int *hface, *twin, *vfirst, *vfan;
    unsigned char *vbound, *vquads;
} mesh;

static void mesh_free(mesh* m) {
    free(m->p); free(m->valid); free(m->complete); free(m->active);
    free(m->off); free(m->fv); free(m->hface); free(m->twin); free(m->vfirst); free(m->vfan);
    free(m->vbound); free(m->vquads);
    memset(m, 0, sizeof *m);
}

static int fsize(const mesh* m, int f) { return m->off[f + 1] - m->off[f]; }
static int hnext(const mesh* m, int h) { int f = m->hface[h]; return h + 1 < m->off[f + 1] ? h + 1 : m->off[f]; }
static int hprev(const mesh* m, int h) { int f = m->hface[h]; return h > m->off[f] ? h - 1 : m->off[f + 1] - 1; }
static int horg(const mesh* m, int h) { return m->fv[h]; }
static int hdst(const mesh* m, int h) { return m->fv[hnext(m, h)]; }
/* next outgoing half-edge around org(h) (the face across the previous edge) */
static int hrot(const mesh* m, int h) { return m->twin[hprev(m, h)]; }

typedef struct { uint64_t key; int h; } hkey;
static int hkey_cmp(const void* a, const void* b) {
    const uint64_t x = ((const hkey*)a)->key, y = ((const hkey*)b)->key;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* twins, fans; full: the control mesh, whose vertices must be manifold */
static int topology(mesh* m, int full) {
    const int nh = m->nh, nv = m->nv;
    m->hface = (int*)malloc(sizeof(int) * (nh + 1));
    m->twin = (int*)malloc(sizeof(int) * (nh + 1));
    for (int f = 0; f < m->nf; ++f)
        for (int h = m->off[f]; h < m->off[f + 1]; ++h) m->hface[h] = f;
    hkey* keys = (hkey*)malloc(sizeof(hkey) * (nh + 1));
    for (int h = 0; h < nh; ++h) {
        keys[h].key = ((uint64_t)(uint32_t)horg(m, h) << 32) | (uint32_t)hdst(m, h);
        keys[h].h = h;
    }
    qsort(keys, (size_t)nh, sizeof(hkey), hkey_cmp);
    for (int i = 1; i < nh; ++i)
        if (keys[i].key == keys[i - 1].key) { free(keys); return -1; }
    for (int h = 0; h < nh; ++h) {
        hkey k;
        k.key = ((uint64_t)(uint32_t)hdst(m, h) << 32) | (uint32_t)horg(m, h);
        const hkey* hit = (const hkey*)bsearch(&k, keys, (size_t)nh, sizeof(hkey), hkey_cmp);
        m->twin[h] = hit ? hit->h : -1;
    }
    free(keys);
    int* outc = (int*)calloc((size_t)nv + 1, sizeof(int));
    int* low = (int*)malloc(sizeof(int) * (nv + 1));
    for (int v = 0; v < nv; ++v) low[v] = -1;
    for (int h = 0; h < nh; ++h) {
        const int v = horg(m, h);
        if (low[v] < 0) low[v] = h;
        outc[v]++;
    }
    m->vfirst = (int*)malloc(sizeof(int) * (nv + 1));
    m->vfan = (int*)calloc((size_t)nv + 1, sizeof(int));
    m->vbound = (unsigned char*)calloc((size_t)nv + 1, 1);
    m->vquads = (unsigned char*)calloc((size_t)nv + 1, 1);
    int rc = 0;
    for (int v = 0; v < nv && rc == 0; ++v) {
        m->vfirst[v] = -1;
        if (low[v] < 0) continue;
        int h = low[v], bnd = 0, steps = 0;
        for (;;) {
            if (++steps > outc[v] + 1) { rc = -1; break; }
            const int t = m->twin[h];
            if (t < 0) { bnd = 1; break; }
            const int g = hnext(m, t);
            if (g == low[v]) break;
            h = g;
        }
        if (rc) break;
        const int first = bnd ? h : low[v];
        int count = 1, quads = fsize(m, m->hface[first]) == 4;
        int g = first;
        for (;;) {
            const int t = hrot(m, g);
            if (t < 0 || t == first) break;
            g = t;
            ++count;
            if (fsize(m, m->hface[g]) != 4) quads = 0;
            if (count > outc[v]) { rc = -1; break; }
        }
        if (full && count != outc[v]) rc = -1;
        m->vfirst[v] = first;
        m->vfan[v] = count;
        m->vbound[v] = (unsigned char)bnd;
        m->vquads[v] = (unsigned char)quads;
    }
    free(outc);
    free(low);
    return rc;
}

static int vregular(const mesh* m, int v) {
    if (!m->complete[v] || !m->vquads[v]) return 0;
    return m->vbound[v] ? m->vfan[v] == 2 : m->vfan[v] == 4;
}

static int fregular(const mesh* m, int f) {
    if (fsize(m, f) != 4) return 0;
    const int h = m->off[f];
    for (int i = 0; i < 4; ++i)
        if (!vregular(m, m->fv[h + i])) return 0;
    const int b0 = m->twin[h] < 0, b1 = m->twin[h + 1] < 0, b2 = m->twin[h + 2] < 0, b3 = m->twin[h + 3] < 0;
    return !((b0 && b2) || (b1 && b3));
}

typedef struct { float* p; int64_t n, cap; } patch_out;
static void put(patch_out* o, const d3 b[16]) {
    if (o->n == o->cap) {
        o->cap = o->cap ? 2 * o->cap : 64;
        o->p = (float*)realloc(o->p, sizeof(float) * 48 * (size_t)o->cap);
    }
    float* q = o->p + 48 * o->n;
    for (int i = 0; i < 16; ++i) {
        q[3 * i] = (float)b[i].x;
        q[3 * i + 1] = (float)b[i].y;
        q[3 * i + 2] = (float)b[i].z;
    }
    o->n++;
}

/* bspline_to_bezier_patch (object/patch_utils.hpp:204-228) as a matrix:
 * each B-spline segment P0..P3 has Bezier points (P0 + 4P1 + P2)/6,
 * (4P1 + 2P2)/6, (2P1 + 4P2)/6, (P1 + 4P2 + P3)/6; rows, then columns */
static void to_bezier(const d3 g[16], d3 b[16]) {
    static const double W[4][4] = {{1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0, 0.0},
                                   {0.0, 4.0 / 6.0, 2.0 / 6.0, 0.0},
                                   {0.0, 2.0 / 6.0, 4.0 / 6.0, 0.0},
                                   {0.0, 1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0}};
    d3 t[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            d3 s = DZERO;
            for (int k = 0; k < 4; ++k) s = dadd(s, dmul(g[r * 4 + k], W[c][k]));
            t[r * 4 + c] = s;
        }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            d3 s = DZERO;
            for (int k = 0; k < 4; ++k) s = dadd(s, dmul(t[k * 4 + c], W[r][k]));
            b[r * 4 + c] = s;
        }
}

/* the 4x4 grid of a regular face: corners c0..c3 at cells 5, 6, 10, 9; the
 * face across side i (c_i -> c_i+1) holds the two points next to c_i and
 * c_i+1 and its neighbour around c_i the diagonal one */
static int bspline_patch(const mesh* m, int f, patch_out* o) {
    static const int corner[4] = {5, 6, 10, 9};
    static const int side[4][3] = {{1, 2, 0}, {7, 11, 3}, {14, 13, 15}, {8, 4, 12}};
    int id[16];
    for (int i = 0; i < 16; ++i) id[i] = -1;
    const int h0 = m->off[f];
    int bside = -1;
    for (int i = 0; i < 4; ++i) id[corner[i]] = m->fv[h0 + i];
    for (int i = 0; i < 4; ++i) {
        const int t = m->twin[h0 + i];
        if (t < 0) { bside = i; continue; }
        const int a = hnext(m, t);
        id[side[i][0]] = hdst(m, a);
        id[side[i][1]] = hdst(m, hnext(m, a));
        const int u = m->twin[a];
        if (u >= 0) id[side[i][2]] = hdst(m, hnext(m, hnext(m, u)));
    }
    d3 g[16];
    for (int i = 0; i < 16; ++i) {
        if (id[i] < 0) continue;
        if (!m->valid[id[i]]) return -1;
        g[i] = m->p[id[i]];
    }
    /* the mirrored row or column of a boundary side
     * (object/subdivision_surface.cpp:167-196: 2 * boundary - inner) */
    for (int k = 0; k < 4 && bside >= 0; ++k) {
        int dst, edge, inner;
        switch (bside) {
            case 0: dst = k; edge = 4 + k; inner = 8 + k; break;
            case 1: dst = 4 * k + 3; edge = 4 * k + 2; inner = 4 * k + 1; break;
            case 2: dst = 12 + k; edge = 8 + k; inner = 4 + k; break;
            default: dst = 4 * k; edge = 4 * k + 1; inner = 4 * k + 2; break;
        }
        if (id[dst] >= 0) continue;
        g[dst] = dsub(dmul(g[edge], 2.0), g[inner]);
        id[dst] = -2;
    }
    for (int i = 0; i < 16; ++i)
        if (id[i] == -1) return -1;
    d3 b[16];
    to_bezier(g, b);
    put(o, b);
    return 0;
}

/* the fan's last outgoing half-edge (boundary vertices) */
static int fan_last(const mesh* m, int v) {
    const int first = m->vfirst[v];
    int last = first;
    for (;;) {
        const int t = hrot(m, last);
        if (t < 0 || t == first) break;
        last = t;
    }
    return last;
}

/* limit position: (n^2 V + 4 sum E + sum D) / (n (n+5)); boundary
 * (E_a + 4 V + E_b) / 6 */
static d3 limit_pos(const mesh* m, int v) {
    const d3 V = m->p[v];
    if (m->vbound[v]) {
        const d3 ea = m->p[hdst(m, m->vfirst[v])];
        const d3 eb = m->p[horg(m, hprev(m, fan_last(m, v)))];
        return ddiv(dadd(dadd(ea, dmul(V, 4.0)), eb), 6.0);
    }
    const int n = m->vfan[v];
    d3 se = DZERO, sd = DZERO;
    int g = m->vfirst[v];
    for (int k = 0; k < n; ++k) {
        se = dadd(se, m->p[hdst(m, g)]);
        sd = dadd(sd, m->p[hdst(m, hnext(m, g))]);
        g = hrot(m, g);
    }
    return ddiv(dadd(dadd(dmul(V, (double)(n * n)), dmul(se, 4.0)), sd), (double)(n * (n + 5)));
}

/* limit tangent along the edge v -> w, per unit parameter of a face */
static int limit_tan(const mesh* m, int v, int w, d3* out) {
    if (m->vbound[v]) {
        int gs[64];
        int k = 0;
        const int first = m->vfirst[v];
        gs[k++] = first;
        for (;;) {
            const int t = hrot(m, gs[k - 1]);
            if (t < 0 || t == first) break;
            if (k == 63) return -1;
            gs[k++] = t;
        }
        int e[65];
        for (int j = 0; j < k; ++j) e[j] = hdst(m, gs[j]);
        e[k] = horg(m, hprev(m, gs[k - 1]));
        if (w == e[0]) { *out = dmul(dsub(m->p[e[0]], m->p[e[k]]), 0.5); return 0; }
        if (w == e[k]) { *out = dmul(dsub(m->p[e[k]], m->p[e[0]]), 0.5); return 0; }
        for (int j = 1; j < k; ++j) {
            if (e[j] != w) continue;
            const d3 da = m->p[hdst(m, hnext(m, gs[j - 1]))];
            const d3 db = m->p[hdst(m, hnext(m, gs[j]))];
            *out = ddiv(dadd(dadd(dsub(da, m->p[e[j - 1]]), dmul(dsub(m->p[w], m->p[v]), 4.0)),
                             dsub(db, m->p[e[j + 1]])), 6.0);
            return 0;
        }
        return -1;
    }
    const int n = m->vfan[v];
    int h = m->vfirst[v], found = 0;
    for (int k = 0; k < n; ++k) {
        if (hdst(m, h) == w) { found = 1; break; }
        h = hrot(m, h);
    }
    if (!found) return -1;
    const double pi = 3.14159265358979323846;
    const double c1 = cos(2.0 * pi / n);
    const double an = 1.0 + c1 + cos(pi / n) * sqrt(2.0 * (9.0 + c1));
    d3 s = DZERO;
    for (int k = 0; k < n; ++k) {
        const double ck = cos(2.0 * pi * k / n), ck1 = cos(2.0 * pi * (k + 1) / n);
        s = dadd(dadd(s, dmul(m->p[hdst(m, h)], an * ck)), dmul(m->p[hdst(m, hnext(m, h))], ck + ck1));
        h = hrot(m, h);
    }
    *out = ddiv(s, 3.0 * n);
    return 0;
}

/* end cap: limit positions at the corners, edge points a third of the limit
 * tangent along each edge, inner points those of the face's B-spline patch */
static int end_cap(const mesh* m, int f, patch_out* o) {
    if (fsize(m, f) != 4) return -1;
    const int h0 = m->off[f];
    int c[4];
    for (int i = 0; i < 4; ++i) {
        c[i] = m->fv[h0 + i];
        if (!m->complete[c[i]] || !m->valid[c[i]]) return -1;
        int g = m->vfirst[c[i]];
        for (int k = 0; k < m->vfan[c[i]]; ++k) {
            if (fsize(m, m->hface[g]) != 4) return -1;
            if (!m->valid[hdst(m, g)] || !m->valid[hdst(m, hnext(m, g))] || !m->valid[horg(m, hprev(m, g))]) return -1;
            g = hrot(m, g);
            if (g < 0) break;
        }
    }
    d3 lim[4], b[16];
    for (int i = 0; i < 4; ++i) lim[i] = limit_pos(m, c[i]);
    b[0] = lim[0]; b[3] = lim[1]; b[15] = lim[2]; b[12] = lim[3];
    static const int ep[8][3] = {{1, 0, 1}, {4, 0, 3}, {2, 1, 0}, {7, 1, 2}, {11, 2, 1}, {14, 2, 3}, {13, 3, 2}, {8, 3, 0}};
    for (int e = 0; e < 8; ++e) {
        d3 t;
        if (limit_tan(m, c[ep[e][1]], c[ep[e][2]], &t)) return -1;
        b[ep[e][0]] = dadd(lim[ep[e][1]], ddiv(t, 3.0));
    }
    const d3 p0 = m->p[c[0]], p1 = m->p[c[1]], p2 = m->p[c[2]], p3 = m->p[c[3]];
    b[5] = ddiv(dadd(dadd(dadd(dmul(p0, 4.0), dmul(p1, 2.0)), dmul(p3, 2.0)), p2), 9.0);
    b[6] = ddiv(dadd(dadd(dadd(dmul(p1, 4.0), dmul(p0, 2.0)), dmul(p2, 2.0)), p3), 9.0);
    b[10] = ddiv(dadd(dadd(dadd(dmul(p2, 4.0), dmul(p1, 2.0)), dmul(p3, 2.0)), p0), 9.0);
    b[9] = ddiv(dadd(dadd(dadd(dmul(p3, 4.0), dmul(p0, 2.0)), dmul(p2, 2.0)), p1), 9.0);
    put(o, b);
    return 0;
}

/* faces within two vertex rings of the marked ones */
static unsigned char* two_rings(const mesh* m, const unsigned char* mark) {
    unsigned char* cur = (unsigned char*)malloc((size_t)m->nf + 1);
    memcpy(cur, mark, (size_t)m->nf);
    unsigned char* vm = (unsigned char*)malloc((size_t)m->nv + 1);
    for (int ring = 0; ring < 2; ++ring) {
        memset(vm, 0, (size_t)m->nv);
        for (int f = 0; f < m->nf; ++f)
            if (cur[f])
                for (int h = m->off[f]; h < m->off[f + 1]; ++h) vm[m->fv[h]] = 1;
        for (int f = 0; f < m->nf; ++f) {
            int touch = 0;
            for (int h = m->off[f]; h < m->off[f + 1]; ++h) touch |= vm[m->fv[h]];
            cur[f] = (unsigned char)touch;
        }
    }
    free(vm);
    return cur;
}

/* one Catmull-Clark step over the faces in S (topology of m built) */
static void refine(const mesh* m, const unsigned char* S, const unsigned char* split, mesh* c) {
    const int nf = m->nf, nh = m->nh, nv = m->nv;
    int* fpi = (int*)malloc(sizeof(int) * (nf + 1));
    int* epi = (int*)malloc(sizeof(int) * (nh + 1));
    int* vpi = (int*)malloc(sizeof(int) * (nv + 1));
    int* eh = (int*)malloc(sizeof(int) * (nh + 1));
    int* vv = (int*)malloc(sizeof(int) * (nv + 1));
    for (int i = 0; i < nf; ++i) fpi[i] = -1;
    for (int i = 0; i < nh; ++i) epi[i] = -1;
    for (int i = 0; i < nv; ++i) vpi[i] = -1;
    int n = 0, ne = 0, nvv = 0, nsf = 0, nsh = 0;
    for (int f = 0; f < nf; ++f)
        if (S[f]) { fpi[f] = n++; nsf++; nsh += fsize(m, f); }
    const int ebase = n;
    for (int f = 0; f < nf; ++f) {
        if (!S[f]) continue;
        for (int h = m->off[f]; h < m->off[f + 1]; ++h) {
            if (epi[h] >= 0) continue;
            epi[h] = n;
            if (m->twin[h] >= 0) epi[m->twin[h]] = n;
            eh[ne++] = h;
            n++;
        }
    }
    const int vbase = n;
    for (int f = 0; f < nf; ++f) {
        if (!S[f]) continue;
        for (int h = m->off[f]; h < m->off[f + 1]; ++h) {
            const int v = m->fv[h];
            if (vpi[v] >= 0) continue;
            vpi[v] = n++;
            vv[nvv++] = v;
        }
    }
    memset(c, 0, sizeof *c);
    c->nv = n;
    c->p = (d3*)calloc((size_t)n + 1, sizeof(d3));
    c->valid = (unsigned char*)calloc((size_t)n + 1, 1);
    c->complete = (unsigned char*)calloc((size_t)n + 1, 1);
    for (int f = 0; f < nf; ++f) {
        if (!S[f]) continue;
        d3 s = DZERO;
        int ok = 1;
        for (int h = m->off[f]; h < m->off[f + 1]; ++h) {
            s = dadd(s, m->p[m->fv[h]]);
            ok = ok && m->valid[m->fv[h]];
        }
        c->p[fpi[f]] = ddiv(s, (double)fsize(m, f));
        c->valid[fpi[f]] = (unsigned char)ok;
        c->complete[fpi[f]] = 1;
    }
    for (int k = 0; k < ne; ++k) {
        const int h = eh[k], id = ebase + k, a = horg(m, h), b = hdst(m, h), t = m->twin[h];
        if (t < 0) {
            if (m->complete[a] || m->complete[b]) {
                c->p[id] = dmul(dadd(m->p[a], m->p[b]), 0.5);
                c->valid[id] = (unsigned char)(m->valid[a] && m->valid[b]);
                c->complete[id] = 1;
            }
        } else if (S[m->hface[t]]) {
            const int fa = fpi[m->hface[h]], fb = fpi[m->hface[t]];
            c->p[id] = dmul(dadd(dadd(dadd(m->p[a], m->p[b]), c->p[fa]), c->p[fb]), 0.25);
            c->valid[id] = (unsigned char)(m->valid[a] && m->valid[b] && c->valid[fa] && c->valid[fb]);
            c->complete[id] = 1;
        }
    }
    for (int k = 0; k < nvv; ++k) {
        const int v = vv[k], id = vbase + k;
        if (!m->complete[v]) continue;
        const int first = m->vfirst[v];
        int all = 1, last = first, g = first;
        for (;;) {
            if (!S[m->hface[g]]) all = 0;
            last = g;
            const int t = hrot(m, g);
            if (t < 0 || t == first) break;
            g = t;
        }
        c->complete[id] = (unsigned char)all;
        if (m->vbound[v]) {
            const int ea = hdst(m, first), eb = horg(m, hprev(m, last));
            c->p[id] = dmul(dadd(dadd(m->p[ea], dmul(m->p[v], 6.0)), m->p[eb]), 0.125);
            c->valid[id] = (unsigned char)(m->valid[v] && m->valid[ea] && m->valid[eb]);
        } else if (all) {
            const int nn = m->vfan[v];
            d3 se = DZERO, sf = DZERO;
            int ok = m->valid[v];
            g = first;
            for (int j = 0; j < nn; ++j) {
                const int fp = fpi[m->hface[g]];
                se = dadd(se, m->p[hdst(m, g)]);
                sf = dadd(sf, c->p[fp]);
                ok = ok && m->valid[hdst(m, g)] && c->valid[fp];
                g = hrot(m, g);
            }
            c->p[id] = dadd(dmul(m->p[v], (double)(nn - 2) / (double)nn), ddiv(dadd(se, sf), (double)(nn * nn)));
            c->valid[id] = (unsigned char)ok;
        }
    }
    c->nf = nsh;
    c->nh = 4 * nsh;
    c->off = (int*)malloc(sizeof(int) * (nsh + 1));
    c->fv = (int*)malloc(sizeof(int) * (4 * nsh + 1));
    c->active = (unsigned char*)malloc((size_t)nsh + 1);
    int cf = 0;
    c->off[0] = 0;
    for (int f = 0; f < nf; ++f) {
        if (!S[f]) continue;
        const int k = fsize(m, f);
        for (int i = 0; i < k; ++i) {
            const int h = m->off[f] + i, hp = m->off[f] + (i + k - 1) % k;
            int* q = c->fv + 4 * cf;
            q[0] = vpi[m->fv[h]];
            q[1] = epi[h];
            q[2] = fpi[f];
            q[3] = epi[hp];
            c->active[cf] = split[f];
            cf++;
            c->off[cf] = 4 * cf;
        }
    }
    (void)nsf;
    free(fpi); free(epi); free(vpi); free(eh); free(vv);
}

int64_t ora_subdiv_patches_ex(const float* verts, int n_verts, const int32_t* counts, int n_faces,
                              const int32_t* indices, int max_level, int min_level, int prune, int cap_all,
                              float** out) {
    *out = NULL;
    if (n_verts < 0 || n_faces < 0) return -1;
    patch_out o = {NULL, 0, 0};
    if (n_faces == 0) return 0;
    mesh m;
    memset(&m, 0, sizeof m);
    m.nv = n_verts;
    m.nf = n_faces;
    int64_t total = 0;
    for (int f = 0; f < n_faces; ++f) {
        if (counts[f] < 3) return -1;
        total += counts[f];
    }
    m.nh = (int)total;
    m.p = (d3*)malloc(sizeof(d3) * ((size_t)n_verts + 1));
    m.valid = (unsigned char*)malloc((size_t)n_verts + 1);
    m.complete = (unsigned char*)malloc((size_t)n_verts + 1);
    for (int i = 0; i < n_verts; ++i) {
        m.p[i].x = verts[3 * i]; m.p[i].y = verts[3 * i + 1]; m.p[i].z = verts[3 * i + 2];
        m.valid[i] = 1; m.complete[i] = 1;
    }
    m.off = (int*)malloc(sizeof(int) * ((size_t)n_faces + 1));
    m.fv = (int*)malloc(sizeof(int) * ((size_t)total + 1));
    m.active = (unsigned char*)malloc((size_t)n_faces + 1);
    m.off[0] = 0;
    int64_t k = 0;
    for (int f = 0; f < n_faces; ++f) {
        for (int i = 0; i < counts[f]; ++i) {
            const int v = indices[k + i];
            if (v < 0 || v >= n_verts) { mesh_free(&m); return -1; }
            for (int j = 0; j < i; ++j)
                if (indices[k + j] == v) { mesh_free(&m); return -1; }
            m.fv[k + i] = v;
        }
        k += counts[f];
        m.off[f + 1] = (int)k;
        m.active[f] = 1;
    }
    int rc = 0;
    for (int level = 0; level <= max_level && rc == 0; ++level) {
        if (topology(&m, level == 0)) { rc = -1; break; }
        unsigned char* split = (unsigned char*)calloc((size_t)m.nf + 1, 1);
        int any = 0;
        for (int f = 0; f < m.nf && rc == 0; ++f) {
            if (!m.active[f]) continue;
            if (level == max_level && cap_all) rc = end_cap(&m, f, &o);
            else if (level >= min_level && fregular(&m, f)) rc = bspline_patch(&m, f, &o);
            else if (level == max_level) rc = end_cap(&m, f, &o);
            else { split[f] = 1; any = 1; }
        }
        if (rc || !any) { free(split); break; }
        unsigned char* S;
        if (prune) {
            S = two_rings(&m, split);
        } else {
            S = (unsigned char*)malloc((size_t)m.nf + 1);
            memset(S, 1, (size_t)m.nf);
        }
        mesh c;
        refine(&m, S, split, &c);
        free(S);
        free(split);
        mesh_free(&m);
        m = c;
    }
    mesh_free(&m);
    if (rc) { free(o.p); return -1; }
    *out = o.p;
    return o.n;
}

int64_t ora_subdiv_patches(const float* verts, int n_verts, const int32_t* counts, int n_faces,
                           const int32_t* indices, float** out) {
    return ora_subdiv_patches_ex(verts, n_verts, counts, n_faces, indices, 5, 0, 1, 0, out);
}

void ora_free(void* p) { free(p); }
