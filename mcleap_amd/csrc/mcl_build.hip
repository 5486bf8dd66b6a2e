// mcl_build.hip -- host-side construction of the initial Delaunay triangulation.
//
// Box domain: the layout of HostTriangulation (Host_Triangulation.h:107-137):
// four corner vertices at +-1e7 (ids 0..3), half-edge twin = i^1,
// HalfEdge{next, v, t, op}, t = op = -1 on the outer face.  Points are
// inserted one at a time (addDelaunayPoint, Host_Triangulation.h:248-280):
// locate by a visibility walk from the previously created triangle, split the
// triangle 1->3 (f1to3 semantics, Common_Triangulation.h:129-209) or, for a
// point exactly on an edge, the two triangles 2->4, then legalise the edges
// opposite the new vertex by flips (delonizeVertex, Host_Triangulation.h:368-
// 395).  Predicates are exact (mcl_geom.hpp).  Points are first ordered along
// a Hilbert curve; the vertex of input point perm[k] is 4 + k.  The reference
// sorts by a Morton key of truncated coordinates, which is constant for
// inputs in [0, 1)^2 (Host_Triangulation.h:287-300).
//
// Torus domain (no reference counterpart): points are copied across the
// boundary within a margin, the planar triangulation of all copies is built
// with shift-exact predicates, and one triangle per torus orbit is kept (the
// copy whose smallest vertex id is untranslated).  The margin doubles until
// the result closes up into a torus (every directed edge has its reverse,
// 2n triangles).  The half-edge mesh has no corners: V = n, HE = 6n, T = 2n.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "mcl_geom.hpp"
#include "mcl_internal.hpp"

namespace mcl {


std::vector<int> hilbert_order(const float* xy, int n) {
    float lo[2] = {1e30f, 1e30f}, hi[2] = {-1e30f, -1e30f};
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 2; k++) {
            lo[k] = std::min(lo[k], xy[2 * i + k]);
            hi[k] = std::max(hi[k], xy[2 * i + k]);
        }
    const uint32_t G = 1u << 16;
    double sx = (hi[0] > lo[0]) ? (G - 1) / ((double)hi[0] - lo[0]) : 0.0;
    double sy = (hi[1] > lo[1]) ? (G - 1) / ((double)hi[1] - lo[1]) : 0.0;
    std::vector<std::pair<uint64_t, int>> k(n);
    for (int i = 0; i < n; i++) {
        uint32_t gx = (uint32_t)(((double)xy[2 * i] - lo[0]) * sx);
        uint32_t gy = (uint32_t)(((double)xy[2 * i + 1] - lo[1]) * sy);
        k[i] = {hilbert_d(G, gx, gy), i};
    }
    std::sort(k.begin(), k.end());
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = k[i].second;
    return order;
}

// ------------------------------------------------- shift-exact predicates
// point (x + sx, y + sy); used by the torus construction on the host
struct SPt {
    float x, y;
    int sx, sy;
};

static int orient_s(const SPt& a, const SPt& b, const SPt& c) {
    if (!a.sx && !a.sy && !b.sx && !b.sy && !c.sx && !c.sy) {
        int st;
        return orient<false>(make_float2(a.x, a.y), make_float2(b.x, b.y), make_float2(c.x, c.y), false, &st);
    }
    Dexp acx = diff_exact(a.x, c.x, a.sx - c.sx, false), acy = diff_exact(a.y, c.y, a.sy - c.sy, false);
    Dexp bcx = diff_exact(b.x, c.x, b.sx - c.sx, false), bcy = diff_exact(b.y, c.y, b.sy - c.sy, false);
    const double e64 = 1.1102230246251565e-16;
    double ax = dexp_value(acx), ay = dexp_value(acy), bx = dexp_value(bcx), by = dexp_value(bcy);
    double dl = ax * by, dr = ay * bx, det = dl - dr, ds = fabs(dl) + fabs(dr);
    // differences may carry two roundings here: use twice the bound
    double bound = 2.0 * (3.0 + 16.0 * e64) * e64 * ds;
    if (ds > 1e-280 && (det > bound || -det > bound)) return det > 0 ? 1 : -1;
    return orient_exact_from(acx, acy, bcx, bcy);
}

static bool rank_gt(const SPt& p, const SPt& q) {
    if (p.x != q.x) return p.x > q.x;
    if (p.y != q.y) return p.y > q.y;
    if (p.sx != q.sx) return p.sx > q.sx;
    return p.sy > q.sy;
}

static int incircle_s(const SPt& a, const SPt& b, const SPt& c, const SPt& d) {
    if (!a.sx && !a.sy && !b.sx && !b.sy && !c.sx && !c.sy && !d.sx && !d.sy) {
        int st;
        return incircle<false, true>(make_float2(a.x, a.y), make_float2(b.x, b.y), make_float2(c.x, c.y),
                                     make_float2(d.x, d.y), false, &st);
    }
    Dexp adx = diff_exact(a.x, d.x, a.sx - d.sx, false), ady = diff_exact(a.y, d.y, a.sy - d.sy, false);
    Dexp bdx = diff_exact(b.x, d.x, b.sx - d.sx, false), bdy = diff_exact(b.y, d.y, b.sy - d.sy, false);
    Dexp cdx = diff_exact(c.x, d.x, c.sx - d.sx, false), cdy = diff_exact(c.y, d.y, c.sy - d.sy, false);
    {
        const double e = 1.1102230246251565e-16;
        double ax = dexp_value(adx), ay = dexp_value(ady), bx = dexp_value(bdx), by = dexp_value(bdy);
        double cx = dexp_value(cdx), cy = dexp_value(cdy);
        double bdxcdy = bx * cy, cdxbdy = cx * by, alift = ax * ax + ay * ay;
        double cdxady = cx * ay, adxcdy = ax * cy, blift = bx * bx + by * by;
        double adxbdy = ax * by, bdxady = bx * ay, clift = cx * cx + cy * cy;
        double det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) + clift * (adxbdy - bdxady);
        double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * alift + (fabs(cdxady) + fabs(adxcdy)) * blift +
                      (fabs(adxbdy) + fabs(bdxady)) * clift;
        double bound = 2.0 * (10.0 + 96.0 * e) * e * perm;
        if (perm > 1e-280 && (det > bound || -det > bound)) return det > 0 ? 1 : -1;
    }
    InSigns s;
    incircle_exact_from<36, 7776>(adx, ady, bdx, bdy, cdx, cdy, s);
    if (s.det != 0) return s.det;
    const SPt* P[4] = {&a, &b, &c, &d};
    int cof[4] = {s.bc, s.ca, s.ab, -s.abc};
    int o[4] = {0, 1, 2, 3};
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++)
            if (rank_gt(*P[o[j]], *P[o[i]])) std::swap(o[i], o[j]);
    for (int i = 0; i < 4; i++)
        if (cof[o[i]] != 0) return cof[o[i]];
    return 0;
}

// ------------------------------------------------ incremental construction
struct Builder {
    std::vector<SPt> P;  // vertex coordinates
    std::vector<int> nx, org, tri, opp;  // half-edges
    std::vector<int> vhe;                // vertex -> outgoing half-edge
    std::vector<int> the;                // triangle -> half-edge
    int last_t = 1;
    long long walk_steps = 0, flips = 0;

    void init_square() {
        const float lo = -10000000.f, hi = 10000000.f;  // Common_Triangulation.h:12-13
        P = {{lo, lo, 0, 0}, {hi, lo, 0, 0}, {hi, hi, 0, 0}, {lo, hi, 0, 0}};
        vhe = {0, 2, 6, 8};
        // next, v, t, op of the 10 initial half-edges (Host_Triangulation.h:119-133)
        const int init[10][4] = {{2, 0, 0, 2}, {9, 1, -1, -1}, {4, 1, 0, 0}, {1, 2, -1, -1}, {0, 2, 0, 1},
                                 {6, 0, 1, 3}, {8, 2, 1, 0},   {3, 3, -1, -1}, {5, 3, 1, 2}, {7, 0, -1, -1}};
        for (auto& r : init) { nx.push_back(r[0]); org.push_back(r[1]); tri.push_back(r[2]); opp.push_back(r[3]); }
        the = {4, 5};
    }
    int new_he_pair() {
        int h = (int)nx.size();
        for (int k = 0; k < 2; k++) { nx.push_back(0); org.push_back(0); tri.push_back(0); opp.push_back(0); }
        return h;
    }
    void set(int h, int n, int v, int t, int o) { nx[h] = n; org[h] = v; tri[h] = t; opp[h] = o; }

    // edge flip of he0 (f2to2 semantics, Common_Triangulation.h:69-127)
    void flip(int he0) {
        int he1 = he0 ^ 1, he2 = nx[he0], he3 = nx[he2], he4 = nx[he1], he5 = nx[he4];
        int t0 = tri[he0], t1 = tri[he1];
        int v0 = org[he0], v1 = org[he1], v2 = org[he3], v3 = org[he5];
        the[t0] = he0; the[t1] = he1;
        vhe[v1] = he2; vhe[v0] = he4;
        set(he0, he5, v2, t0, v1); set(he5, he2, v3, t0, v2); set(he2, he0, v1, t0, v3);
        set(he1, he3, v3, t1, v0); set(he3, he4, v2, t1, v3); set(he4, he1, v0, t1, v2);
        flips++;
    }
    bool illegal(int h) {
        if (tri[h] < 0 || tri[h ^ 1] < 0) return false;
        // inCircle(org(he), op(twin), org(twin), op(he)) > 0 (Host_Triangulation.h:330-340)
        return incircle_s(P[org[h]], P[opp[h ^ 1]], P[org[h ^ 1]], P[opp[h]]) > 0;
    }
    void legalize_vertex(int v) {
        int initial = vhe[v], cur = initial;
        do {
            int link = nx[cur];
            while (illegal(link)) { flip(link); link = nx[cur]; }
            cur = nx[nx[cur]] ^ 1;
        } while (cur != initial);
    }
    // visibility walk; returns triangle and the half-edge p lies on (or -1)
    int locate(const SPt& p, int t, int* on_edge) {
        long long cap = 4LL * (long long)the.size() + 100, steps = 0;
        while (steps++ < cap) {
            int h0 = the[t], hs[3] = {h0, nx[h0], nx[nx[h0]]}, zero = -1;
            bool moved = false;
            for (int k = 0; k < 3; k++) {
                int h = hs[(k + (int)(steps % 3)) % 3];
                int o = orient_s(P[org[h]], P[org[h ^ 1]], p);
                if (o < 0) {
                    if (tri[h ^ 1] < 0) return -1;
                    t = tri[h ^ 1];
                    moved = true;
                    break;
                }
                if (o == 0) zero = h;
            }
            walk_steps++;
            if (!moved) { *on_edge = zero; return t; }
        }
        for (int tt = 0; tt < (int)the.size(); tt++) {  // the walk cycled: scan
            int h0 = the[tt], hs[3] = {h0, nx[h0], nx[nx[h0]]}, zero = -1;
            bool in = true;
            for (int k = 0; k < 3 && in; k++) {
                int o = orient_s(P[org[hs[k]]], P[org[hs[k] ^ 1]], p);
                if (o < 0) in = false;
                else if (o == 0) zero = hs[k];
            }
            if (in) { *on_edge = zero; return tt; }
        }
        return -1;
    }
    bool same(const SPt& a, const SPt& b) { return a.x == b.x && a.y == b.y && a.sx == b.sx && a.sy == b.sy; }
    // returns 0 ok, 1 duplicate, 2 not located
    int insert(const SPt& p) {
        int on_edge = -1;
        int t = locate(p, last_t, &on_edge);
        if (t < 0) return 2;
        int h0 = the[t];
        int hs[3] = {h0, nx[h0], nx[nx[h0]]};
        for (int k = 0; k < 3; k++)
            if (same(P[org[hs[k]]], p)) return 1;
        int v = (int)P.size();
        P.push_back(p);
        vhe.push_back(0);
        int tA = (int)the.size(), tB = tA + 1;
        the.push_back(0); the.push_back(0);
        int n0 = new_he_pair(), n1 = new_he_pair(), n2 = new_he_pair();
        if (on_edge < 0) {
            // 1 -> 3 split of t = (a, b, c) by he0 = a->b, he1 = b->c, he2 = c->a
            int e0 = hs[0], e1 = hs[1], e2 = hs[2];
            int a = org[e0], b = org[e1], c = org[e2];
            // t = (a, b, v): e0, n0 (b->v), n2^1 (v->a)
            // tA = (b, c, v): e1, n1 (c->v), n0^1 (v->b)
            // tB = (c, a, v): e2, n2 (a->v), n1^1 (v->c)
            set(e0, n0, a, t, v); set(n0, n2 ^ 1, b, t, a); set(n2 ^ 1, e0, v, t, b);
            set(e1, n1, b, tA, v); set(n1, n0 ^ 1, c, tA, b); set(n0 ^ 1, e1, v, tA, c);
            set(e2, n2, c, tB, v); set(n2, n1 ^ 1, a, tB, c); set(n1 ^ 1, e2, v, tB, a);
            the[t] = e0; the[tA] = e1; the[tB] = e2;
            vhe[v] = n2 ^ 1;
            last_t = tB;
        } else {
            int he0 = on_edge, he1 = he0 ^ 1;
            if (tri[he1] < 0) return 2;
            int he2 = nx[he0], he3 = nx[he2], he4 = nx[he1], he5 = nx[he4];
            int t0 = tri[he0], t1 = tri[he1];
            int a = org[he0], b = org[he1], c = org[he3], d = org[he5];
            // t0 = (a, v, c), tA = (v, b, c), t1 = (v, a, d), tB = (b, v, d)
            set(he0, n1, a, t0, c); set(n1, he3, v, t0, a); set(he3, he0, c, t0, v);
            set(n0, he2, v, tA, c); set(he2, n1 ^ 1, b, tA, v); set(n1 ^ 1, n0, c, tA, b);
            set(he1, he4, v, t1, d); set(he4, n2 ^ 1, a, t1, v); set(n2 ^ 1, he1, d, t1, a);
            set(n0 ^ 1, n2, b, tB, d); set(n2, he5, v, tB, b); set(he5, n0 ^ 1, d, tB, v);
            the[t0] = he0; the[tA] = n0; the[t1] = he1; the[tB] = n0 ^ 1;
            vhe[v] = n0; vhe[a] = he0; vhe[b] = he2;
            last_t = tB;
        }
        legalize_vertex(v);
        return 0;
    }
};

// ------------------------------------------------------------------ entry
int build_box(const float* xy, int n, HostMesh& m) {
    Builder B;
    B.init_square();
    std::vector<int> order = hilbert_order(xy, n);
    B.P.reserve(n + 4);
    B.nx.reserve(6 * (size_t)n + 10); B.org.reserve(6 * (size_t)n + 10);
    B.tri.reserve(6 * (size_t)n + 10); B.opp.reserve(6 * (size_t)n + 10);
    for (int k = 0; k < n; k++) {
        int i = order[k];
        SPt p{xy[2 * i], xy[2 * i + 1], 0, 0};
        if (!(p.x > -1e7f && p.x < 1e7f && p.y > -1e7f && p.y < 1e7f)) return MCL_ERR_INPUT;
        int r = B.insert(p);
        if (r == 1) return MCL_ERR_DUPLICATE;
        if (r) return MCL_ERR_INPUT;
    }
    m.periodic = false;
    m.nv = (int)B.P.size();
    m.pos.resize(2 * (size_t)m.nv);
    for (int v = 0; v < m.nv; v++) { m.pos[2 * v] = B.P[v].x; m.pos[2 * v + 1] = B.P[v].y; }
    m.vrec.resize(2 * (size_t)m.nv);
    for (int v = 0; v < m.nv; v++) { m.vrec[2 * v] = v; m.vrec[2 * v + 1] = B.vhe[v]; }
    size_t nhe = B.nx.size();
    m.he.resize(4 * nhe);
    for (size_t h = 0; h < nhe; h++) {
        m.he[4 * h] = B.nx[h]; m.he[4 * h + 1] = B.org[h]; m.he[4 * h + 2] = B.tri[h]; m.he[4 * h + 3] = B.opp[h];
    }
    m.t = B.the;
    m.perm = order;
    m.flips = B.flips;
    return MCL_OK;
}

// half-edge mesh on the torus from a closed CCW triangle list over n vertices
static int mesh_from_triangles(int n, const std::vector<int>& tri3, HostMesh& m) {
    size_t nt = tri3.size() / 3;
    std::unordered_map<uint64_t, int> edge;
    edge.reserve(nt * 2);
    std::vector<int> he_of(3 * nt);
    int ne = 0;
    for (size_t t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++) {
            int a = tri3[3 * t + k], b = tri3[3 * t + (k + 1) % 3];
            uint64_t key = ((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
            auto it = edge.find(key);
            int e;
            if (it == edge.end()) { e = ne++; edge.emplace(key, e); }
            else e = it->second;
            he_of[3 * t + k] = 2 * e + (a < b ? 0 : 1);
        }
    if ((size_t)ne * 2 != 3 * nt) return MCL_ERR_TOPOLOGY;
    m.he.assign(4 * (size_t)ne * 2, -7);
    m.vrec.assign(2 * (size_t)n, -1);
    m.t.resize(nt);
    for (size_t t = 0; t < nt; t++) {
        for (int k = 0; k < 3; k++) {
            int h = he_of[3 * t + k], hn = he_of[3 * t + (k + 1) % 3];
            if (m.he[4 * (size_t)h] != -7) return MCL_ERR_TOPOLOGY;  // directed edge used twice
            m.he[4 * (size_t)h] = hn;
            m.he[4 * (size_t)h + 1] = tri3[3 * t + k];
            m.he[4 * (size_t)h + 2] = (int)t;
            m.he[4 * (size_t)h + 3] = tri3[3 * t + (k + 2) % 3];
            m.vrec[2 * (size_t)tri3[3 * t + k] + 1] = h;
        }
        m.t[t] = he_of[3 * t];
    }
    for (int v = 0; v < n; v++) {
        m.vrec[2 * (size_t)v] = v;
        if (m.vrec[2 * (size_t)v + 1] < 0) return MCL_ERR_TOPOLOGY;
    }
    return MCL_OK;
}

int build_torus(const float* xy_in, int n, HostMesh& m) {
    if (n < 16) return MCL_ERR_INPUT;
    for (int i = 0; i < 2 * n; i++)
        if (!(xy_in[i] >= 0.0f && xy_in[i] < 1.0f)) return MCL_ERR_INPUT;
    // vertex ids follow the Hilbert order of the points
    std::vector<int> order = hilbert_order(xy_in, n);
    std::vector<float> xy(2 * (size_t)n);
    for (int k = 0; k < n; k++) { xy[2 * k] = xy_in[2 * order[k]]; xy[2 * k + 1] = xy_in[2 * order[k] + 1]; }
    double margin = std::min(0.5, std::max(0.02, 6.0 / std::sqrt((double)n)));
    for (;;) {
        std::vector<SPt> pts;
        std::vector<int> vid;
        for (int i = 0; i < n; i++) { pts.push_back({xy[2 * i], xy[2 * i + 1], 0, 0}); vid.push_back(i); }
        for (int sx = -1; sx <= 1; sx++)
            for (int sy = -1; sy <= 1; sy++) {
                if (!sx && !sy) continue;
                for (int i = 0; i < n; i++) {
                    double x = (double)xy[2 * i] + sx, y = (double)xy[2 * i + 1] + sy;
                    if (x < -margin || x >= 1.0 + margin || y < -margin || y >= 1.0 + margin) continue;
                    pts.push_back({xy[2 * i], xy[2 * i + 1], sx, sy});
                    vid.push_back(i);
                }
            }
        // insertion order: Hilbert order of the unwrapped coordinates
        std::vector<float> uxy(2 * pts.size());
        for (size_t k = 0; k < pts.size(); k++) {
            uxy[2 * k] = (float)((double)pts[k].x + pts[k].sx);
            uxy[2 * k + 1] = (float)((double)pts[k].y + pts[k].sy);
        }
        std::vector<int> ins = hilbert_order(uxy.data(), (int)pts.size());
        Builder B;
        B.init_square();
        std::vector<int> bvid(4, -1), bpid(4, -1);
        bool ok = true;
        for (int k : ins) {
            if (B.insert(pts[k]) != 0) { ok = false; break; }
            bvid.push_back(vid[k]);
            bpid.push_back(k);
        }
        if (!ok) return MCL_ERR_DUPLICATE;
        std::vector<int> tri3;
        for (size_t t = 0; t < B.the.size(); t++) {
            int h = B.the[t];
            int vs[3] = {B.org[h], B.org[B.nx[h]], B.org[B.nx[B.nx[h]]]};
            if (vs[0] < 4 || vs[1] < 4 || vs[2] < 4) continue;
            int kmin = 0;
            for (int q = 1; q < 3; q++)
                if (bvid[vs[q]] < bvid[vs[kmin]]) kmin = q;
            const SPt& pm = pts[bpid[vs[kmin]]];
            if (pm.sx || pm.sy) continue;
            for (int q = 0; q < 3; q++) tri3.push_back(bvid[vs[q]]);
        }
        if ((int)tri3.size() == 6 * n && mesh_from_triangles(n, tri3, m) == MCL_OK) break;
        if (margin >= 0.5) return MCL_ERR_TOPOLOGY;
        margin = std::min(0.5, 2.0 * margin);
    }
    m.periodic = true;
    m.nv = n;
    m.pos = xy;
    m.perm = order;
    m.flips = 0;
    return MCL_OK;
}

}  // namespace mcl
