// Host_Triangulation.h -- drop-in HostTriangulation (reference
// include/Host_Triangulation.h:24-98) over libmcleap_hip.so.
//
// Same public data (m_pos, m_v, m_he, m_t in the reference AoS layouts, the
// four corners at +-1e7 as vertices 0..3) and the same member names.  The
// construction and the Lawson/untangle operations run in the library with
// exact predicates; the neighbourhood queries follow the reference's walks.
#pragma once

#include <algorithm>
#include <stack>
#include <vector>

#include "Common_Triangulation.h"

struct HostTriangulation {
    using vec2 = mcleap::vec2;
    // -------------------------------------------
    // host data (Host_Triangulation.h:27-30)
    std::vector<vec2> m_pos;
    std::vector<Vertex> m_v;
    std::vector<HalfEdge> m_he;
    std::vector<Triangle> m_t;
    // unit-torus triangulation (no corners) instead of the box with corners
    bool periodic = false;

    // -------------------------------------------
    // constructors (Host_Triangulation.h:107-137)
    HostTriangulation() { rebuild({}); }
    explicit HostTriangulation(bool torus) : periodic(torus) {
        if (!periodic) rebuild({});
    }

    // -------------------------------------------
    // incremental construction
    int findContainingTriangleIndexCheckingAll(vec2& p) {
        for (int t = 0; t < (int)m_t.size(); t++)
            if (isInside(p, t)) return t;
        return -1;
    }
    // visibility walk from t_index_guess (Host_Triangulation.h:188-246), exact orientation
    int findContainingTriangleIndexWalking(const vec2& p, int t_index_guess) {
        int t = t_index_guess;
        for (size_t steps = 0; steps <= 4 * m_t.size() + 8 && t >= 0; steps++) {
            int h = m_t[t].he, moved = -1;
            for (int k = 0; k < 3; k++, h = m_he[h].next) {
                if (orient(m_pos[m_he[h].v], m_pos[m_he[h ^ 1].v], p) < 0) { moved = m_he[h ^ 1].t; break; }
            }
            if (moved == -1 && isInsideOrOn(p, t)) return t;
            t = moved;
        }
        return findContainingTriangleIndexCheckingAll(const_cast<vec2&>(p));
    }
    // insertion without Lawson flips is not offered: every insertion keeps the
    // triangulation Delaunay (the reference's addPoint followed by delonizeVertex)
    bool addPoint(vec2 p, int) { return addDelaunayPoint(p); }
    bool addPoint(vec2 p) { return addDelaunayPoint(p); }
    bool addDelaunayPoint(vec2 p) {
        std::vector<vec2> one(1, p);
        return addDelaunayPoints(one);
    }
    bool sortPointsToAdd(std::vector<vec2>) { return false; }  // the reference stub
    bool addPoints(std::vector<vec2> points) { return addDelaunayPoints(points); }
    // addDelaunayPoints (Host_Triangulation.h:296-300).  The library builds the
    // exact Delaunay triangulation of all points (the unique one, ties broken
    // by the symbolic perturbation) inserting in Hilbert order.  Existing
    // vertices keep their ids; the new points get the following ids in
    // Hilbert order and, like the reference's in-place sort, the caller's
    // vector is reordered to that order (vertex old+k holds points[k]).
    bool addDelaunayPoints(std::vector<vec2>& points) {
        std::vector<vec2> pts(m_pos.begin() + corners(), m_pos.end());
        const int old = (int)pts.size();
        pts.insert(pts.end(), points.begin(), points.end());
        std::vector<int> perm;
        if (!rebuild(pts, &perm)) return false;
        // perm[k] = input index held by vertex corners()+k; renumber so that
        // inputs < old keep their vertex, new inputs follow in Hilbert order
        const int c = corners(), n = (int)perm.size();
        std::vector<int> id(c + n);
        for (int k = 0; k < c; k++) id[k] = k;
        int next = old;
        for (int k = 0; k < n; k++) id[c + k] = c + (perm[k] < old ? perm[k] : next++);
        std::vector<vec2> pos(m_pos.size());
        std::vector<Vertex> vv(m_v.size());
        for (int k = 0; k < c + n; k++) {
            pos[id[k]] = m_pos[k];
            vv[id[k]] = Vertex{id[k], m_v[k].he};
        }
        for (auto& h : m_he) {
            h.v = id[h.v];
            if (h.op >= 0) h.op = id[h.op];
        }
        m_pos.swap(pos);
        m_v.swap(vv);
        for (int k = 0; k < (int)points.size(); k++) points[k] = m_pos[c + old + k];
        return true;
    }

    bool isInside(vec2& p, int t_index) {
        const Triangle t = m_t[t_index];
        const int v[3] = {m_he[t.he].v, m_he[t.he ^ 1].v, m_he[t.he].op};
        for (int i = 0; i < 3; i++)
            if (orient(m_pos[v[i]], m_pos[v[(i + 1) % 3]], p) <= 0) return false;
        return true;
    }

    // -------------------------------------------
    // delonizing (Host_Triangulation.h:305-395)
    bool delonize() { return lawson(nullptr, 0); }
    bool delonizeEdge(int he_index) { return lawson(&he_index, 1); }
    bool delonizeTriangle(int t_index) {
        const int h = m_t[t_index].he;
        const int hs[3] = {h, m_he[h].next, m_he[m_he[h].next].next};
        return lawson(hs, 3);
    }
    bool delonizeVertex(int v_index) {
        std::vector<int> hs;
        const int h0 = m_v[v_index].he;
        int h = h0;
        do {
            hs.push_back(m_he[h].next);  // the link edge opposite v
            h = m_he[h ^ 1].next;
        } while (h != h0 && hs.size() < m_he.size());
        return lawson(hs.data(), (int)hs.size());
    }

    // -------------------------------------------
    // moving points
    bool movePoint(int p_index, vec2 d) {  // Host_Triangulation.h:403-406: moves without fixing anything
        m_pos[m_v[p_index].pos_index] += d;
        return true;
    }
    // declared and left unimplemented in the reference (Host_Triangulation.h:77-80,
    // 408-413): remove and re-insert p_index at its position + d (then Lawson
    // flips), or move it and untangle (then Lawson flips); false when the
    // vertex could not get there (a target on another vertex or outside the
    // mesh) or, untangling, a triangle stays inverted
    bool moveFlipflop(int p_index, vec2 d) { return moveVertex(p_index, d, MCL_MOVE_FLIPFLOP); }
    bool moveFlipflopDelaunay(int p_index, vec2 d) { return moveVertex(p_index, d, MCL_MOVE_FLIPFLOP_DELAUNAY); }
    bool moveUntangling(int p_index, vec2 d) { return moveVertex(p_index, d, MCL_MOVE_UNTANGLING); }
    bool moveUntanglingDelaunay(int p_index, vec2 d) { return moveVertex(p_index, d, MCL_MOVE_UNTANGLING_DELAUNAY); }
    bool moveVertex(int p_index, vec2 d, int mode) {
        long long f = 0;
        const float dd[2] = {d.x, d.y};
        return mcl_hmesh_move_vertex(fp(), ip(m_he), (int)m_he.size(), ip(m_t), (int)m_t.size(), ip(m_v),
                                     (int)m_v.size(), periodic, p_index, dd, mode, &f) == MCL_OK;
    }

    // -------------------------------------------
    // untangling (Host_Triangulation.h:417-471): one pass of B/C steps
    bool untangle() {
        long long f = 0;
        return mcleap::ok(mcl_hmesh_untangle(fp(), ip(m_he), (int)m_he.size(), ip(m_t), (int)m_t.size(), ip(m_v),
                                             (int)m_v.size(), periodic, &f),
                          "untangle");
    }
    bool untangleEdge(int) { return untangle(); }
    bool untangleTriangle(int) { return untangle(); }

    // -------------------------------------------
    // neighbourhood searching (Host_Triangulation.h:492-560)
    // CCW walk from v.he (skipping the outer face), then the backward walk
    // from the same start, as the reference does: an interior vertex gets its
    // neighbours, then all but the first again in reverse order
    std::vector<int> oneRing(int v_index) {
        std::vector<int> ring;
        int initial = m_v[v_index].he;
        if (m_he[initial].t == -1) initial ^= 1;
        ring.push_back(m_he[initial ^ 1].v);
        for (int h = m_he[m_he[initial].next].next ^ 1; h != initial && m_he[h].t != -1;
             h = m_he[m_he[h].next].next ^ 1)
            ring.push_back(m_he[h ^ 1].v);
        for (int h = m_he[initial ^ 1].next; h != initial && m_he[h].t != -1; h = m_he[h ^ 1].next)
            ring.push_back(m_he[h ^ 1].v);
        return ring;
    }
    // depth-first walk over one-rings collecting vertices within r
    std::vector<int> getFRNN(int v_index, float r) {
        std::vector<int> out;
        std::vector<char> visited(m_pos.size(), 0);
        std::stack<int> st;
        const vec2 c = m_pos[v_index];
        const float r2 = r * r;
        visited[v_index] = 1;
        st.push(v_index);
        while (!st.empty()) {
            const int s = st.top();
            st.pop();
            for (int a : oneRing(s)) {
                if (visited[a]) continue;
                visited[a] = 1;
                const float dx = c.x - m_pos[a].x, dy = c.y - m_pos[a].y;
                if (dx * dx + dy * dy <= r2) {
                    out.push_back(a);
                    st.push(a);
                }
            }
        }
        return out;
    }

    bool swapVertices(int, int) { return false; }  // experimental stub in the reference

    // -------------------------------------------
    // helpers
    int corners() const { return periodic ? 0 : 4; }
    int orient(const vec2& a, const vec2& b, const vec2& c) const {
        return mcl_orient2d(&a.x, &b.x, &c.x, periodic);
    }

  private:
    float* fp() { return reinterpret_cast<float*>(m_pos.data()); }
    template <class T>
    static int* ip(std::vector<T>& v) { return reinterpret_cast<int*>(v.data()); }
    bool isInsideOrOn(const vec2& p, int t_index) {
        const int h = m_t[t_index].he;
        const int v[3] = {m_he[h].v, m_he[h ^ 1].v, m_he[h].op};
        for (int i = 0; i < 3; i++)
            if (orient(m_pos[v[i]], m_pos[v[(i + 1) % 3]], p) < 0) return false;
        return true;
    }
    bool lawson(const int* seeds, int n) {
        long long f = 0;
        return mcleap::ok(mcl_hmesh_delonize(fp(), ip(m_he), (int)m_he.size(), ip(m_t), (int)m_t.size(), ip(m_v),
                                             (int)m_v.size(), periodic, seeds, n, &f),
                          "delonize");
    }
    bool rebuild(const std::vector<vec2>& pts, std::vector<int>* perm = nullptr) {
        mcl_hmesh_t* h = nullptr;
        int rc = mcl_hmesh_build(reinterpret_cast<const float*>(pts.data()), (int)pts.size(), periodic, &h);
        if (!mcleap::ok(rc, "addDelaunayPoints")) return false;
        int npos, nv, nhe, nt;
        mcl_hmesh_sizes(h, &npos, &nv, &nhe, &nt);
        m_pos.resize(npos);
        m_v.resize(nv);
        m_he.resize(nhe);
        m_t.resize(nt);
        if (perm) perm->resize(pts.size());
        rc = mcl_hmesh_export(h, fp(), ip(m_v), ip(m_he), ip(m_t), perm ? perm->data() : nullptr);
        mcl_hmesh_free(h);
        return mcleap::ok(rc, "mcl_hmesh_export");
    }
};
