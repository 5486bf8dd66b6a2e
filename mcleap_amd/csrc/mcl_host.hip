// mcl_host.hip -- host-side operations of HostTriangulation on the caller's
// AoS arrays (Vertex{pos_index, he}, HalfEdge{next, v, t, op},
// Triangle{he}), with the library's exact predicates.  Host code only.
#include <vector>

#include "mcl_geom.hpp"
#include "mcl_internal.hpp"

namespace mcl {
namespace {

struct Aos {
    float2* pos;
    int* he;  // 4 ints per half-edge
    int nhe;
    int* t;
    int nt;
    int* v;  // 2 ints per vertex
    int nv;
    bool periodic;
    int& next(int h) { return he[4 * h]; }
    int& org(int h) { return he[4 * h + 1]; }
    int& tri(int h) { return he[4 * h + 2]; }
    int& opp(int h) { return he[4 * h + 3]; }
    float2 P(int vi) { return pos[v[2 * vi]]; }
};

int orient_h(const Aos& m, float2 a, float2 b, float2 c) {
    int st;
    return orient<false>(a, b, c, m.periodic, &st);
}

// f2to2 (Common_Triangulation.h:69-127): he0/he1 become the other diagonal
// (v2 -> v3 / v3 -> v2), the outer half-edges keep their origins
void flip_aos(Aos& m, int he0) {
    const int he1 = he0 ^ 1, he2 = m.next(he0), he3 = m.next(he2), he4 = m.next(he1), he5 = m.next(he4);
    const int t0 = m.tri(he0), t1 = m.tri(he1);
    const int v0 = m.org(he0), v1 = m.org(he1), v2 = m.org(he3), v3 = m.org(he5);
    m.t[t0] = he0;
    m.t[t1] = he1;
    m.v[2 * v1 + 1] = he2;
    m.v[2 * v0 + 1] = he4;
    auto set = [&](int h, int n, int vv, int tt, int o) { m.next(h) = n; m.org(h) = vv; m.tri(h) = tt; m.opp(h) = o; };
    set(he0, he5, v2, t0, v1); set(he5, he2, v3, t0, v2); set(he2, he0, v1, t0, v3);
    set(he1, he3, v3, t1, v0); set(he3, he4, v2, t1, v3); set(he4, he1, v0, t1, v2);
}

bool interior(Aos& m, int h) { return m.tri(h) >= 0 && m.tri(h ^ 1) >= 0; }

// inCircle(org(he), op(twin), org(twin), op(he)) > 0 (Host_Triangulation.h:330-340), exact with the tie rule
bool illegal(Aos& m, int h) {
    if (!interior(m, h)) return false;
    int st;
    return incircle<false, true>(m.P(m.org(h)), m.P(m.opp(h ^ 1)), m.P(m.org(h ^ 1)), m.P(m.opp(h)), m.periodic, &st) > 0;
}

// the flip of edge h keeps both triangles strictly CCW
bool flippable(Aos& m, int h) {
    const float2 a = m.P(m.org(h)), b = m.P(m.org(h ^ 1)), c = m.P(m.opp(h)), d = m.P(m.opp(h ^ 1));
    return orient_h(m, c, d, b) > 0 && orient_h(m, d, c, a) > 0;
}

}  // namespace
}  // namespace mcl

using mcl::Aos;

static bool aos_ok(const float* pos, int* he4, int nhe, int* t1, int nt, int* v2, int nv) {
    return pos && he4 && t1 && v2 && nhe > 0 && nt > 0 && nv > 0 && (nhe & 1) == 0;
}

extern "C" int mcl_orient2d(const float* a, const float* b, const float* c, int periodic) {
    int st;
    return mcl::orient<false>(make_float2(a[0], a[1]), make_float2(b[0], b[1]), make_float2(c[0], c[1]), periodic != 0,
                              &st);
}

extern "C" int mcl_incircle(const float* a, const float* b, const float* c, const float* d, int periodic) {
    int st;
    return mcl::incircle<false, true>(make_float2(a[0], a[1]), make_float2(b[0], b[1]), make_float2(c[0], c[1]),
                                      make_float2(d[0], d[1]), periodic != 0, &st);
}

extern "C" int mcl_hmesh_flip(float* pos, int* he4, int nhe, int* t1, int nt, int* v2, int nv, int periodic,
                              int he_index, int check) {
    if (!aos_ok(pos, he4, nhe, t1, nt, v2, nv) || he_index < 0 || he_index >= nhe) return MCL_ERR_INPUT;
    Aos m{(float2*)pos, he4, nhe, t1, nt, v2, nv, periodic != 0};
    if (!mcl::interior(m, he_index)) return MCL_ERR_INPUT;
    if (check && !mcl::flippable(m, he_index)) return MCL_ERR_TOPOLOGY;
    mcl::flip_aos(m, he_index);
    return MCL_OK;
}

extern "C" int mcl_hmesh_delonize(float* pos, int* he4, int nhe, int* t1, int nt, int* v2, int nv, int periodic,
                                  const int* seeds, int nseeds, long long* flips) {
    if (!aos_ok(pos, he4, nhe, t1, nt, v2, nv)) return MCL_ERR_INPUT;
    Aos m{(float2*)pos, he4, nhe, t1, nt, v2, nv, periodic != 0};
    std::vector<int> stack;
    std::vector<char> queued(nhe / 2, 0);
    if (seeds) {
        for (int k = 0; k < nseeds; k++)
            if (seeds[k] >= 0 && seeds[k] < nhe && !queued[seeds[k] >> 1]) { queued[seeds[k] >> 1] = 1; stack.push_back(seeds[k] >> 1); }
    } else {
        for (int e = nhe / 2 - 1; e >= 0; e--) { queued[e] = 1; stack.push_back(e); }
    }
    long long nf = 0, skipped = 0;
    while (!stack.empty()) {
        const int e = stack.back();
        stack.pop_back();
        queued[e] = 0;
        const int h = 2 * e;
        if (!mcl::illegal(m, h)) continue;
        if (!mcl::flippable(m, h)) { skipped++; continue; }  // reflex quad (a tangled mesh): leave it
        const int q[4] = {m.next(h), m.next(m.next(h)), m.next(h ^ 1), m.next(m.next(h ^ 1))};
        mcl::flip_aos(m, h);
        nf++;
        for (int k = 0; k < 4; k++)
            if (!queued[q[k] >> 1]) { queued[q[k] >> 1] = 1; stack.push_back(q[k] >> 1); }
    }
    if (flips) *flips = nf;
    return skipped ? MCL_ERR_UNRESOLVED : MCL_OK;
}

extern "C" int mcl_hmesh_untangle(float* pos, int* he4, int nhe, int* t1, int nt, int* v2, int nv, int periodic,
                                  long long* flips) {
    if (!aos_ok(pos, he4, nhe, t1, nt, v2, nv)) return MCL_ERR_INPUT;
    Aos m{(float2*)pos, he4, nhe, t1, nt, v2, nv, periodic != 0};
    long long nf = 0;
    // one pass over the edges (Host_Triangulation.h:417-423) of the B/C steps
    // (:425-471): for a creased edge (exactly one of its triangles CCW) flip
    // when the inverted triangle's opposite vertex lies strictly inside the
    // upright triangle (B) or the upright triangle's opposite vertex lies
    // strictly inside the inverted one (C); orientation exact
    for (int e = 0; e < nhe / 2; e++) {
        const int h = 2 * e;
        if (!mcl::interior(m, h)) continue;
        const float2 p0 = m.P(m.org(h)), p1 = m.P(m.org(h ^ 1)), a0 = m.P(m.opp(h)), a1 = m.P(m.opp(h ^ 1));
        const bool ccw0 = mcl::orient_h(m, p0, p1, a0) >= 0, ccw1 = mcl::orient_h(m, p1, p0, a1) >= 0;
        if (ccw0 == ccw1) continue;
        const float2 u0 = ccw0 ? p0 : p1, u1 = ccw0 ? p1 : p0, ua = ccw0 ? a0 : a1, ia = ccw0 ? a1 : a0;
        const bool inB = mcl::orient_h(m, u0, u1, ia) > 0 && mcl::orient_h(m, u1, ua, ia) > 0 &&
                         mcl::orient_h(m, ua, u0, ia) > 0;
        const bool inC = !inB && mcl::orient_h(m, u1, u0, ua) < 0 && mcl::orient_h(m, u0, ia, ua) < 0 &&
                         mcl::orient_h(m, ia, u1, ua) < 0;
        if (inB || inC) {
            mcl::flip_aos(m, h);
            nf++;
        }
    }
    if (flips) *flips = nf;
    return MCL_OK;
}
