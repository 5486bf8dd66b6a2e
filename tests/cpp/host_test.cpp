// host_test.cpp -- the host half of the drop-in interface (include/mcleap/
// Host_Triangulation.h) without a GPU: construction, incremental insertion
// with stable vertex ids, point location, the host Lawson and untangle
// passes, and the neighbourhood queries.  Writes the final mesh as text for
// the caller to compare with the oracle:
//   npos nhe nt / pos lines / he lines / t lines
#include <cstdio>
#include <random>
#include <vector>

#include "mcleap/Host_Triangulation.h"

using mcleap::vec2;

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 500;
    std::mt19937 rng{10};
    std::uniform_real_distribution<float> u(0.f, 1.f);
    std::vector<vec2> pts;
    for (int i = 0; i < n; i++) pts.push_back(vec2(u(rng), u(rng)));
    HostTriangulation h;
    std::vector<vec2> first(pts.begin(), pts.begin() + n / 2), second(pts.begin() + n / 2, pts.end());
    if (!h.addDelaunayPoints(first)) return 2;
    const vec2 keep = h.m_pos[7];
    if (!h.addDelaunayPoints(second)) return 3;
    if (h.m_pos[7].x != keep.x || h.m_pos[7].y != keep.y) return 4;  // existing ids are stable
    if (!h.addDelaunayPoint(vec2(0.123f, 0.456f))) return 5;
    if (h.m_pos.back().x != 0.123f) return 6;
    // point location
    vec2 q(0.5f, 0.5f);
    const int ta = h.findContainingTriangleIndexCheckingAll(q), tb = h.findContainingTriangleIndexWalking(q, 1);
    if (ta < 0 || ta != tb) return 7;
    // perturb every third point slightly and legalise on the host
    std::normal_distribution<float> g(0.f, 2e-4f);
    for (int i = 4; i < (int)h.m_pos.size(); i += 3) h.movePoint(i, vec2(g(rng), g(rng)));
    if (!h.untangle()) return 8;
    if (!h.delonize()) return 9;
    // remove-and-reinsert moves (moveFlipflop*, stubs in the reference), ending Delaunay
    for (int i = 10; i < (int)h.m_pos.size(); i += 37) {
        const vec2 p = h.m_pos[i], d(0.03f * g(rng) / 2e-4f, 0.03f * g(rng) / 2e-4f);
        if (p.x + d.x <= 0.f || p.x + d.x >= 1.f || p.y + d.y <= 0.f || p.y + d.y >= 1.f) continue;
        if (!h.moveFlipflop(i, d)) return 11;
        if (h.m_pos[i].x != p.x + d.x || h.m_pos[i].y != p.y + d.y) return 12;
    }
    if (!h.delonize()) return 13;
    if (!h.moveFlipflopDelaunay(9, vec2(0.001f, -0.002f))) return 14;
    // queries
    if (h.oneRing(20).size() < 5) return 10;
    std::printf("%zu %zu %zu\n", h.m_pos.size(), h.m_he.size(), h.m_t.size());
    for (auto& p : h.m_pos) std::printf("%.9g %.9g\n", p.x, p.y);
    for (auto& e : h.m_he) std::printf("%d %d %d %d\n", e.next, e.v, e.t, e.op);
    for (auto& t : h.m_t) std::printf("%d\n", t.he);
    return 0;
}
