// dropin_test.cpp -- the reference's C++ call sequence (moving_device_mesh.cu,
// tests/test_frnn.cu, tests/test_closest_neighbors.cu, tests/grid2d_frnn.cu)
// written against the drop-in headers include/mcleap/*.h, checked on the
// host with brute force and the library's exact predicates.
//   hipcc --offload-arch=gfx950 -I include tests/cpp/dropin_test.cpp -L mcleap_amd -lmcleap_hip
// Prints "DROPIN OK" and exits 0 when every check holds.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "mcleap/Device_Triangulation.h"
#include "mcleap/Host_Triangulation.h"
#include "mcleap/gridCount2d.h"

using mcleap::ivec2;
using mcleap::vec2;

static int failures = 0;
#define CHECK(cond, ...)                                       \
    do {                                                       \
        if (!(cond)) {                                         \
            if (failures < 20) {                               \
                std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
                std::printf(__VA_ARGS__);                      \
                std::printf("\n");                             \
            }                                                  \
            failures++;                                        \
        }                                                      \
    } while (0)

// tests/grid2d_frnn.cu:26-60
struct SaveNeighborsFunctor {
    float rad;
    int max_neighbors;
    int* num_neighbors;
    int* neighbors;
    __device__ void operator()(const int& i, const int& j, const vec2, const double dist) {
        if (i != j && dist <= rad) {
            const int k = atomicAdd(&num_neighbors[i], 1);
            if (k < max_neighbors) neighbors[i * max_neighbors + k] = j;
        }
    }
};

// every functor call on a periodic grid: its largest |d| and dist (float bits
// of non-negative values order like ints), and the pair count per slot
struct TorusProbeFunctor {
    float rad;
    int* max_bits;  // [0] max dist, [1] max |d.x|, [2] max |d.y|
    int* count;
    __device__ void operator()(const int& i, const int& j, const vec2 d, const double dist) {
        if (i == j) return;
        atomicMax(&max_bits[0], __float_as_int((float)dist));
        atomicMax(&max_bits[1], __float_as_int(fabsf(d.x)));
        atomicMax(&max_bits[2], __float_as_int(fabsf(d.y)));
        if (dist <= rad) atomicAdd(&count[i], 1);
    }
};

static float d2(const vec2& a, const vec2& b) {
    const float dx = a.x - b.x, dy = a.y - b.y;
    return dx * dx + dy * dy;
}

static void check_mesh(HostTriangulation& h, const char* tag) {
    int bad_ccw = 0, bad_del = 0;
    for (size_t t = 0; t < h.m_t.size(); t++) {
        const int e = h.m_t[t].he;
        const int a = h.m_he[e].v, b = h.m_he[h.m_he[e].next].v, c = h.m_he[h.m_he[h.m_he[e].next].next].v;
        if (h.orient(h.m_pos[a], h.m_pos[b], h.m_pos[c]) <= 0) bad_ccw++;
    }
    for (size_t e = 0; e < h.m_he.size(); e += 2) {
        const HalfEdge& x = h.m_he[e];
        const HalfEdge& y = h.m_he[e + 1];
        if (x.t < 0 || y.t < 0) continue;
        if (mcl_incircle(&h.m_pos[x.v].x, &h.m_pos[y.op].x, &h.m_pos[y.v].x, &h.m_pos[x.op].x, h.periodic) > 0) bad_del++;
    }
    CHECK(bad_ccw == 0 && bad_del == 0, "%s: %d non-CCW triangles, %d non-Delaunay edges", tag, bad_ccw, bad_del);
}

int main() {
    const int numP = 2000;
    const float bounds = 1000.f, movement = 0.5f;
    std::mt19937 rng{3};
    std::uniform_real_distribution<float> pos_r(-bounds, bounds);
    std::uniform_real_distribution<float> move_r(-movement, movement);
    std::vector<vec2> h_pos, h_move;
    for (int i = 0; i < numP; i++) h_pos.push_back(vec2(pos_r(rng), pos_r(rng)));
    for (int i = 0; i < numP + 4; i++) h_move.push_back(i < 4 ? vec2(0.f, 0.f) : vec2(move_r(rng), move_r(rng)));

    HostTriangulation* ht = new HostTriangulation();
    CHECK(ht->addDelaunayPoints(h_pos), "addDelaunayPoints");
    CHECK((int)ht->m_pos.size() == numP + 4, "m_pos size %zu", ht->m_pos.size());
    for (int k = 0; k < numP; k++) CHECK(ht->m_pos[4 + k].x == h_pos[k].x && ht->m_pos[4 + k].y == h_pos[k].y, "order %d", k);
    check_mesh(*ht, "host build");

    // host queries
    {
        const std::vector<int> ring = ht->oneRing(10);
        CHECK(ring.size() >= 5, "oneRing size %zu", ring.size());
        std::vector<int> fr = ht->getFRNN(10, 80.f);
        int brute = 0;
        for (int j = 4; j < (int)ht->m_pos.size(); j++)
            if (j != 10 && d2(ht->m_pos[10], ht->m_pos[j]) <= 80.f * 80.f) brute++;
        CHECK((int)fr.size() == brute, "host getFRNN %zu vs %d", fr.size(), brute);
        vec2 q(12.5f, -40.25f);
        const int t0 = ht->findContainingTriangleIndexCheckingAll(q);
        const int t1 = ht->findContainingTriangleIndexWalking(q, 0);
        CHECK(t0 >= 0 && t0 == t1, "findContaining %d vs %d", t0, t1);
    }
    // host kinetic path: move a few points a little, then delonize
    {
        HostTriangulation h2 = *ht;
        for (int i = 4; i < 4 + numP; i += 3) h2.movePoint(i, vec2(0.3f * h_move[i].x, 0.3f * h_move[i].y));
        CHECK(h2.untangle(), "host untangle");
        bool ok = h2.delonize();
        CHECK(ok, "host delonize");
        if (ok) check_mesh(h2, "host delonize");
    }

    // device
    DeviceTriangulation dt(ht);
    CHECK(dt.m_pos_size == numP + 4, "m_pos_size");
    vec2 *d_move, *d_pos;
    (void)hipMalloc(&d_move, sizeof(vec2) * (numP + 4));
    (void)hipMalloc(&d_pos, sizeof(vec2) * (numP + 4));
    (void)hipMemcpy(d_move, h_move.data(), sizeof(vec2) * (numP + 4), hipMemcpyHostToDevice);

    constexpr int K = 20, MF = 100;
    int *d_ring, *d_closest, *d_frnn;
    (void)hipMalloc(&d_ring, sizeof(int) * (numP + 4) * K);
    (void)hipMalloc(&d_closest, sizeof(int) * (numP + 4));
    (void)hipMalloc(&d_frnn, sizeof(int) * (numP + 4) * MF);

    const float rad = 60.f;
    GridCount2d gc(numP + 4, vec2(-bounds, -bounds), vec2(rad, rad),
                   ivec2((int)std::ceil(2.0 * bounds / rad), (int)std::ceil(2.0 * bounds / rad)));
    SaveNeighborsFunctor sn{rad, 64, nullptr, nullptr};
    (void)hipMalloc(&sn.num_neighbors, sizeof(int) * (numP + 4));
    (void)hipMalloc(&sn.neighbors, sizeof(int) * (numP + 4) * 64);

    for (int it = 0; it < 5; it++) {
        if (it % 2 == 0) {
            // the reference sequence: movePoints, untangle2, delonize2.  The
            // B/C steps of untangle2 need not clear every inversion (the
            // reference prints "couldn't fix everything"); the bool says so
            dt.transferToHost();
            HostTriangulation before = *ht;
            CHECK(dt.movePoints(d_move), "movePoints");
            const bool clean = dt.untangle2();
            const long long inv = dt.countInverted();
            CHECK(clean == (inv == 0), "untangle2 returned %d with %lld inverted", (int)clean, inv);
            if (inv == 0) {
                CHECK(dt.delonize2(), "delonize2");
            } else {
                // restore the valid mesh (transferToDevice) and take the move
                // through the complete kinetic step instead
                *ht = before;
                CHECK(dt.transferToDevice(), "transferToDevice");
                CHECK(dt.moveAndLegalize(d_move), "moveAndLegalize after a tangle");
            }
        } else {
            CHECK(dt.moveAndLegalize(d_move), "moveAndLegalize");
            CHECK(dt.last_stats.stalled == 0, "stalled %lld", dt.last_stats.stalled);
        }
        CHECK(dt.countInverted() == 0, "inverted at step %d", it);
        CHECK(dt.countNonDelaunay() == 0, "non-Delaunay at step %d", it);

        CHECK(dt.oneRing<K>(d_ring), "oneRing");
        CHECK(dt.closestNeighbors<K>(d_ring, d_closest), "closestNeighbors");
        CHECK((dt.getFRNN<K, MF>(rad, d_ring, d_frnn)), "getFRNN");
        dt.transferToHost();
        check_mesh(*ht, "device step");
        std::vector<int> closest(numP + 4), frnn((size_t)(numP + 4) * MF);
        (void)hipMemcpy(closest.data(), d_closest, sizeof(int) * (numP + 4), hipMemcpyDeviceToHost);
        (void)hipMemcpy(frnn.data(), d_frnn, sizeof(int) * (numP + 4) * MF, hipMemcpyDeviceToHost);
        const int nv = numP + 4;
        for (int i = 4; i < nv; i++) {
            float best = 3e38f;
            for (int j = 4; j < nv; j++)
                if (j != i) best = std::min(best, d2(ht->m_pos[i], ht->m_pos[j]));
            CHECK(d2(ht->m_pos[i], ht->m_pos[closest[i]]) == best, "closest of %d", i);
            int brute = 0;
            for (int j = 4; j < nv; j++)
                if (j != i && d2(ht->m_pos[i], ht->m_pos[j]) <= rad * rad) brute++;
            CHECK(frnn[i] == brute, "graph FRNN count of %d: %d vs %d", i, frnn[i], brute);
        }

        // grid FRNN on a copy of the positions (moving_device_mesh.cu:125-138)
        (void)hipMemcpy(d_pos, dt.m_pos, sizeof(vec2) * nv, hipMemcpyDeviceToDevice);
        (void)hipMemset(sn.num_neighbors, 0, sizeof(int) * nv);
        gc.update(d_pos);
        gc.apply_f_frnn<SaveNeighborsFunctor>(sn, d_pos, rad);
        std::vector<vec2> sorted(nv);
        std::vector<int> cnt(nv);
        (void)hipMemcpy(sorted.data(), d_pos, sizeof(vec2) * nv, hipMemcpyDeviceToHost);
        (void)hipMemcpy(cnt.data(), sn.num_neighbors, sizeof(int) * nv, hipMemcpyDeviceToHost);
        for (int i = 0; i < nv; i++) {
            int brute = 0;
            for (int j = 0; j < nv; j++)
                if (j != i && d2(sorted[i], sorted[j]) <= rad * rad) brute++;
            CHECK(cnt[i] == brute, "grid FRNN count of slot %d: %d vs %d", i, cnt[i], brute);
        }
        CHECK(gc.mean_num_particle_in_cell() > 0.f, "mean occupancy");
    }
    // built on the GPU from device points, saved and resumed
    {
        vec2* d_pts = nullptr;
        (void)hipMalloc(&d_pts, sizeof(vec2) * numP);
        (void)hipMemcpy(d_pts, h_pos.data(), sizeof(vec2) * numP, hipMemcpyHostToDevice);
        HostTriangulation hb(false), hh(false), hr(false);
        std::vector<vec2> pts = h_pos;
        CHECK(hh.addDelaunayPoints(pts), "host build for the device build");
        DeviceTriangulation db(&hb, d_pts, numP);
        CHECK(hb.m_pos.size() == hh.m_pos.size(), "device build size");
        bool same = hb.m_pos.size() == hh.m_pos.size();
        for (size_t v = 0; same && v < hb.m_pos.size(); v++)
            same = hb.m_pos[v].x == hh.m_pos[v].x && hb.m_pos[v].y == hh.m_pos[v].y;
        CHECK(same, "device build: the host build's vertex numbering");
        check_mesh(hb, "device build");
        const char* path = "dropin_state.mcl";
        CHECK(db.saveState(path), "saveState");
        DeviceTriangulation dr(&hr, path);
        bool eq = hr.m_he.size() == hb.m_he.size() && hr.m_pos.size() == hb.m_pos.size();
        for (size_t h = 0; eq && h < hr.m_he.size(); h++)
            eq = hr.m_he[h].next == hb.m_he[h].next && hr.m_he[h].v == hb.m_he[h].v && hr.m_he[h].t == hb.m_he[h].t &&
                 hr.m_he[h].op == hb.m_he[h].op;
        CHECK(eq, "resumed state equals the saved one");
        std::remove(path);
        (void)hipFree(d_pts);
    }
    // periodic grid: pairs across x = 0 and y = 0 reach the functor with the
    // minimum-image d and dist <= rad (gridCount2d.h:284-290 on the torus)
    {
        const int n = 4000;
        const float r = 0.03f;
        std::mt19937 g(21);
        std::uniform_real_distribution<float> u(0.f, 1.f);
        std::vector<vec2> p(n);
        for (int i = 0; i < n; i++) {
            p[i] = vec2(u(g), u(g));
            if (i % 4 == 0) p[i].x = (i % 8 == 0) ? 0.01f * u(g) : 1.f - 0.01f * u(g);  // near the seam x = 0
            if (i % 5 == 0) p[i].y = (i % 10 == 0) ? 0.01f * u(g) : 1.f - 0.01f * u(g);  // near y = 0
            if (p[i].x >= 1.f) p[i].x = 0.f;
            if (p[i].y >= 1.f) p[i].y = 0.f;
        }
        const int nc = (int)(1.0f / r);
        GridCount2d gt(n, vec2(0.f, 0.f), vec2(1.0f / nc, 1.0f / nc), ivec2(nc, nc), MCL_GRID_MORTON | MCL_GRID_PERIODIC);
        vec2* dp = nullptr;
        TorusProbeFunctor tp{r, nullptr, nullptr};
        (void)hipMalloc(&dp, sizeof(vec2) * n);
        (void)hipMalloc(&tp.max_bits, sizeof(int) * 3);
        (void)hipMalloc(&tp.count, sizeof(int) * n);
        (void)hipMemcpy(dp, p.data(), sizeof(vec2) * n, hipMemcpyHostToDevice);
        (void)hipMemset(tp.max_bits, 0, sizeof(int) * 3);
        (void)hipMemset(tp.count, 0, sizeof(int) * n);
        gt.update(dp);
        gt.apply_f_frnn<TorusProbeFunctor>(tp, dp, r);
        int mb[3];
        std::vector<int> cnt(n);
        std::vector<vec2> sp(n);
        (void)hipMemcpy(mb, tp.max_bits, sizeof(mb), hipMemcpyDeviceToHost);
        (void)hipMemcpy(cnt.data(), tp.count, sizeof(int) * n, hipMemcpyDeviceToHost);
        (void)hipMemcpy(sp.data(), dp, sizeof(vec2) * n, hipMemcpyDeviceToHost);
        float mx[3];
        std::memcpy(mx, mb, sizeof(mx));
        CHECK(mx[0] <= r && mx[1] <= r && mx[2] <= r, "periodic functor: max dist %g, |dx| %g, |dy| %g > rad %g",
              mx[0], mx[1], mx[2], r);
        auto wrap = [](float d) { return d > 0.5f ? d - 1.f : (d < -0.5f ? d + 1.f : d); };
        int seam_pairs = 0;
        for (int i = 0; i < n; i++) {
            int brute = 0;
            for (int j = 0; j < n; j++) {
                if (j == i) continue;
                const float dx = wrap(sp[j].x - sp[i].x), dy = wrap(sp[j].y - sp[i].y);
                if (dx * dx + dy * dy <= r * r) {
                    brute++;
                    seam_pairs += std::fabs(sp[j].x - sp[i].x) > 0.5f || std::fabs(sp[j].y - sp[i].y) > 0.5f;
                }
            }
            CHECK(cnt[i] == brute, "periodic grid FRNN count of slot %d: %d vs %d", i, cnt[i], brute);
        }
        CHECK(seam_pairs > 100, "pairs across the seams: %d", seam_pairs);
        (void)hipFree(dp);
        (void)hipFree(tp.max_bits);
        (void)hipFree(tp.count);
    }
    (void)hipFree(d_move);
    (void)hipFree(d_pos);
    (void)hipFree(d_ring);
    (void)hipFree(d_closest);
    (void)hipFree(d_frnn);
    (void)hipFree(sn.num_neighbors);
    (void)hipFree(sn.neighbors);
    if (failures) {
        std::printf("DROPIN FAILED (%d checks)\n", failures);
        return 1;
    }
    std::printf("DROPIN OK\n");
    return 0;
}
