// Device_Triangulation.h -- drop-in DeviceTriangulation (reference
// include/Device_Triangulation.h:128-223) over libmcleap_hip.so.
//
// Same constructor (borrows a HostTriangulation, allocates and uploads),
// public m_pos (device vec2*, the caller may read it in its own kernels and
// copy from it) and sizes, and the same member functions returning bool.
// The mesh lives on the device as SoA arrays owned by the library: the
// reference's device m_v / m_he / m_t AoS pointers are not exposed (no
// caller of the reference touches them); transferToHost fills the host
// object's AoS vectors.  Every call is synchronous, like the reference.
#pragma once

#include "Host_Triangulation.h"

struct DeviceTriangulation {
    using vec2 = mcleap::vec2;
    HostTriangulation* h_triangulation;
    int m_pos_size = 0, m_v_size = 0, m_he_size = 0, m_t_size = 0;
    vec2* m_pos = nullptr;  // device positions (float2), owned by the library
    int blocksize = 256;
    mcl_stats_t last_stats{};  // statistics of the last delonize/untangle/step
    int ring_overflow = 0;  // vertices with more than maxRingSize - 1 neighbours in the last oneRing / oneRing2

    explicit DeviceTriangulation(HostTriangulation* h) : h_triangulation(h) {
        mcleap::must(mcl_tri_create(reinterpret_cast<const float*>(h->m_pos.data()), (int)h->m_pos.size(),
                                    reinterpret_cast<const int*>(h->m_v.data()), (int)h->m_v.size(),
                                    reinterpret_cast<const int*>(h->m_he.data()), (int)h->m_he.size(),
                                    reinterpret_cast<const int*>(h->m_t.data()), (int)h->m_t.size(), h->periodic, &t_),
                     "DeviceTriangulation");
        mcl_tri_sizes(t_, &m_pos_size, &m_v_size, &m_he_size, &m_t_size);
        m_pos = static_cast<vec2*>(mcl_tri_pos_ptr(t_));
    }
    // The same object built on the GPU from n device points (mcl_tri_build:
    // HostTriangulation::addDelaunayPoints, Host_Triangulation.h:296-300, and
    // this class's constructor in one device pass; a torus when h->periodic).
    // h receives the arrays, with the vertex numbering the host build gives.
    DeviceTriangulation(HostTriangulation* h, const vec2* d_points, int n) : h_triangulation(h) {
        mcleap::must(mcl_tri_build(d_points, n, h->periodic ? 1 : 0, nullptr, &t_, nullptr),
                     "DeviceTriangulation(points)");
        adopt();
    }
    // Resumed from a state file written by saveState (mcl_tri_load).
    DeviceTriangulation(HostTriangulation* h, const char* state_path) : h_triangulation(h) {
        mcleap::must(mcl_tri_load(state_path, &t_), "DeviceTriangulation(state)");
        adopt();
    }
    bool saveState(const char* path) { return mcleap::ok(mcl_tri_save(t_, path), "saveState"); }
    ~DeviceTriangulation() { mcl_tri_destroy(t_); }
    DeviceTriangulation(const DeviceTriangulation&) = delete;
    DeviceTriangulation& operator=(const DeviceTriangulation&) = delete;

    // -------------------------------------------
    // transfers (Device_Triangulation.h:245-263)
    bool transferToHost() {
        HostTriangulation* h = h_triangulation;
        return mcleap::ok(mcl_tri_download(t_, reinterpret_cast<float*>(h->m_pos.data()),
                                           reinterpret_cast<int*>(h->m_v.data()), reinterpret_cast<int*>(h->m_he.data()),
                                           reinterpret_cast<int*>(h->m_t.data())),
                          "transferToHost");
    }
    bool transferToDevice() {
        HostTriangulation* h = h_triangulation;
        return mcleap::ok(mcl_tri_upload(t_, reinterpret_cast<const float*>(h->m_pos.data()),
                                         reinterpret_cast<const int*>(h->m_v.data()),
                                         reinterpret_cast<const int*>(h->m_he.data()),
                                         reinterpret_cast<const int*>(h->m_t.data())),
                          "transferToDevice");
    }
    bool allocateMemory() { return true; }  // done by the constructor
    bool freeMemory() { return true; }      // done by the destructor

    // -------------------------------------------
    // delonizing (Device_Triangulation.h:324-390)
    bool delonize() { return mcleap::ok(mcl_tri_delonize(t_, 1, 0, &last_stats, nullptr), "delonize"); }
    bool delonize2() { return mcleap::ok(mcl_tri_delonize(t_, 0, 0, &last_stats, nullptr), "delonize2"); }

    // -------------------------------------------
    // moving points (Device_Triangulation.h:395-410); d, dt, v are device arrays
    bool movePoints(vec2* d) { return mcleap::ok(mcl_tri_move(t_, d, nullptr), "movePoints"); }
    bool integrateMovePoints(float* dt, vec2* v) {
        return mcleap::ok(mcl_tri_integrate(t_, dt, v, nullptr), "integrateMovePoints");
    }
    // declared in the reference and left as stubs returning false
    // (Device_Triangulation.h:183-186, 411-417): remove and re-insert
    // p_index at its position + d (then Lawson rounds), or move it and
    // untangle (then Lawson rounds); moveVertices does the same for many
    // vertices at once (host arrays ids, d)
    bool moveFlipflop(int p_index, vec2 d) { return moveVertices(&p_index, &d, 1, MCL_MOVE_FLIPFLOP); }
    bool moveFlipflopDelaunay(int p_index, vec2 d) { return moveVertices(&p_index, &d, 1, MCL_MOVE_FLIPFLOP_DELAUNAY); }
    bool moveUntangling(int p_index, vec2 d) { return moveVertices(&p_index, &d, 1, MCL_MOVE_UNTANGLING); }
    bool moveUntanglingDelaunay(int p_index, vec2 d) {
        return moveVertices(&p_index, &d, 1, MCL_MOVE_UNTANGLING_DELAUNAY);
    }
    bool moveVertices(const int* ids, const vec2* d, int n, int mode) {
        return mcl_tri_move_vertices(t_, ids, reinterpret_cast<const float*>(d), n, mode, &last_stats, nullptr) == MCL_OK;
    }

    // -------------------------------------------
    // untangling (Device_Triangulation.h:422-495): false when inverted triangles remain
    bool untangle() { return mcl_tri_untangle(t_, 0, &last_stats, nullptr) == MCL_OK; }
    bool untangle2() { return untangle(); }

    // movePoints(d) + untangle + delonize in one call that always ends on a
    // valid, exactly Delaunay mesh (d: device displacement per vertex)
    bool moveAndLegalize(vec2* d, int variant = 0) {
        return mcleap::ok(mcl_tri_move_legalize(t_, d, variant, &last_stats, nullptr), "moveAndLegalize");
    }
    // one kinetic step of a synthetic workload: move, keep the mesh valid,
    // flip back to the exact Delaunay triangulation (the library's driver)
    bool step(const mcl_step_params_t& p) { return mcleap::ok(mcl_tri_step(t_, &p, &last_stats, nullptr), "step"); }

    // -------------------------------------------
    // neighbourhood searching (Device_Triangulation.h:499-557); all buffers
    // are caller-allocated device memory in the reference layouts
    template <int maxRingSize>
    bool oneRing(int* ring_neighbors) {
        return mcleap::ok(mcl_tri_one_ring(t_, maxRingSize, 0, ring_neighbors, &ring_overflow, nullptr), "oneRing");
    }
    template <int maxRingSize>
    bool oneRing2(int* ring_neighbors) {
        return mcleap::ok(mcl_tri_one_ring(t_, maxRingSize, 1, ring_neighbors, &ring_overflow, nullptr), "oneRing2");
    }
    template <int maxRingSize>
    bool closestNeighbors(int* ring_neighbors, int* closest_neighbors) {
        return mcleap::ok(mcl_tri_closest_ring(t_, maxRingSize, 0, ring_neighbors, closest_neighbors, nullptr),
                          "closestNeighbors");
    }
    template <int maxRingSize>
    bool closestNeighbors2(int* ring_neighbors, int* closest_neighbors) {
        return mcleap::ok(mcl_tri_closest_ring(t_, maxRingSize, 1, ring_neighbors, closest_neighbors, nullptr),
                          "closestNeighbors2");
    }
    template <int maxRingSize, int maxFRNNSize>
    bool getFRNN(float r, int* ring_neighbors, int* neighbors) {
        int ov = 0;
        return mcleap::ok(mcl_tri_frnn_graph(t_, r, maxRingSize, 0, ring_neighbors, maxFRNNSize, neighbors, &ov, nullptr),
                          "getFRNN");
    }
    template <int maxRingSize, int maxFRNNSize>
    bool getFRNN2(float r, int* ring_neighbors, int* neighbors) {
        int ov = 0;
        return mcleap::ok(mcl_tri_frnn_graph(t_, r, maxRingSize, 1, ring_neighbors, maxFRNNSize, neighbors, &ov, nullptr),
                          "getFRNN2");
    }

    bool swapVertices(int, int) { return false; }  // experimental stub in the reference

    // diagnostics used by the reference experiments (moving_device_mesh.cu:212-242)
    long long countInverted() {
        long long n = -1;
        mcl_tri_count_inverted(t_, &n, nullptr);
        return n;
    }
    long long countNonDelaunay() {
        long long n = -1;
        mcl_tri_count_non_delaunay(t_, &n, nullptr);
        return n;
    }
    mcl_tri_t* handle() { return t_; }

  private:
    mcl_tri_t* t_ = nullptr;
    // sizes, m_pos and the host object's arrays for a triangulation made by the library
    void adopt() {
        mcl_tri_sizes(t_, &m_pos_size, &m_v_size, &m_he_size, &m_t_size);
        m_pos = static_cast<vec2*>(mcl_tri_pos_ptr(t_));
        int per = 0;
        mcl_tri_is_periodic(t_, &per);
        HostTriangulation* h = h_triangulation;
        h->periodic = per != 0;
        h->m_pos.resize(m_pos_size);
        h->m_v.resize(m_v_size);
        h->m_he.resize(m_he_size);
        h->m_t.resize(m_t_size);
        transferToHost();
    }
};
