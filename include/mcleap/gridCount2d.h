// gridCount2d.h -- drop-in GridCount2d (reference include/gridCount2d.h:23-48)
// over libmcleap_hip.so.
//
// update(pos[, vel]) sorts the caller's device arrays in place by cell and
// exposes d_ind (new slot of each old index), as the reference does.
// apply_f_frnn<F>(f, pos, rad) calls the caller's device functor
//   __device__ void operator()(const int& i, const int& j, const vec2 d, const double dist)
// for every pair of sorted slots within rad (d = pos[j] - pos[i], the
// minimum image on a periodic grid; FP32 squared distance <= rad*rad),
// i == j included, like the reference.  The
// pairs come from the library's CSR query; a small adapter kernel defined
// here runs the functor, so a translation unit that uses apply_f_frnn is
// compiled with hipcc.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "Common_Triangulation.h"

namespace mcleap {
// periodic != 0: d is the minimum image on the unit torus (each component
// wrapped into [-0.5, 0.5]), as the library's query measured it
template <class Functor, class V>
__global__ void apply_f_frnn_csr_kernel(Functor f, int n, const V* __restrict__ pos, const int* __restrict__ off,
                                        const int* __restrict__ nbr, int periodic) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V pi = pos[i];
    V zero;
    zero.x = 0.f;
    zero.y = 0.f;
    f(i, i, zero, 0.0);
    for (int k = off[i]; k < off[i + 1]; k++) {
        const int j = nbr[k];
        V d;
        d.x = __fsub_rn(pos[j].x, pi.x);
        d.y = __fsub_rn(pos[j].y, pi.y);
        if (periodic) {
            d.x = d.x > 0.5f ? d.x - 1.0f : (d.x < -0.5f ? d.x + 1.0f : d.x);
            d.y = d.y > 0.5f ? d.y - 1.0f : (d.y < -0.5f ? d.y + 1.0f : d.y);
        }
        const float r2 = __fadd_rn(__fmul_rn(d.x, d.x), __fmul_rn(d.y, d.y));
        f(i, j, d, (double)sqrtf(r2));
    }
}
}  // namespace mcleap

struct GridCount2d {
    using vec2 = mcleap::vec2;
    using ivec2 = mcleap::ivec2;
    int numP;
    int blocksize = 256;
    int* d_ind = nullptr;  // new slot of old index i after update() (device)

    // flags: MCL_GRID_ROWMAJOR (the reference's cell key), MCL_GRID_MORTON,
    // optionally | MCL_GRID_PERIODIC (wrapped cells, minimum-image distances)
    GridCount2d(int numP_, vec2 min, vec2 cell_size, ivec2 num_cells, int flags = MCL_GRID_ROWMAJOR)
        : numP(numP_), periodic_((flags & MCL_GRID_PERIODIC) ? 1 : 0) {
        mcleap::must(mcl_grid_create(numP, min.x, min.y, cell_size.x, cell_size.y, num_cells.x, num_cells.y, flags, &g_),
                     "GridCount2d");
        void *ind = nullptr, *perm = nullptr, *start = nullptr, *cnt = nullptr;
        mcl_grid_index_ptrs(g_, &ind, &perm, &start, &cnt);
        d_ind = static_cast<int*>(ind);
        mcleap::must(hipMalloc(&off_, sizeof(int) * (numP + 1)) == hipSuccess ? MCL_OK : MCL_ERR_HIP, "GridCount2d");
    }
    ~GridCount2d() {
        if (nbr_) (void)hipFree(nbr_);
        if (off_) (void)hipFree(off_);
        mcl_grid_destroy(g_);
    }
    GridCount2d(const GridCount2d&) = delete;
    GridCount2d& operator=(const GridCount2d&) = delete;

    void update(vec2* pos, vec2* vel) { mcleap::ok(mcl_grid_update(g_, pos, vel, nullptr), "GridCount2d::update"); }
    void update(vec2* pos) { update(pos, nullptr); }

    template <class Functor>
    void apply_f_frnn(Functor& f, vec2* pos, const float rad) {
        long long total = 0;
        const int rc = mcl_grid_frnn_csr(g_, pos, rad, off_, nullptr, 0, &total, nullptr);  // sizing pass
        if (rc != MCL_ERR_CAPACITY && !mcleap::ok(rc, "apply_f_frnn")) return;
        if (total > cap_) {
            if (nbr_) (void)hipFree(nbr_);
            cap_ = total + total / 4 + 1024;
            if (hipMalloc(&nbr_, sizeof(int) * cap_) != hipSuccess) { nbr_ = nullptr; cap_ = 0; return; }
        }
        if (!mcleap::ok(mcl_grid_frnn_csr(g_, pos, rad, off_, nbr_, (int)cap_, &total, nullptr), "apply_f_frnn")) return;
        hipLaunchKernelGGL((mcleap::apply_f_frnn_csr_kernel<Functor, vec2>), dim3((numP + blocksize - 1) / blocksize),
                           dim3(blocksize), 0, 0, f, numP, pos, off_, nbr_, periodic_);
        (void)hipDeviceSynchronize();
    }

    float mean_num_particle_in_cell() {
        float m = 0.f;
        mcl_grid_mean_occupancy(g_, &m, nullptr);
        return m;
    }
    mcl_grid_t* handle() { return g_; }

  private:
    int periodic_ = 0;
    mcl_grid_t* g_ = nullptr;
    int* off_ = nullptr;
    int* nbr_ = nullptr;
    long long cap_ = 0;
};
