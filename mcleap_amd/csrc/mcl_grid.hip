// mcl_grid.hip -- uniform-grid fixed-radius neighbours (GridCount2d,
// gridCount2d.h:23-302) as a counting sort plus a CSR query.
//
// update():   key -> histogram (global atomics) -> exclusive scan ->
//             slot assignment (atomic order) -> each point's rank by index
//             within its cell (deterministic and stable; the reference's
//             atomic order is not) ->
//             gather into the sorted arrays -> copy back over the caller's
//             buffers (the reference sorts the caller's array in place,
//             gridCount2d.h:207-208, 256).
// frnn_csr(): per sorted slot i, the 3x3 cells around i's cell; pairs with
//             r2 = dx*dx + dy*dy <= rad*rad in FP32 without contraction
//             (gridCount2d.h:284-290), j != i.  Two passes (count, scan,
//             fill), so the CSR is deterministic.
// Cell keys: MCL_GRID_ROWMAJOR reproduces calcHash / calc_hashxy
// (gridCount2d.h:89-108) including the unsigned wrap and the query's
// `hh.x > nx` bound (gridCount2d.h:278-281); MCL_GRID_MORTON orders cells
// along a Morton curve (cells of a tile are contiguous);
// MCL_GRID_PERIODIC wraps cells and distances on the unit torus.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <vector>

#include "mcl_geom.hpp"
#include "mcl_internal.hpp"
#include "mcl_scan.hpp"

using namespace mcl;

namespace {

constexpr int kT = 256;

struct GridParams {
    float minx, miny, csx, csy;
    int nx, ny, tot, nkeys;
    int morton, periodic;
    int block = 256;  // threads per block of this grid's launches (64, 128 or 256)
    int reach = 1;    // cells searched on each side of a point's cell (set per query)
};

__host__ __device__ __forceinline__ uint32_t spread16(uint32_t x) {
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// calc_hashxy (gridCount2d.h:94-102): FP32 subtract, divide, truncate, C %
__device__ __forceinline__ int2 cell_of(const GridParams& g, float2 p) {
    float qx = __fsub_rn(p.x, g.minx), qy = __fsub_rn(p.y, g.miny);
    int x = (int)__fdiv_rn(qx, g.csx), y = (int)__fdiv_rn(qy, g.csy);
    // C remainder (a no-op inside the grid: the division runs only outside it)
    if ((unsigned)x >= (unsigned)g.nx) x %= g.nx;
    if ((unsigned)y >= (unsigned)g.ny) y %= g.ny;
    if (g.periodic || g.morton) {
        if (g.periodic) {
            // (x + nx) % nx for x in (-nx, nx)
            x += x < 0 ? g.nx : 0;
            y += y < 0 ? g.ny : 0;
        } else {
            x = min(max(x, 0), g.nx - 1);
            y = min(max(y, 0), g.ny - 1);
        }
    }
    return make_int2(x, y);
}
__device__ __forceinline__ int key_of(const GridParams& g, int x, int y) {
    if (g.morton) return (int)(spread16((uint32_t)x) | (spread16((uint32_t)y) << 1));
    return (int)(((unsigned)y * (unsigned)g.nx + (unsigned)x) % (unsigned)g.tot);  // calcHash
}

__global__ void k_zero(int* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}
// two arrays of n ints each, one launch
__global__ void k_zero2(int* __restrict__ p, int* __restrict__ q, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        p[i] = 0;
        q[i] = 0;
    }
}

// key and histogram; with slot given, each point's place among its cell's
// points in atomic (arbitrary) order, the count before its own increment
__global__ void __launch_bounds__(kT) k_key_count(GridParams g, const float2* __restrict__ pos, int n,
                                                  int* __restrict__ key, int* __restrict__ count,
                                                  int* __restrict__ slot) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int2 c = cell_of(g, pos[i]);
        int k = key_of(g, c.x, c.y);
        key[i] = k;
        if (slot) slot[i] = atomicAdd(&count[k], 1);
        else atomicAdd(&count[k], 1);
    }
}

// each point into a slot of its cell, in atomic (arbitrary) order: the place
// its key pass took (slot given), else a second round of atomics on fill
__global__ void __launch_bounds__(kT) k_slot(const int* __restrict__ key, int n, const int* __restrict__ start,
                                             int* __restrict__ fill, const int* __restrict__ slot,
                                             int* __restrict__ perm) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int k = key[i];
        perm[start[k] + (slot ? slot[i] : atomicAdd(&fill[k], 1))] = i;
    }
}

// the stable order within each cell: the point in slot s of the arbitrary
// order goes to its cell's start plus the number of the cell's points with a
// smaller index (one thread per point; a cell's slots are contiguous and
// read by all of its threads, so they stay in cache)
__global__ void __launch_bounds__(kT) k_cell_rank(const int* __restrict__ key, int n, const int* __restrict__ start,
                                                  const int* __restrict__ perm_any, int* __restrict__ perm) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const int i = perm_any[s];
        const int k = key[i];
        const int b = start[k], e = start[k + 1];
        int r = 0;
        for (int j = b; j < e; j++) r += perm_any[j] < i;
        perm[b + r] = i;
    }
}

__global__ void __launch_bounds__(kT) k_gather(const int* __restrict__ perm, int n, const float2* __restrict__ pos,
                                               const float2* __restrict__ vel, float2* __restrict__ spos,
                                               float2* __restrict__ svel, int* __restrict__ ind) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        int i = perm[s];
        spos[s] = pos[i];
        if (vel) svel[s] = vel[i];
        ind[i] = s;
    }
}

// pass 0 counts, pass 1 writes.  The rows of a wave's 64 consecutive points
// are one contiguous range of the output; in pass 1 a wave whose range fits
// kWaveRow ints gathers it in LDS and writes it out with contiguous stores
// (each thread storing its own short row directly writes partial cache lines,
// several times the bytes of the list); a longer range is written directly.
// With hits given, pass 0 also records which candidates (the points of the
// searched cells other than i, in scan order) are within the radius: bit k
// for the k-th, bit 63 set when there are more than 63 (then pass 1 computes
// the distances again); pass 1 then writes the recorded ones without reading
// any position.
constexpr int kWaveRow = 1024;
constexpr unsigned long long kHitsOverflow = 1ull << 63;
template <int kPass>
__global__ void __launch_bounds__(kT) k_frnn(GridParams g, const float2* __restrict__ pos, int n,
                                             const int* __restrict__ start, float rad2, int* __restrict__ cnt,
                                             const int* __restrict__ off, int* __restrict__ nbr, int cap,
                                             unsigned long long* __restrict__ hits) {
    __shared__ int s_row[kPass ? kT / 64 : 1][kPass ? kWaveRow : 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // each wave stages its own rows (wave-level syncs); each XCD takes one
    // contiguous range of the sorted points
    for (int bb = xcd_block() * blockDim.x; bb < n; bb += gridDim.x * blockDim.x) {
        const int i = bb + threadIdx.x;
        const bool act = i < n;
        const int wb = bb + (threadIdx.x & ~63);  // this wave's first point
        int w0 = 0, wn = 0;
        if (kPass && wb < n) {
            w0 = off[wb];
            wn = off[min(wb + 64, n)] - w0;
        }
        const bool staged = kPass && wn <= kWaveRow;
        unsigned long long hm = 0;  // pass 0: hits so far; pass 1: hits left to write
        bool from_hits = false;
        if (kPass && hits && act) {
            hm = hits[i];
            from_hits = !(hm & kHitsOverflow);
        }
        const float2 pi = act ? pos[i] : make_float2(0.0f, 0.0f);
        const int2 h = cell_of(g, pi);
        int c = 0, w = (kPass && act) ? off[i] : 0, nb = 0;
        const int R = g.reach;
        for (int ddx = -R; act && ddx <= R; ddx++) {
            for (int ddy = -R; ddy <= R; ddy++) {
                if (from_hits && hm == 0) break;
                int x = h.x + ddx, y = h.y + ddy;
                if (g.periodic) {
                    // wrapped without a division: |ddx|, |ddy| <= reach, and
                    // 2 * reach + 1 <= nx, ny (checked at the launch)
                    x += x < 0 ? g.nx : 0;
                    x -= x >= g.nx ? g.nx : 0;
                    y += y < 0 ? g.ny : 0;
                    y -= y >= g.ny ? g.ny : 0;
                } else if (g.morton) {
                    if (x < 0 || x >= g.nx || y < 0 || y >= g.ny) continue;
                } else {
                    if (x > g.nx || x < 0) continue;  // gridCount2d.h:278-281 bound test
                    if (y > g.ny || y < 0) continue;
                }
                const int k = key_of(g, x, y);
                const int b = start[k], e = start[k + 1];
                if (from_hits) {
                    for (int j = b; j < e && hm; j++) {
                        if (j == i) continue;
                        if (hm & 1ull) {
                            if (staged) s_row[wv][w - w0] = j;
                            else if (w < cap) nbr[w] = j;
                            w++;
                        }
                        hm >>= 1;
                    }
                    continue;
                }
                for (int j = b; j < e; j++) {
                    if (j == i) continue;
                    const float2 pj = pos[j];
                    float dx = __fsub_rn(pj.x, pi.x), dy = __fsub_rn(pj.y, pi.y);
                    if (g.periodic) { dx = min_image_f(dx); dy = min_image_f(dy); }
                    const float r2 = __fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy));
                    const bool hit = r2 <= rad2;
                    if (!kPass && hit && nb < 63) hm |= 1ull << nb;
                    nb++;
                    if (hit) {
                        if (kPass) {
                            if (staged) s_row[wv][w - w0] = j;
                            else if (w < cap) nbr[w] = j;
                        }
                        w++;
                        c++;
                    }
                }
            }
        }
        if (!kPass && act) {
            cnt[i] = c;
            if (hits) hits[i] = nb <= 63 ? hm : kHitsOverflow;
        }
        if (kPass) {
            wave_sync();
            if (staged)
                for (int k = lane; k < wn; k += 64)
                    if (w0 + k < cap) nbr[w0 + k] = s_row[wv][k];
            wave_sync();
        }
    }
}

__global__ void k_occupancy(const int* __restrict__ start, int nkeys, unsigned long long* out) {
    unsigned long long nonempty = 0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nkeys; c += gridDim.x * blockDim.x)
        nonempty += (start[c + 1] > start[c]);
    if (nonempty) atomicAdd(out, nonempty);
}

int blocks(long long n, int block = kT) {
    long long b = (n + block - 1) / block;
    return (int)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

// The buffers of one counting sort by cell.  key, perm, ind: n + 1 ints;
// start, fill: nkeys + 1 ints; scan_tmp: scan_scratch_ints(max(nkeys, n) + 1)
// ints; spos (and svel when vel is given): n float2.
struct SortBufs {
    int *key, *start, *fill, *perm, *ind, *scan_tmp;
    float2 *spos, *svel;
};

// Counting sort of n points by cell: key + histogram (with each point's place
// in its cell in atomic order), exclusive scan of the histogram, slots in
// that order (into ind), the stable order within each cell (into perm), then
// the gather of the sorted positions and ind[i] = the slot of point i.
// Asynchronous on s; uses only the given buffers.
int grid_sort(const GridParams& g, const float2* pos, const float2* vel, int n, const SortBufs& b, hipStream_t s) {
    const int nk = g.nkeys, B = g.block;
    // each point's place in its cell comes from the histogram's own atomics,
    // kept in the sorted positions' buffer until the gather fills it (unless
    // that buffer is the input: then a second round of atomics)
    int* slot = (const void*)pos == (const void*)b.spos ? nullptr : reinterpret_cast<int*>(b.spos);
    if (slot) hipLaunchKernelGGL(k_zero, dim3(blocks(nk + 1)), dim3(kT), 0, s, b.start, nk + 1);
    else hipLaunchKernelGGL(k_zero2, dim3(blocks(nk + 1)), dim3(kT), 0, s, b.start, b.fill, nk + 1);
    hipLaunchKernelGGL(k_key_count, dim3(blocks(n, B)), dim3(B), 0, s, g, pos, n, b.key, b.start, slot);
    MCL_TRY(scan_exclusive(b.start, b.start, nk, b.scan_tmp, s));
    hipLaunchKernelGGL(k_slot, dim3(blocks(n, B)), dim3(B), 0, s, b.key, n, b.start, b.fill, (const int*)slot, b.ind);
    hipLaunchKernelGGL(k_cell_rank, dim3(blocks(n, B)), dim3(B), 0, s, b.key, n, b.start, (const int*)b.ind, b.perm);
    hipLaunchKernelGGL(k_gather, dim3(blocks(n, B)), dim3(B), 0, s, b.perm, n, pos, vel, b.spos, b.svel, b.ind);
    MCL_HIP(hipGetLastError());
    return MCL_OK;
}

// the count pass, its scan and (with d_nbrs) the fill pass of the CSR query
// over sorted positions spos; total == NULL: no host read, writes past cap
// dropped; else *total = the pair count after a synchronisation
int frnn_csr(const GridParams& g0, const float2* spos, int n, const int* start, float rad, int* off, int* nbrs,
             int cap, long long* total, int* scan_tmp, hipStream_t s, KTimer* tc, KTimer* tf,
             unsigned long long* hits = nullptr) {
    const float rad2 = rad * rad;
    // cells searched on each side: one for the reference's row-major grid
    // (gridCount2d.h:272-282, whatever the radius); with the Morton or
    // periodic layouts as many as a radius longer than a cell needs
    GridParams g = g0;
    g.reach = 1;
    if (g.morton || g.periodic) {
        const double c = std::min(g.csx, g.csy);
        g.reach = std::max(1, (int)std::ceil((double)rad / c * (1.0 - 1e-6)));
        if (g.periodic && 2 * g.reach + 1 > std::min(g.nx, g.ny)) return MCL_ERR_INPUT;
    }
    const int B = g.block;
    if (tc) tc->mark(s);
    hipLaunchKernelGGL(k_frnn<0>, dim3(blocks(n, B)), dim3(B), 0, s, g, spos, n, start, rad2, off, (const int*)nullptr,
                       (int*)nullptr, 0, hits);
    if (tc) tc->mark(s);
    MCL_TRY(scan_exclusive(off, off, n, scan_tmp, s));
    if (!total) {
        if (tf) tf->mark(s);
        if (nbrs && cap > 0)
            hipLaunchKernelGGL(k_frnn<1>, dim3(blocks(n, B)), dim3(B), 0, s, g, spos, n, start, rad2, (int*)nullptr,
                               (const int*)off, nbrs, cap, hits);
        if (tf) tf->mark(s);
        MCL_HIP(hipGetLastError());
        return MCL_OK;
    }
    int tot = 0;
    MCL_HIP(hipMemcpyAsync(&tot, off + n, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    *total = tot;
    if (nbrs && cap > 0) {
        hipLaunchKernelGGL(k_frnn<1>, dim3(blocks(n, B)), dim3(B), 0, s, g, spos, n, start, rad2, (int*)nullptr,
                           (const int*)off, nbrs, cap, hits);
        MCL_HIP(hipGetLastError());
        MCL_HIP(hipStreamSynchronize(s));
    }
    return tot > cap ? MCL_ERR_CAPACITY : MCL_OK;
}

// grid geometry from the mcl_grid_create arguments (MCL_ERR_INPUT if invalid)
int grid_params(float minx, float miny, float csx, float csy, int ncx, int ncy, int flags, GridParams* g) {
    if (ncx <= 0 || ncy <= 0 || !(csx > 0) || !(csy > 0)) return MCL_ERR_INPUT;
    const int morton = (flags & 1) ? 1 : 0, periodic = (flags & MCL_GRID_PERIODIC) ? 1 : 0;
    if (morton && (ncx > 65536 || ncy > 65536)) return MCL_ERR_INPUT;
    if (periodic && (ncx < 3 || ncy < 3)) return MCL_ERR_INPUT;
    if ((long long)ncx * ncy >= (1LL << 31)) return MCL_ERR_INPUT;
    *g = {minx, miny, csx, csy, ncx, ncy, ncx * ncy, ncx * ncy, morton, periodic};
    if (morton) {
        long long side = 1;
        while (side < ncx || side < ncy) side <<= 1;
        if (side * side >= (1LL << 31)) return MCL_ERR_INPUT;
        g->nkeys = (int)(side * side);
    }
    return MCL_OK;
}

}  // namespace

struct mcl_grid {
    GridParams g;
    int nump = 0;
    int* key = nullptr;
    int* start = nullptr;  // nkeys + 1 (counts, then exclusive offsets)
    int* fill = nullptr;
    int* perm = nullptr;
    int* ind = nullptr;
    int* cnt = nullptr;    // nump + 1
    int* scan_tmp = nullptr;
    float2* spos = nullptr;
    float2* svel = nullptr;
    unsigned long long* scalar = nullptr;
    unsigned long long* hits = nullptr;  // nump: the FRNN count pass's hit masks, read by its fill pass
    mcl::KTimer timers[3];  // FRNN count, FRNN fill, build
    SortBufs bufs() const { return {key, start, fill, perm, ind, scan_tmp, spos, svel}; }
};

static void grid_free(mcl_grid* G) {
    void* p[] = {G->key, G->start, G->fill, G->perm, G->ind, G->cnt, G->scan_tmp, G->spos, G->svel, G->scalar,
                 G->hits};
    for (void* q : p)
        if (q) (void)hipFree(q);
}

extern "C" int mcl_grid_create(int nump, float minx, float miny, float csx, float csy, int ncx, int ncy, int flags,
                               mcl_grid_t** out) {
    if (!out || nump < 0) return MCL_ERR_INPUT;
    GridParams gp;
    MCL_TRY(grid_params(minx, miny, csx, csy, ncx, ncy, flags, &gp));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MCL_ERR_NO_DEVICE;
    mcl_grid* G = new mcl_grid();
    G->nump = nump;
    G->g = gp;
    const int nk = G->g.nkeys;
    bool ok = hipMalloc(&G->key, sizeof(int) * (nump + 1)) == hipSuccess &&
              hipMalloc(&G->start, sizeof(int) * (nk + 1)) == hipSuccess &&
              hipMalloc(&G->fill, sizeof(int) * (nk + 1)) == hipSuccess &&
              hipMalloc(&G->perm, sizeof(int) * (nump + 1)) == hipSuccess &&
              hipMalloc(&G->ind, sizeof(int) * (nump + 1)) == hipSuccess &&
              hipMalloc(&G->cnt, sizeof(int) * (nump + 1)) == hipSuccess &&
              hipMalloc(&G->scan_tmp, sizeof(int) * (scan_scratch_ints((nk > nump ? nk : nump) + 1) + 16)) == hipSuccess &&
              hipMalloc(&G->spos, sizeof(float2) * (nump + 1)) == hipSuccess &&
              hipMalloc(&G->svel, sizeof(float2) * (nump + 1)) == hipSuccess &&
              hipMalloc(&G->scalar, sizeof(unsigned long long) * 2) == hipSuccess &&
              hipMalloc(&G->hits, sizeof(unsigned long long) * (nump + 1)) == hipSuccess;
    if (!ok) { grid_free(G); delete G; return MCL_ERR_HIP; }
    hipLaunchKernelGGL(k_zero, dim3(blocks(nk + 1)), dim3(kT), 0, 0, G->start, nk + 1);
    if (hipDeviceSynchronize() != hipSuccess) { grid_free(G); delete G; return MCL_ERR_HIP; }
    tune_grid(G, G->g.periodic, nump);
    *out = G;
    return MCL_OK;
}

extern "C" int mcl_grid_destroy(mcl_grid_t* G) {
    if (!G) return MCL_OK;
    (void)hipDeviceSynchronize();
    grid_free(G);
    delete G;
    return MCL_OK;
}

extern "C" int mcl_grid_update(mcl_grid_t* G, void* d_pos, void* d_vel, void* stream) {
    if (!G || !d_pos) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    const int n = G->nump;
    float2* pos = (float2*)d_pos;
    float2* vel = (float2*)d_vel;
    MCL_TRY(grid_sort(G->g, pos, vel, n, G->bufs(), s));
    MCL_HIP(hipMemcpyAsync(pos, G->spos, sizeof(float2) * n, hipMemcpyDeviceToDevice, s));
    if (vel) MCL_HIP(hipMemcpyAsync(vel, G->svel, sizeof(float2) * n, hipMemcpyDeviceToDevice, s));
    MCL_HIP(hipStreamSynchronize(s));
    return MCL_OK;
}

// same counting sort, but the source stays untouched: the sorted copy lives
// in the grid (mcl_grid_sorted_ptr).  Asynchronous on `stream`.
extern "C" int mcl_grid_build(mcl_grid_t* G, const void* d_src, void* stream) {
    if (!G || !d_src) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    const float2* pos = (const float2*)d_src;
    G->timers[2].mark(s);
    MCL_TRY(grid_sort(G->g, pos, nullptr, G->nump, G->bufs(), s));
    G->timers[2].mark(s);
    return MCL_OK;
}

extern "C" void* mcl_grid_sorted_ptr(mcl_grid_t* G) { return G ? (void*)G->spos : nullptr; }

extern "C" int mcl_grid_index_ptrs(mcl_grid_t* G, void** d_ind, void** d_perm, void** d_cell_start,
                                   void** d_cell_count) {
    if (!G) return MCL_ERR_INPUT;
    if (d_ind) *d_ind = G->ind;
    if (d_perm) *d_perm = G->perm;
    if (d_cell_start) *d_cell_start = G->start;
    if (d_cell_count) *d_cell_count = nullptr;  // counts are start[c+1] - start[c]
    return MCL_OK;
}

extern "C" int mcl_grid_frnn_csr(mcl_grid_t* G, const void* d_pos, float rad, void* d_offsets, void* d_nbrs, int cap,
                                 long long* total, void* stream) {
    if (!G || !d_pos || !d_offsets) return MCL_ERR_INPUT;
    return frnn_csr(G->g, (const float2*)d_pos, G->nump, G->start, rad, (int*)d_offsets, (int*)d_nbrs, cap, total,
                    G->scan_tmp, (hipStream_t)stream, &G->timers[0], &G->timers[1], G->hits);
}

// the counting sort of pos into the grid and the CSR query over the sorted
// copy, asynchronous and without host reads (writes past cap dropped), as a
// timestep's tail enqueues (or captures) it
int mcl::grid_enqueue_frnn(mcl_grid_t* G, const float2* pos, float rad, int* off, int* nbrs, int cap,
                           hipStream_t s) {
    if (!G || !pos || !off) return MCL_ERR_INPUT;
    G->timers[2].mark(s);
    MCL_TRY(grid_sort(G->g, pos, nullptr, G->nump, G->bufs(), s));
    G->timers[2].mark(s);
    return frnn_csr(G->g, G->spos, G->nump, G->start, rad, off, nbrs, cap, nullptr, G->scan_tmp, s, &G->timers[0],
                    &G->timers[1], G->hits);
}

// ------------------------------------------------------------------------
// The grid FRNN without any allocation (mcl_grid_frnn_csr_noalloc): the
// counting sort's buffers live in `scratch`, a file-scope device array of
// fixed size that every GPU holds from module load until the process ends.
// Its device address is obtained once per device with hipGetSymbolAddress
// (the array's name in host code is not a device address).  All calls on one
// GPU share the array: calls that overlap on one GPU (concurrent streams or
// host threads) race; each call clears what must start at zero.
namespace {
constexpr int kNaMaxPoints = 1 << 20;  // points per call
constexpr int kNaMaxKeys = 1 << 20;    // cell keys (cells, or the Morton square)
constexpr int kNaScanInts = (((kNaMaxPoints > kNaMaxKeys ? kNaMaxPoints : kNaMaxKeys) + 1 + 1023) / 1024) + 1 + 16;
constexpr size_t kNaInts = 3 * (size_t)(kNaMaxPoints + 1)       // key, perm, ind
                           + 2 * (size_t)(kNaMaxKeys + 1)       // start, fill
                           + (size_t)kNaScanInts                // scan partials
                           + 2 * (size_t)(kNaMaxPoints + 64);   // sorted positions (float2), 256-byte aligned
constexpr int kNaMaxDevices = 64;
}  // namespace

static __device__ __align__(256) int scratch[kNaInts];

static std::mutex g_na_mu;
static int* g_na_ptr[kNaMaxDevices];

// scratch's device address on the current device (once per device)
static int na_scratch(int** out) {
    int dev = -1;
    MCL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) return MCL_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "mcleap: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return MCL_ERR_HIP;
        }
        g_na_ptr[dev] = (int*)p;
    }
    *out = g_na_ptr[dev];
    return MCL_OK;
}

extern "C" int mcl_grid_frnn_csr_noalloc(const void* d_pos, int nump, float minx, float miny, float csx, float csy,
                                         int ncx, int ncy, int flags, float rad, void* d_perm, void* d_offsets,
                                         void* d_nbrs, int cap, long long* total, void* stream) {
    if (!d_pos || !d_offsets || nump < 0) return MCL_ERR_INPUT;
    GridParams g;
    MCL_TRY(grid_params(minx, miny, csx, csy, ncx, ncy, flags, &g));
    if (nump > kNaMaxPoints || g.nkeys > kNaMaxKeys) {
        fprintf(stderr,
                "mcleap: mcl_grid_frnn_csr_noalloc takes at most %d points and %d cell keys (got %d points, %d keys)\n",
                kNaMaxPoints, kNaMaxKeys, nump, g.nkeys);
        return MCL_ERR_CAPACITY;
    }
    int* base = nullptr;
    MCL_TRY(na_scratch(&base));
    hipStream_t s = (hipStream_t)stream;
    SortBufs b;
    int* q = base;
    b.key = q; q += kNaMaxPoints + 1;
    b.perm = q; q += kNaMaxPoints + 1;
    b.ind = q; q += kNaMaxPoints + 1;
    b.start = q; q += kNaMaxKeys + 1;
    b.fill = q; q += kNaMaxKeys + 1;
    b.scan_tmp = q; q += kNaScanInts;
    q = (int*)(((uintptr_t)q + 255) & ~(uintptr_t)255);
    b.spos = (float2*)q;
    b.svel = nullptr;
    if (nump == 0) {
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, (int*)d_offsets, 1);
        MCL_HIP(hipGetLastError());
        if (total) {
            MCL_HIP(hipStreamSynchronize(s));
            *total = 0;
        }
        return MCL_OK;
    }
    MCL_TRY(grid_sort(g, (const float2*)d_pos, nullptr, nump, b, s));
    if (d_perm) MCL_HIP(hipMemcpyAsync(d_perm, b.perm, sizeof(int) * nump, hipMemcpyDeviceToDevice, s));
    return frnn_csr(g, b.spos, nump, b.start, rad, (int*)d_offsets, (int*)d_nbrs, cap, total, b.scan_tmp, s, nullptr,
                    nullptr);
}

extern "C" int mcl_grid_set_option(mcl_grid_t* G, int option, long long value) {
    if (!G) return MCL_ERR_INPUT;
    if (option == MCL_GRID_OPT_BLOCK) {
        if (value != 64 && value != 128 && value != 256) return MCL_ERR_INPUT;
        G->g.block = (int)value;
        return MCL_OK;
    }
    return MCL_ERR_INPUT;
}

extern "C" int mcl_grid_get_option(const mcl_grid_t* G, int option, long long* value) {
    if (!G || !value) return MCL_ERR_INPUT;
    if (option == MCL_GRID_OPT_BLOCK) {
        *value = G->g.block;
        return MCL_OK;
    }
    return MCL_ERR_INPUT;
}

extern "C" int mcl_grid_mean_occupancy(mcl_grid_t* G, float* out, void* stream) {
    if (!G || !out) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, (int*)G->scalar, 2);
    hipLaunchKernelGGL(k_occupancy, dim3(blocks(G->g.nkeys)), dim3(kT), 0, s, G->start, G->g.nkeys, G->scalar);
    MCL_HIP(hipGetLastError());
    unsigned long long ne = 0;
    MCL_HIP(hipMemcpyAsync(&ne, G->scalar, sizeof(ne), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    *out = ne ? (float)G->nump / (float)ne : 0.0f;
    return MCL_OK;
}

extern "C" int mcl_grid_kernel_timer(mcl_grid_t* G, int enable) {
    if (!G) return MCL_ERR_INPUT;
    for (auto& k : G->timers) k.reset(enable != 0);
    return MCL_OK;
}
extern "C" int mcl_grid_kernel_timer_read(mcl_grid_t* G, int which, double* total_ms, long long* launches) {
    if (!G || which < 0 || which > 2) return MCL_ERR_INPUT;
    double ms = G->timers[which].total_ms(launches);
    if (total_ms) *total_ms = ms;
    return MCL_OK;
}

// ------------------------------------------------------------------------
// Batched grids: one uniform grid per instance of an ensemble held as one
// point array (instance i: points vstart[i]..vstart[i+1]-1, all in [0,1)^2).
// Instance i has nc_i = max(3, floor(1/r_i)) cells per side of side
// cs_i = 1/nc_i, cells clamped at the border (as MCL_GRID_MORTON), and its
// cells are keys cbase[i] .. cbase[i] + nc_i^2 - 1 in row-major order, so
// one counting sort orders every instance's points by (instance, cell).
// The query is the reference's per-grid test (gridCount2d.h:284-290) with
// the instance's own radius: j != i in the 3x3 cells of i's instance grid,
// r2 = dx*dx + dy*dy <= r_i*r_i in FP32 without contraction.
namespace {

struct BGridParams {
    const int* vstart;  // ninst + 1
    const int* nc;      // ninst
    const float* cs;    // ninst
    const float* rad2;  // ninst
    const int* cbase;   // ninst + 1
    int ninst, v0, n;   // points v0 .. v0 + n - 1
};

// the instance holding point v: the largest i with vstart[i] <= v
__device__ __forceinline__ int bgrid_inst(const BGridParams& b, int v) {
    int lo = 0, hi = b.ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b.vstart[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int2 bgrid_cell(float cs, int nc, float2 p) {
    int x = (int)__fdiv_rn(p.x, cs), y = (int)__fdiv_rn(p.y, cs);
    x = min(max(x, 0), nc - 1);
    y = min(max(y, 0), nc - 1);
    return make_int2(x, y);
}

__global__ void __launch_bounds__(kT) k_bkey_count(BGridParams b, const float2* __restrict__ pos,
                                                   int* __restrict__ key, int* __restrict__ count,
                                                   int* __restrict__ slot) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x) {
        const int v = b.v0 + i;
        const int t = bgrid_inst(b, v);
        const int nc = b.nc[t];
        const int2 c = bgrid_cell(b.cs[t], nc, pos[v]);
        const int k = b.cbase[t] + c.y * nc + c.x;
        key[i] = k;
        if (slot) slot[i] = atomicAdd(&count[k], 1);
        else atomicAdd(&count[k], 1);
    }
}

// pass 0 counts, pass 1 writes (sorted slots; perm[s] = point index - v0);
// pass 1 stages a wave's rows in LDS when they fit (as k_frnn)
template <int kPass>
__global__ void __launch_bounds__(kT) k_bfrnn(BGridParams b, const float2* __restrict__ spos,
                                              const int* __restrict__ perm, const int* __restrict__ start,
                                              int* __restrict__ cnt, const int* __restrict__ off, int* __restrict__ nbr,
                                              int cap) {
    __shared__ int s_row[kPass ? kT / 64 : 1][kPass ? kWaveRow : 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // each wave stages its own rows (wave-level syncs); each XCD takes one
    // contiguous range of the sorted points
    for (int bb = xcd_block() * blockDim.x; bb < b.n; bb += gridDim.x * blockDim.x) {
        const int i = bb + threadIdx.x;
        const bool act = i < b.n;
        const int wb = bb + (threadIdx.x & ~63);
        int w0 = 0, wn = 0;
        if (kPass && wb < b.n) {
            w0 = off[wb];
            wn = off[min(wb + 64, b.n)] - w0;
        }
        const bool staged = kPass && wn <= kWaveRow;
        int c = 0, w = (kPass && act) ? off[i] : 0;
        if (act) {
            const float2 pi = spos[i];
            const int t = bgrid_inst(b, b.v0 + perm[i]);
            const int nc = b.nc[t], cb = b.cbase[t];
            const float rad2 = b.rad2[t];
            const int2 h = bgrid_cell(b.cs[t], nc, pi);
            // keys are row-major within an instance, so the up to 3 cells of one
            // row are one contiguous range of the sorted points (same order as
            // cell by cell)
            const int x0 = max(h.x - 1, 0), x1 = min(h.x + 1, nc - 1);
            for (int y = max(h.y - 1, 0); y <= min(h.y + 1, nc - 1); y++) {
                const int k = cb + y * nc;
                const int e = start[k + x1 + 1];
                for (int j = start[k + x0]; j < e; j++) {
                    if (j == i) continue;
                    const float2 pj = spos[j];
                    const float dx = __fsub_rn(pj.x, pi.x), dy = __fsub_rn(pj.y, pi.y);
                    const float r2 = __fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy));
                    if (r2 <= rad2) {
                        if (kPass) {
                            if (staged) s_row[wv][w - w0] = j;
                            else if (w < cap) nbr[w] = j;
                        }
                        w++;
                        c++;
                    }
                }
            }
        }
        if (!kPass && act) cnt[i] = c;
        if (kPass) {
            wave_sync();
            if (staged)
                for (int k = lane; k < wn; k += 64)
                    if (w0 + k < cap) nbr[w0 + k] = s_row[wv][k];
            wave_sync();
        }
    }
}

}  // namespace

struct mcl_bgrid {
    int ninst = 0, v0 = 0, n = 0, nkeys = 0;
    int *vstart = nullptr, *nc = nullptr, *cbase = nullptr;
    float *cs = nullptr, *rad2 = nullptr;
    int *key = nullptr, *start = nullptr, *fill = nullptr, *perm = nullptr, *ind = nullptr, *scan_tmp = nullptr;
    float2* spos = nullptr;
    mcl::KTimer timers[3];  // FRNN count, FRNN fill, build
    BGridParams params() const { return {vstart, nc, cs, rad2, cbase, ninst, v0, n}; }
};

static void bgrid_free(mcl_bgrid* G) {
    void* p[] = {G->vstart, G->nc, G->cbase, G->cs, G->rad2, G->key, G->start, G->fill, G->perm, G->ind,
                 G->scan_tmp, G->spos};
    for (void* q : p)
        if (q) (void)hipFree(q);
}

extern "C" int mcl_bgrid_create(int ninst, const int* vstart, const float* radius, mcl_bgrid_t** out) {
    if (!out || ninst <= 0 || !vstart || !radius) return MCL_ERR_INPUT;
    std::vector<int> nc(ninst), cbase(ninst + 1, 0);
    std::vector<float> cs(ninst), rad2(ninst);
    long long keys = 0;
    for (int i = 0; i < ninst; i++) {
        if (vstart[i + 1] < vstart[i] || vstart[i] < 0 || !(radius[i] > 0.0f) || !(radius[i] < 1e30f))
            return MCL_ERR_INPUT;
        const double k = std::floor(1.0 / (double)radius[i]);
        nc[i] = (int)std::max(3.0, std::min(k, 4096.0));
        cs[i] = 1.0f / (float)nc[i];
        if (!(cs[i] >= radius[i])) return MCL_ERR_INPUT;  // a 3x3 query must cover the radius
        rad2[i] = radius[i] * radius[i];
        cbase[i] = (int)keys;
        keys += (long long)nc[i] * nc[i];
        if (keys >= (1LL << 30)) return MCL_ERR_CAPACITY;
    }
    cbase[ninst] = (int)keys;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MCL_ERR_NO_DEVICE;
    mcl_bgrid* G = new mcl_bgrid();
    G->ninst = ninst;
    G->v0 = vstart[0];
    G->n = vstart[ninst] - vstart[0];
    G->nkeys = (int)keys;
    const int n = G->n, nk = G->nkeys;
    bool ok = hipMalloc(&G->vstart, sizeof(int) * (ninst + 1)) == hipSuccess &&
              hipMalloc(&G->nc, sizeof(int) * ninst) == hipSuccess &&
              hipMalloc(&G->cbase, sizeof(int) * (ninst + 1)) == hipSuccess &&
              hipMalloc(&G->cs, sizeof(float) * ninst) == hipSuccess &&
              hipMalloc(&G->rad2, sizeof(float) * ninst) == hipSuccess &&
              hipMalloc(&G->key, sizeof(int) * (n + 1)) == hipSuccess &&
              hipMalloc(&G->start, sizeof(int) * (nk + 1)) == hipSuccess &&
              hipMalloc(&G->fill, sizeof(int) * (nk + 1)) == hipSuccess &&
              hipMalloc(&G->perm, sizeof(int) * (n + 1)) == hipSuccess &&
              hipMalloc(&G->ind, sizeof(int) * (n + 1)) == hipSuccess &&
              hipMalloc(&G->scan_tmp, sizeof(int) * (scan_scratch_ints((nk > n ? nk : n) + 1) + 16)) == hipSuccess &&
              hipMalloc(&G->spos, sizeof(float2) * (n + 1)) == hipSuccess;
    ok = ok && hipMemcpy(G->vstart, vstart, sizeof(int) * (ninst + 1), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(G->nc, nc.data(), sizeof(int) * ninst, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(G->cbase, cbase.data(), sizeof(int) * (ninst + 1), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(G->cs, cs.data(), sizeof(float) * ninst, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(G->rad2, rad2.data(), sizeof(float) * ninst, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_zero, dim3(blocks(nk + 1)), dim3(kT), 0, 0, G->start, nk + 1);
        ok = hipDeviceSynchronize() == hipSuccess;
    }
    if (!ok) { bgrid_free(G); delete G; return MCL_ERR_HIP; }
    *out = G;
    return MCL_OK;
}

extern "C" int mcl_bgrid_destroy(mcl_bgrid_t* G) {
    if (!G) return MCL_OK;
    (void)hipDeviceSynchronize();
    bgrid_free(G);
    delete G;
    return MCL_OK;
}

extern "C" int mcl_bgrid_sizes(const mcl_bgrid_t* G, int* npoints, int* first, int* ncells) {
    if (!G) return MCL_ERR_INPUT;
    if (npoints) *npoints = G->n;
    if (first) *first = G->v0;
    if (ncells) *ncells = G->nkeys;
    return MCL_OK;
}

extern "C" int mcl_bgrid_build(mcl_bgrid_t* G, const void* d_pos, void* stream) {
    if (!G || !d_pos) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    const int n = G->n, nk = G->nkeys;
    const BGridParams b = G->params();
    G->timers[2].mark(s);
    // each point's place in its cell from the histogram's atomics, kept in the
    // sorted positions' buffer until the gather fills it (unless that buffer
    // is the input: then a second round of atomics)
    int* slot = (const float2*)d_pos + G->v0 == G->spos ? nullptr : reinterpret_cast<int*>(G->spos);
    if (slot) hipLaunchKernelGGL(k_zero, dim3(blocks(nk + 1)), dim3(kT), 0, s, G->start, nk + 1);
    else hipLaunchKernelGGL(k_zero2, dim3(blocks(nk + 1)), dim3(kT), 0, s, G->start, G->fill, nk + 1);
    hipLaunchKernelGGL(k_bkey_count, dim3(blocks(n)), dim3(kT), 0, s, b, (const float2*)d_pos, G->key, G->start,
                       slot);
    int rc = scan_exclusive(G->start, G->start, nk, G->scan_tmp, s);
    if (rc != MCL_OK) return rc;
    hipLaunchKernelGGL(k_slot, dim3(blocks(n)), dim3(kT), 0, s, G->key, n, G->start, G->fill, (const int*)slot,
                       G->ind);
    hipLaunchKernelGGL(k_cell_rank, dim3(blocks(n)), dim3(kT), 0, s, G->key, n, G->start, (const int*)G->ind,
                       G->perm);
    hipLaunchKernelGGL(k_gather, dim3(blocks(n)), dim3(kT), 0, s, G->perm, n, (const float2*)d_pos + G->v0,
                       (const float2*)nullptr, G->spos, (float2*)nullptr, G->ind);
    G->timers[2].mark(s);
    MCL_HIP(hipGetLastError());
    return MCL_OK;
}

extern "C" void* mcl_bgrid_sorted_ptr(mcl_bgrid_t* G) { return G ? (void*)G->spos : nullptr; }

extern "C" int mcl_bgrid_index_ptrs(mcl_bgrid_t* G, void** d_ind, void** d_perm, void** d_cell_start) {
    if (!G) return MCL_ERR_INPUT;
    if (d_ind) *d_ind = G->ind;
    if (d_perm) *d_perm = G->perm;
    if (d_cell_start) *d_cell_start = G->start;
    return MCL_OK;
}

extern "C" int mcl_bgrid_frnn_csr(mcl_bgrid_t* G, void* d_offsets, void* d_nbrs, int cap, long long* total,
                                  void* stream) {
    if (!G || !d_offsets) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    const int n = G->n;
    const BGridParams b = G->params();
    int* off = (int*)d_offsets;
    G->timers[0].mark(s);
    hipLaunchKernelGGL(k_bfrnn<0>, dim3(blocks(n)), dim3(kT), 0, s, b, (const float2*)G->spos, (const int*)G->perm,
                       (const int*)G->start, off, (const int*)nullptr, (int*)nullptr, 0);
    G->timers[0].mark(s);
    int rc = scan_exclusive(off, off, n, G->scan_tmp, s);
    if (rc != MCL_OK) return rc;
    auto fill = [&]() {
        G->timers[1].mark(s);
        if (d_nbrs && cap > 0)
            hipLaunchKernelGGL(k_bfrnn<1>, dim3(blocks(n)), dim3(kT), 0, s, b, (const float2*)G->spos,
                               (const int*)G->perm, (const int*)G->start, (int*)nullptr, (const int*)off,
                               (int*)d_nbrs, cap);
        G->timers[1].mark(s);
    };
    if (!total) {  // asynchronous form: the fill pass keeps writes below cap
        fill();
        MCL_HIP(hipGetLastError());
        return MCL_OK;
    }
    int tot = 0;
    MCL_HIP(hipMemcpyAsync(&tot, off + n, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    *total = tot;
    fill();
    MCL_HIP(hipGetLastError());
    MCL_HIP(hipStreamSynchronize(s));
    return tot > cap ? MCL_ERR_CAPACITY : MCL_OK;
}

extern "C" int mcl_bgrid_kernel_timer(mcl_bgrid_t* G, int enable) {
    if (!G) return MCL_ERR_INPUT;
    for (auto& k : G->timers) k.reset(enable != 0);
    return MCL_OK;
}
extern "C" int mcl_bgrid_kernel_timer_read(mcl_bgrid_t* G, int which, double* total_ms, long long* launches) {
    if (!G || which < 0 || which > 2) return MCL_ERR_INPUT;
    double ms = G->timers[which].total_ms(launches);
    if (total_ms) *total_ms = ms;
    return MCL_OK;
}
