// mcl_internal.hpp -- declarations shared by the translation units of
// libmcleap_hip.so.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/mcleap_hip.h"

#define MCL_HIP(call)                                  \
    do {                                               \
        hipError_t _e = (call);                        \
        if (_e != hipSuccess) return mcl::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

// return a non-OK status from the enclosing function
#define MCL_TRY(x)                   \
    do {                             \
        int _r = (x);                \
        if (_r != MCL_OK) return _r; \
    } while (0)

namespace mcl {

int hip_fail(hipError_t e, const char* what, const char* file, int line);

// ---------------------------------------------------------------- host mesh
struct HostMesh {
    bool periodic = false;
    int nv = 0;
    std::vector<float> pos;   // 2 * nv
    std::vector<int> vrec;    // 2 * nv  (pos_index, he)
    std::vector<int> he;      // 4 * nhe (next, v, t, op)
    std::vector<int> t;       // nt
    std::vector<int> perm;    // n input indices in vertex order
    long long flips = 0;
    // ensemble of box instances (empty for a single mesh): instance i owns
    // corners 4i..4i+3, points vstart[i]..vstart[i+1]-1, half-edges
    // hestart[i].. and triangles tstart[i]..
    std::vector<int> vstart, hestart, tstart;
};
int build_box(const float* xy, int n, HostMesh& m);

// Hilbert index of cell (x, y) on an n x n grid (n a power of two); the
// vertex order of every box build (host and device)
__host__ __device__ inline uint64_t hilbert_d(uint32_t n, uint32_t x, uint32_t y) {
    uint64_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        const uint32_t rx = (x & s) > 0, ry = (y & s) > 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            const uint32_t tmp = x;
            x = y;
            y = tmp;
        }
    }
    return d;
}
int build_torus(const float* xy, int n, HostMesh& m);
std::vector<int> hilbert_order(const float* xy, int n);

// ------------------------------------------------------------- device mesh
// Structure-of-arrays half-edge mesh.  Half-edge h and its twin h^1 are
// adjacent, so the pair (2e, 2e+1) of edge e is one 8-byte load in every
// array.  Vertex records and triangle records are separate arrays.
struct DMesh {
    int nv = 0, nhe = 0, nt = 0, ne = 0;
    int periodic = 0;
    int ncorner = 0;  // corner vertices come first (box: ids 0..3, outer face through them)
    float2* pos = nullptr;
    int* he_next = nullptr;
    int* he_v = nullptr;
    int* he_t = nullptr;
    int* he_op = nullptr;
    int* v_he = nullptr;
    int* v_pidx = nullptr;
    int* t_he = nullptr;
};

// the flip rounds' round and stamp counters of a triangulation (read, or
// set when set is true); saved in state files so that a resumed run repeats
int tri_counters(mcl_tri_t* t, int* round, int* stamp, bool set);
// the autotuner's table (mcl_tune.hip) applied to a new triangulation / grid
void tune_tri(mcl_tri_t* t, int periodic, int ncorner, int nv);
void tune_grid(mcl_grid_t* g, int periodic, int nump);
// grid sort of pos + the asynchronous CSR FRNN query, no host reads
int grid_enqueue_frnn(mcl_grid_t* g, const float2* pos, float rad, int* off, int* nbrs, int cap, hipStream_t s);

// a valid triangulation of points at their row positions, built on the
// device into m: ninst box instances of counts[i] points (nv = N + 4 ninst,
// nhe = 6N + 10 ninst, nt = 2N + 2 ninst, every instance's corners first) or
// one torus (m.periodic, ninst = 1: nv = N, nhe = 6N, nt = 2N); target gets
// every vertex's own position, perm (or null) the input index of each point
// vertex in order
int gbuild(DMesh& m, const float2* xy, const int* counts, int ninst, float2* target, int* perm, hipStream_t s);

// statistics slots kept on the device (unsigned long long)
enum StatSlot : int {
    kStRounds = 0, kStFlips, kStUntangleFlips, kStReverts, kStIterations, kStStalled, kStInverted,
    kStOrient0, kStOrient1, kStOrient2, kStOrient3,
    kStIncircle0, kStIncircle1, kStIncircle2, kStIncircle3,
    kStRelocated, kStRelocFailed,
    kStRelWalk, kStRelStar, kStRelStuck, kStRelPutBack,  // reasons of the failures
    kStRelDup,     // vertices left where they were because another vertex holds the same target
    kStRelocGrid,  // relocations done by the full-grid apply launches (timed by timer 3)
    kStCount
};

// device counters (int)
enum CountSlot : int {
    kCnCand = 0,      // 3 slots: candidate edge list sizes, by round % 3
    kCnWants = 3,     // 3 slots: edges that want to flip, by round % 3
    kCnBad = 6,       // 3 slots: inverted triangles found, by pass % 3
    kCnRev = 9,       // 3 slots: vertices reverted in the pass
    kCnMove = 12,     // vertices moving in this iteration
    kCnPend = 13,     // vertices still short of their target
    kCnErr = 14,      // error flags
    kCnMisc = 15,
    kCnTouched = 16,  // vertices of quads flipped by repair rounds
    kCnDefer = 17,    // relocations deferred past the Lawson rounds
    kCnCount = 32
};

}  // namespace mcl

namespace mcl {
// GPU time of one kernel class, measured with event pairs recorded on the
// launch stream around each launch; read back after the timed region.
struct KTimer {
    bool on = false;
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    void reset(bool enable) {
        on = enable;
        used = 0;
    }
    void mark(hipStream_t s) {
        if (!on) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return;  // not in graphs
        if (used == ev.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) { on = false; return; }
            ev.push_back(e);
        }
        (void)hipEventRecord(ev[used++], s);
    }
    double total_ms(long long* launches) {
        double tot = 0.0;
        long long n = 0;
        for (size_t i = 0; i + 1 < used; i += 2) {
            float ms = 0.f;
            if (hipEventSynchronize(ev[i + 1]) == hipSuccess && hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) {
                tot += ms;
                n++;
            }
        }
        if (launches) *launches = n;
        return tot;
    }
    ~KTimer() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};
// XCD-aware split of n items over the grid: the blocks of a launch are dealt
// round-robin to the 8 XCDs, each with its own L2, so block b runs on XCD
// b % 8.  Each XCD gets one contiguous share of the (Hilbert-ordered) items,
// so that neighbouring items, which gather the same rows, meet in one L2.
// Yields this block's first item, the end of its XCD's share and the stride
// (all block-uniform); falls back to the plain grid-stride split when the
// grid is not a multiple of 8.
constexpr int kXcds = 8;
// Logical block index for a grid-stride loop: consecutive blocks are dealt
// round-robin to the XCDs, so with gridDim.x a multiple of kXcds this maps the
// blocks of one XCD to one contiguous range (its L2 then holds that part of
// the data and its halo).
__device__ __forceinline__ int xcd_block() {
    const int nb = gridDim.x;
    if (nb % kXcds != 0) return blockIdx.x;
    return (blockIdx.x % kXcds) * (nb / kXcds) + blockIdx.x / kXcds;
}

// ints of LDS per wave for staging a wave's CSR one-ring rows
constexpr int kRingWave = 1024;

// Block-buffered appends to a global list: items collect in an LDS buffer
// and go out with one atomic reservation and a coalesced copy per flush.  A
// single counter serialises its atomics: one reservation per block per
// iteration of a grid-stride loop (thousands of blocks, hundreds of
// iterations) costs milliseconds by itself.
// Orders the LDS accesses of one wave where LDS data belong to one wave (a
// wave's staged rows, a vertex's candidates): the LDS operations of a wave
// execute in order, and the fences keep the compiler from moving accesses
// across.  Unlike a block barrier it does not make the block's other waves
// wait.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kAppendBuf = 8192;  // ints of LDS per block (the star checks' and scatter's buffer)
constexpr int kListBuf = 2048;    // ints of LDS per block (the flip rounds' lists: up to 4 per thread a tile)
template <int kN = kAppendBuf>
struct AppendBuf {
    int buf[kN];
    int fill, base;
};
template <int kN>
__device__ __forceinline__ void append_init(AppendBuf<kN>& ab) {
    if (threadIdx.x == 0) ab.fill = 0;
    __syncthreads();
}
// this thread's offset for n items (a wave prefix sum, one LDS atomic per wave)
template <int kN>
__device__ __forceinline__ int append_reserve(AppendBuf<kN>& ab, int n) {
    const int lane = threadIdx.x & 63;
    int x = n;
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
        const int y = __shfl_up(x, dd);
        if (lane >= dd) x += y;
    }
    const int wtot = __shfl(x, 63);
    int woff = 0;
    if (lane == 63 && wtot) woff = atomicAdd(&ab.fill, wtot);
    woff = __shfl(woff, 63);
    return woff + x - n;
}
// writes the buffer out (list null: only counts it) and empties it; every
// thread of the block, after a barrier
template <int kN>
__device__ __forceinline__ void append_flush(AppendBuf<kN>& ab, int* count, int* __restrict__ list) {
    if (threadIdx.x == 0 && ab.fill) ab.base = atomicAdd(count, ab.fill);
    __syncthreads();
    const int nf = ab.fill, base = ab.base;
    if (list)
        for (int i = threadIdx.x; i < nf; i += blockDim.x) list[base + i] = ab.buf[i];
    __syncthreads();
    if (threadIdx.x == 0) ab.fill = 0;
    __syncthreads();
}
// after a tile's items are written (and a barrier): flush when the next tile,
// up to max_tile items, might not fit
template <int kN>
__device__ __forceinline__ void append_tile_end(AppendBuf<kN>& ab, int max_tile, int* count, int* __restrict__ list) {
    __syncthreads();
    if (ab.fill > kN - max_tile) append_flush(ab, count, list);
}
template <int kN>
__device__ __forceinline__ void append_finish(AppendBuf<kN>& ab, int* count, int* __restrict__ list) {
    __syncthreads();
    if (ab.fill > 0) append_flush(ab, count, list);
}

// The same buffer without barriers between tiles: a wave's items go to the
// LDS buffer while it has room and the rest straight to the list (one global
// reservation per wave for what does not fit), and the buffer is written out
// once, by append_finish_nb.  Positions are handed out by one LDS atomic per
// wave, so the waves of a block never wait for each other until the end.
// For blocks that append less than the buffer holds (a few tiles each, as
// the star checks and scatter over their lists); a block that sweeps many
// tiles fills it early, and every later wave then takes a global atomic on
// the one counter (the flip rounds' grids keep the per-tile flush).
struct AppendSlot {
    int at;     // buffer position of this thread's first item
    int gbase;  // list index of the item at buffer position p >= kN: gbase + p
};
template <int kN>
__device__ __forceinline__ AppendSlot append_reserve_nb(AppendBuf<kN>& ab, int n, int* count) {
    const int lane = threadIdx.x & 63;
    int x = n;
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
        const int y = __shfl_up(x, dd);
        if (lane >= dd) x += y;
    }
    const int wtot = __shfl(x, 63);
    int old = 0;
    if (lane == 63 && wtot) old = atomicAdd(&ab.fill, wtot);
    old = __shfl(old, 63);
    const int cut = old > kN ? old : kN;             // first position of this wave past the buffer
    const int over = old + wtot > cut ? old + wtot - cut : 0;
    int g = 0;
    if (lane == 63 && over) g = atomicAdd(count, over);
    g = __shfl(g, 63);
    return AppendSlot{old + x - n, g - cut};
}
// (list null: the items are only counted)
template <int kN>
__device__ __forceinline__ void append_put(AppendBuf<kN>& ab, const AppendSlot& sl, int k, int v,
                                           int* __restrict__ list) {
    const int p = sl.at + k;
    if (p < kN) ab.buf[p] = v;
    else if (list) list[sl.gbase + p] = v;
}
// every thread of the block, once, after its last append_put (list null:
// only the count is added)
template <int kN>
__device__ __forceinline__ void append_finish_nb(AppendBuf<kN>& ab, int* count, int* __restrict__ list) {
    __syncthreads();
    const int nf = ab.fill < kN ? ab.fill : kN;
    if (threadIdx.x == 0 && nf) ab.base = atomicAdd(count, nf);
    __syncthreads();
    const int base = ab.base;
    if (list)
        for (int i = threadIdx.x; i < nf; i += blockDim.x) list[base + i] = ab.buf[i];
}

__device__ __forceinline__ void xcd_split(int n, int per_block, int& first, int& end, int& stride) {
    const int nb = gridDim.x;
    if (nb % kXcds != 0) {
        first = blockIdx.x * per_block;
        end = n;
        stride = nb * per_block;
        return;
    }
    const int x = blockIdx.x % kXcds, k = blockIdx.x / kXcds, per_xcd = nb / kXcds;
    const int share = (((n + kXcds - 1) / kXcds) + per_block - 1) / per_block * per_block;
    const int lo = min(n, x * share);
    first = lo + k * per_block;
    end = min(n, lo + share);
    stride = per_xcd * per_block;
}
}  // namespace mcl
