// mcl_scan.hip -- device-wide exclusive prefix sum of int32 counts.
//
// Up to kScanTwoPassTiles tiles of 1024 elements: two launches without
// inter-block waiting: (1) each tile's total; (2) each 256-thread block sums
// the totals of the tiles before its own and scans its tile (4 elements per
// thread, wave-level scan with __shfl_up over the 64 lanes, then across the
// 4 waves in LDS) from there.  Longer inputs take three launches: (1) each
// tile is scanned and its total stored; (2) one block scans the tile totals
// with a running carry; (3) tiles add their carry.  out[n] receives the
// grand total.
#include "mcl_internal.hpp"
#include "mcl_scan.hpp"

namespace mcl {

constexpr int kScanThreads = 256;
constexpr int kScanPer = 4;
constexpr int kScanTile = kScanThreads * kScanPer;

// exclusive scan of one value per thread across a 256-thread block
__device__ __forceinline__ int block_exclusive(int v, int* lds_wave, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds_wave[wave] = x;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; w++) {
        int s = lds_wave[w];
        if (w < wave) base += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const int* in, int* out,
                                                             int n, int* __restrict__ tile_sum) {
    __shared__ int lw[kScanThreads / 64];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    int v[kScanPer], s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int tot;
    int ex = block_exclusive(s, lw, &tot);
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_sums(int* __restrict__ tile_sum, int ntiles,
                                                            int* __restrict__ grand) {
    __shared__ int lw[kScanThreads / 64];
    int carry = 0;
    for (int off = 0; off < ntiles; off += kScanThreads) {
        int i = off + threadIdx.x;
        int v = i < ntiles ? tile_sum[i] : 0;
        int tot;
        int ex = block_exclusive(v, lw, &tot);
        if (i < ntiles) tile_sum[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *grand = carry;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_add(int* __restrict__ out, int n,
                                                           const int* __restrict__ tile_sum) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    const int c = tile_sum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanPer; k++)
        if (base + k < n) out[base + k] += c;
}

// two-launch form: tile totals, then each tile scanned from the sum of the
// totals before it (every block reads at most kScanTwoPassTiles totals)
constexpr int kScanTwoPassTiles = 4096;

// block-wide sum of one value per thread (256 threads)
__device__ __forceinline__ int block_sum(int v, int* lds_wave) {
    int tot;
    (void)block_exclusive(v, lds_wave, &tot);
    return tot;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const int* __restrict__ in, int n,
                                                              int* __restrict__ tile_sum) {
    __shared__ int lw[kScanThreads / 64];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) s += (base + k < n) ? in[base + k] : 0;
    const int tot = block_sum(s, lw);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const int* in, int* out, int n,
                                                             const int* __restrict__ tile_sum, int ntiles) {
    __shared__ int lw[kScanThreads / 64];
    int c = 0;
    for (int i = threadIdx.x; i < (int)blockIdx.x; i += kScanThreads) c += tile_sum[i];
    const int carry = block_sum(c, lw);
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    int v[kScanPer], s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int tot;
    int ex = carry + block_exclusive(s, lw, &tot);
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if ((int)blockIdx.x == ntiles - 1 && threadIdx.x == 0) out[n] = carry + tot;
}

int scan_scratch_ints(int n) { return (n + kScanTile - 1) / kScanTile + 1; }

int scan_exclusive(const int* d_in, int* d_out, int n, int* d_scratch, hipStream_t s) {
    if (n <= 0) {
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, d_scratch, 0, d_out);
        MCL_HIP(hipGetLastError());
        return MCL_OK;
    }
    int ntiles = (n + kScanTile - 1) / kScanTile;
    if (ntiles <= kScanTwoPassTiles) {
        hipLaunchKernelGGL(k_scan_reduce, dim3(ntiles), dim3(kScanThreads), 0, s, d_in, n, d_scratch);
        hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(kScanThreads), 0, s, d_in, d_out, n,
                           (const int*)d_scratch, ntiles);
        MCL_HIP(hipGetLastError());
        return MCL_OK;
    }
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kScanThreads), 0, s, d_in, d_out, n, d_scratch);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, d_scratch, ntiles, d_out + n);
    hipLaunchKernelGGL(k_scan_add, dim3(ntiles), dim3(kScanThreads), 0, s, d_out, n, d_scratch);
    MCL_HIP(hipGetLastError());
    return MCL_OK;
}

}  // namespace mcl
