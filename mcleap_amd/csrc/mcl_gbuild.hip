// mcl_gbuild.hip -- the exact Delaunay triangulation of a point set built on
// the GPU (in place of HostTriangulation::addDelaunayPoints and the upload,
// Host_Triangulation.h:296-300, Device_Triangulation.h:232-306), for one
// triangulation or for an ensemble of independent box triangulations held as
// one mesh (every instance's 4 corners first, as mcl_hmesh_build_batch lays
// them out).
//
//  1. Vertex order: the host build's Hilbert keys (each instance on the grid
//     of its own bounding box), one stable radix sort of (instance, key), so
//     every vertex is the same point as in the host build.
//  2. A valid triangulation of each instance's points at "row" positions.
//     Rows by (y, x) rank: R = floor(sqrt n) rows of n / R or n / R + 1
//     points.  Each row lies on a horizontal line (the mean y of its points,
//     raised by one ulp where needed so that rows strictly increase).  In a
//     box a row's points keep their own x except the first and last, which
//     sit at the instance's min and max x; consecutive rows are joined by
//     zipper strips and the rectangle they fill is joined to the instance's
//     4 corners by one fan per side, the hull bowed out slightly so that no
//     three hull points are collinear.  On the torus rows are closed loops
//     and strips annuli.  Every triangle is checked strictly CCW (exact).
//  3. Half-edges from the triangle list: one radix sort of the undirected
//     edge keys pairs every half-edge with its twin.
//  4. (build_impl in mcl_tri.hip) the kinetic step moves every vertex from
//     its row position to its point -- at most about one row height -- and
//     ends on the exact Delaunay triangulation.
#include <algorithm>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "mcl_geom.hpp"
#include "mcl_internal.hpp"

#define GB_TRY(x)                      \
    do {                               \
        const int rc_ = (x);           \
        if (rc_ != MCL_OK) return rc_; \
    } while (0)

namespace mcl {
namespace {

constexpr int kB = 256;

int nblocks(long long n) {
    long long b = (n + kB - 1) / kB;
    return (int)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

// bits needed for the values 0 .. x (x >= 0)
int bit_width(long long x) {
    int b = 0;
    while (x > 0) {
        b++;
        x >>= 1;
    }
    return b;
}

// order-preserving unsigned image of a float
__host__ __device__ __forceinline__ uint32_t ford(float f) {
#pragma clang fp contract(off)
    const uint32_t u = __builtin_bit_cast(uint32_t, f + 0.0f);  // -0 -> +0
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float funord(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __builtin_bit_cast(float, u);
}

// One instance.  Its points are sorted positions pstart .. pstart + n - 1
// (vertices base + pstart ..., base = 4 ninst in a box, 0 on the torus), its
// rows the global rows rowbase .. rowbase + R - 1 (local rows 0 .. extra-1
// hold q + 1 points, the others q), its triangles tbase .. tbase + 2n + 1
// (2n - 1 on the torus).
struct Inst {
    int pstart, n, R, q, extra, rowbase, tbase, pad;
    float X0, X1;    // row x range (box)
    float lox, loy;  // Hilbert grid origin
    double sx, sy;   // Hilbert grid scale
    __host__ __device__ int start(int r) const {  // first position of local row r, from pstart
        return r < extra ? r * (q + 1) : extra * (q + 1) + (r - extra) * q;
    }
    __host__ __device__ int count(int r) const { return r < extra ? q + 1 : q; }
    __host__ __device__ int row_of(int j) const {  // local row of local rank j
        const int big = extra * (q + 1);
        return j < big ? j / (q + 1) : extra + (j - big) / q;
    }
};

struct Ens {
    const Inst* inst;
    const int* pstart;   // ninst + 1 point offsets
    const int* rowinst;  // global row -> instance
    int ninst, base, periodic, nrows;
};

// the instance of sorted position j: the last i with pstart[i] <= j
__device__ __forceinline__ int inst_of(const Ens& e, int j) {
    int lo = 0, hi = e.ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (e.pstart[mid] <= j) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// per instance (blockIdx.y) min and max of x and y as ordered keys: a block
// reduction, then 4 atomics per block
__global__ void __launch_bounds__(kB) k_minmax(const float2* __restrict__ xy, Ens e, uint32_t* __restrict__ mm) {
    __shared__ uint32_t red[4][kB / 64];
    for (int i = blockIdx.y; i < e.ninst; i += gridDim.y) {
        const int p1 = e.pstart[i + 1];
        uint32_t v0 = 0xffffffffu, v1 = 0xffffffffu, v2 = 0u, v3 = 0u;
        for (int p = e.pstart[i] + blockIdx.x * kB + threadIdx.x; p < p1; p += gridDim.x * kB) {
            const float2 q = xy[p];
            v0 = min(v0, ford(q.x));
            v1 = min(v1, ford(q.y));
            v2 = max(v2, ford(q.x));
            v3 = max(v3, ford(q.y));
        }
        for (int o = 32; o > 0; o >>= 1) {
            v0 = min(v0, (uint32_t)__shfl_xor((int)v0, o));
            v1 = min(v1, (uint32_t)__shfl_xor((int)v1, o));
            v2 = max(v2, (uint32_t)__shfl_xor((int)v2, o));
            v3 = max(v3, (uint32_t)__shfl_xor((int)v3, o));
        }
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            red[0][w] = v0;
            red[1][w] = v1;
            red[2][w] = v2;
            red[3][w] = v3;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int k = 1; k < kB / 64; k++) {
                v0 = min(v0, red[0][k]);
                v1 = min(v1, red[1][k]);
                v2 = max(v2, red[2][k]);
                v3 = max(v3, red[3][k]);
            }
            atomicMin(&mm[4 * i], v0);
            atomicMin(&mm[4 * i + 1], v1);
            atomicMax(&mm[4 * i + 2], v2);
            atomicMax(&mm[4 * i + 3], v3);
        }
        __syncthreads();
    }
}

// (instance, the host build's Hilbert key): 16-bit grid coordinates in double
__global__ void __launch_bounds__(kB) k_hilbert_keys(const float2* __restrict__ xy, Ens e,
                                                     unsigned long long* __restrict__ key, int* __restrict__ idx) {
#pragma clang fp contract(off)
    const int n = e.pstart[e.ninst];
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int i = inst_of(e, p);
        const Inst I = e.inst[i];
        const float2 q = xy[p];
        const uint32_t gx = (uint32_t)(((double)q.x - I.lox) * I.sx);
        const uint32_t gy = (uint32_t)(((double)q.y - I.loy) * I.sy);
        key[p] = ((unsigned long long)i << 32) | (uint32_t)hilbert_d(1u << 16, gx, gy);
        idx[p] = p;
    }
}

// sorted position j (vertex base + j) = point order[j]: its target and x key
__global__ void __launch_bounds__(kB) k_place(const float2* __restrict__ xy, const int* __restrict__ order, Ens e,
                                              float2* __restrict__ target, uint32_t* __restrict__ xk,
                                              int* __restrict__ vid) {
    const int n = e.pstart[e.ninst];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const float2 p = xy[order[j]];
        target[e.base + j] = p;
        xk[j] = ford(p.x);
        vid[j] = e.base + j;
    }
}

// (instance, y) keys of the vertices in x order, for the stable second pass
__global__ void __launch_bounds__(kB) k_iy_keys(const int* __restrict__ vid, Ens e, const float2* __restrict__ target,
                                                unsigned long long* __restrict__ key) {
    const int n = e.pstart[e.ninst];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const int v = vid[j];
        key[j] = ((unsigned long long)inst_of(e, v - e.base) << 32) | ford(target[v].y);
    }
}

// in (instance, y, x) order adjacent equal points are duplicates; then the
// (global row, x) keys
__global__ void __launch_bounds__(kB) k_row_keys(const int* __restrict__ vid, Ens e, const float2* __restrict__ target,
                                                 unsigned long long* __restrict__ rx, int* __restrict__ rvid,
                                                 int* __restrict__ dup) {
    const int n = e.pstart[e.ninst];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const int v = vid[j];
        const int i = inst_of(e, j);
        const float2 p = target[v];
        if (j + 1 < e.pstart[i + 1]) {
            const float2 q = target[vid[j + 1]];
            if (ford(p.x) == ford(q.x) && ford(p.y) == ford(q.y)) atomicAdd(dup, 1);
        }
        const Inst& I = e.inst[i];
        const int row = I.rowbase + I.row_of(j - I.pstart);
        rx[j] = ((unsigned long long)row << 32) | ford(p.x);
        rvid[j] = v;
    }
}

// one thread per row: the mean y of its points (double)
__global__ void __launch_bounds__(kB) k_row_mean(const int* __restrict__ lat, Ens e,
                                                 const float2* __restrict__ target, float* __restrict__ ry) {
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < e.nrows; g += gridDim.x * blockDim.x) {
        const Inst& I = e.inst[e.rowinst[g]];
        const int r = g - I.rowbase, s = I.pstart + I.start(r), m = I.count(r);
        double acc = 0.0;
        for (int c = 0; c < m; c++) acc += (double)target[lat[s + c]].y;
        ry[g] = (float)(acc / m);
    }
}

// one thread per instance: its row lines strictly increasing; gap[i] = the
// smallest row gap of instance i
__global__ void __launch_bounds__(kB) k_row_fix(Ens e, float* __restrict__ ry, float* __restrict__ gap) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < e.ninst; i += gridDim.x * blockDim.x) {
        const Inst& I = e.inst[i];
        float* y = ry + I.rowbase;
        float gmin = 3.0e38f;
        for (int r = 1; r < I.R; r++) {
            if (!(y[r] > y[r - 1])) y[r] = nextafterf(y[r - 1], 3.0e38f);
            gmin = fminf(gmin, y[r] - y[r - 1]);
        }
        gap[i] = gmin;
    }
}

// One thread per box row: row positions.  First point at X0, last at X1, the
// others at their own x raised to one ulp above their left neighbour; a row
// where that reaches X1 is spread evenly instead.  With bulge set (R >= 3)
// the hull gets no three collinear points: the first and last points of the
// inner rows move out along a parabola in r (zero at the end rows), and the
// bottom and top rows bow out by a parabola in x of height a quarter of
// their gap to the next row.  The caller checks every triangle and rebuilds
// without bulge if one is not strictly CCW.
__global__ void __launch_bounds__(kB) k_row_pos(const int* __restrict__ lat, Ens e, const float* __restrict__ ry,
                                                const float* __restrict__ gap, const float2* __restrict__ target,
                                                float2* __restrict__ pos, int* __restrict__ err, int bulge) {
#pragma clang fp contract(off)
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < e.nrows; g += gridDim.x * blockDim.x) {
        const int i = e.rowinst[g];
        const Inst& I = e.inst[i];
        const int R = I.R, r = g - I.rowbase, s = I.pstart + I.start(r), m = I.count(r);
        const float X0 = I.X0, X1 = I.X1, W = X1 - X0, y = ry[g];
        if (m == 1) {
            pos[lat[s]] = make_float2(target[lat[s]].x, y);
            continue;
        }
        float prev = X0;
        bool ok = true;
        pos[lat[s]] = make_float2(X0, y);
        for (int c = 1; c + 1 < m; c++) {
            float x = target[lat[s + c]].x;
            if (!(x > prev)) x = nextafterf(prev, 3.0e38f);
            if (!(x < X1)) {
                ok = false;
                break;
            }
            pos[lat[s + c]] = make_float2(x, y);
            prev = x;
        }
        if (!ok) {
            prev = X0;
            for (int c = 1; c + 1 < m; c++) {
                const float x = X0 + W * ((float)c / (float)(m - 1));
                if (!(x > prev && x < X1)) atomicAdd(err, 1);  // x range too narrow for the row
                pos[lat[s + c]] = make_float2(x, y);
                prev = x;
            }
        }
        pos[lat[s + m - 1]] = make_float2(X1, y);
        if (bulge && R >= 3) {
            if (r > 0 && r + 1 < R) {  // side chains
                // height A: a quarter of the smallest row gap, or more where
                // float resolution at |x| would straighten the parabola, but
                // never so much that a chain step leans out by a row gap
                const float gmin = gap[i];
                const float mx = fmaxf(fabsf(X0), fabsf(X1));
                const float ulp = nextafterf(mx, 3.0e38f) - mx;
                const float A = fminf(0.0625f * gmin * (float)(R - 1),
                                      fmaxf(0.25f * gmin, 0.25f * ulp * (float)R * (float)R));
                const float b = (float)r * (float)(R - 1 - r) / (0.25f * (float)(R - 1) * (float)(R - 1));
                const float dx = A * b;
                pos[lat[s]].x = X0 - dx;
                pos[lat[s + m - 1]].x = X1 + dx;
            } else {  // bottom row down, top row up
                const float h = r == 0 ? ry[g + 1] - ry[g] : ry[g] - ry[g - 1];
                const float sgn = r == 0 ? -1.0f : 1.0f;
                for (int c = 1; c + 1 < m; c++) {
                    const float x = pos[lat[s + c]].x;
                    const float q = (x - X0) * (X1 - x) / (0.25f * W * W);
                    pos[lat[s + c]].y = y + sgn * 0.25f * h * q;
                }
            }
        }
    }
}

// torus rows: every point at its own x raised to one ulp above its left
// neighbour (a row is a closed loop, so no point is pinned); a row where
// that reaches 1 is spread evenly instead
__global__ void __launch_bounds__(kB) k_row_pos_torus(const int* __restrict__ lat, Ens e,
                                                      const float* __restrict__ ry, const float2* __restrict__ target,
                                                      float2* __restrict__ pos) {
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < e.nrows; g += gridDim.x * blockDim.x) {
        const Inst& I = e.inst[e.rowinst[g]];
        const int r = g - I.rowbase, s = I.pstart + I.start(r), m = I.count(r);
        const float y = ry[g];
        float prev = -1.0f;
        bool ok = true;
        for (int c = 0; c < m; c++) {
            float x = target[lat[s + c]].x;
            if (!(x > prev)) x = nextafterf(prev, 3.0e38f);
            if (!(x < 1.0f)) {
                ok = false;
                break;
            }
            pos[lat[s + c]] = make_float2(x, y);
            prev = x;
        }
        if (!ok)
            for (int c = 0; c < m; c++) pos[lat[s + c]] = make_float2(((float)c + 0.5f) / (float)m, y);
    }
}

// One thread per strip.  Box: local rows r and r+1 of an instance, from
// triangle tbase + start(r) + start(r+1) - m0 - 2r, a zipper that advances
// along the row whose next point has the smaller x.  Torus: row r and row
// r+1 (mod R) bound an annulus, zipped once around from the rung between
// their first points (m_r + m_r+1 triangles, from start(r) + start(r+1) -
// m0), x unrolled past 1.
__global__ void __launch_bounds__(kB) k_strips(const int* __restrict__ lat, Ens e, const float2* __restrict__ pos,
                                               int3* __restrict__ tri) {
    const bool torus = e.periodic != 0;
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < e.nrows; g += gridDim.x * blockDim.x) {
        const Inst& I = e.inst[e.rowinst[g]];
        const int R = I.R, r = g - I.rowbase, m0 = I.count(0);
        if (!torus && r + 1 >= R) continue;
        const int rt = r + 1 == R ? 0 : r + 1;
        const int sb = I.pstart + I.start(r), mb = I.count(r), st = I.pstart + I.start(rt), mt = I.count(rt);
        int o = I.tbase + I.start(r) + I.start(r + 1) - m0 - (torus ? 0 : 2 * r);
        if (torus) {
            auto bx = [&](int k) { return k < mb ? pos[lat[sb + k]].x : pos[lat[sb + k - mb]].x + 1.0f; };
            auto tx = [&](int k) { return k < mt ? pos[lat[st + k]].x : pos[lat[st + k - mt]].x + 1.0f; };
            auto bv = [&](int k) { return lat[sb + (k < mb ? k : k - mb)]; };
            auto tv = [&](int k) { return lat[st + (k < mt ? k : k - mt)]; };
            int a = 0, b = 0;
            while (a < mb || b < mt) {
                if (b == mt || (a < mb && bx(a + 1) <= tx(b + 1))) {
                    tri[o++] = make_int3(bv(a), bv(a + 1), tv(b));
                    a++;
                } else {
                    tri[o++] = make_int3(bv(a), tv(b + 1), tv(b));
                    b++;
                }
            }
        } else {
            int a = 0, b = 0;
            while (a < mb - 1 || b < mt - 1) {
                if (b == mt - 1 || (a < mb - 1 && pos[lat[sb + a + 1]].x <= pos[lat[st + b + 1]].x)) {
                    tri[o++] = make_int3(lat[sb + a], lat[sb + a + 1], lat[st + b]);
                    a++;
                } else {
                    tri[o++] = make_int3(lat[sb + a], lat[st + b + 1], lat[st + b]);
                    b++;
                }
            }
        }
    }
}

// box fans to the corners of instance i = blockIdx.y (4i + 0 BL, 1 BR, 2 TR,
// 3 TL) and its 4 corner triangles; an instance without points is the
// square's two triangles
__global__ void __launch_bounds__(kB) k_fans(const int* __restrict__ lat, Ens e, int3* __restrict__ tri) {
    for (int i = blockIdx.y; i < e.ninst; i += gridDim.y) {
        const Inst& I = e.inst[i];
        const int c0 = 4 * i;
        if (I.n == 0) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                tri[I.tbase] = make_int3(c0, c0 + 1, c0 + 2);
                tri[I.tbase + 1] = make_int3(c0, c0 + 2, c0 + 3);
            }
            continue;
        }
        const int R = I.R, m0 = I.count(0), mt = I.count(R - 1), s0 = I.pstart, sT = I.pstart + I.start(R - 1);
        const int nstrip = I.start(R - 1) + I.start(R) - m0 - 2 * (R - 1);
        const int ob = nstrip, orr = ob + (m0 - 1), ot = orr + (R - 1), ol = ot + (mt - 1), oc = ol + (R - 1);
        auto first = [&](int r) { return lat[I.pstart + I.start(r)]; };
        auto last = [&](int r) { return lat[I.pstart + I.start(r) + I.count(r) - 1]; };
        for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < oc + 4 - nstrip; k += gridDim.x * blockDim.x) {
            const int o = nstrip + k;
            int3 t;
            if (o < orr) {
                const int j = o - ob;
                t = make_int3(lat[s0 + j + 1], lat[s0 + j], c0);
            } else if (o < ot) {
                const int r = o - orr;
                t = make_int3(last(r + 1), last(r), c0 + 1);
            } else if (o < ol) {
                const int j = o - ot;
                t = make_int3(lat[sT + j], lat[sT + j + 1], c0 + 2);
            } else if (o < oc) {
                const int r = o - ol;
                t = make_int3(first(r), first(r + 1), c0 + 3);
            } else {
                const int c = o - oc;
                if (c == 0) t = make_int3(c0, c0 + 1, last(0));
                else if (c == 1) t = make_int3(c0 + 1, c0 + 2, last(R - 1));
                else if (c == 2) t = make_int3(c0 + 2, c0 + 3, first(R - 1));
                else t = make_int3(c0 + 3, c0, first(0));
            }
            tri[I.tbase + o] = t;
        }
    }
}

// triangles that are not strictly CCW (exact)
__global__ void __launch_bounds__(kB) k_tri_check(const int3* __restrict__ tri, int nt, const float2* __restrict__ pos,
                                                  int periodic, int* __restrict__ bad) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
        const int3 q = tri[t];
        int st;
        if (orient<true>(pos[q.x], pos[q.y], pos[q.z], periodic != 0, &st) <= 0) atomicAdd(bad, 1);
    }
}

// half-edge slots: 3 per triangle (corner k: org tri[k], dst tri[k+1]), then
// in a box 4 per instance for its outer face BR->BL, BL->TL, TL->TR, TR->BR
__device__ __forceinline__ void slot_ends(const int3* tri, int nt, int s, int& a, int& b) {
    if (s < 3 * nt) {
        const int3 t = tri[s / 3];
        const int k = s % 3;
        a = k == 0 ? t.x : (k == 1 ? t.y : t.z);
        b = k == 0 ? t.y : (k == 1 ? t.z : t.x);
    } else {
        const int o = s - 3 * nt, c0 = o & ~3, k = o & 3;
        const int org[4] = {1, 0, 3, 2}, dst[4] = {0, 3, 2, 1};
        a = c0 + org[k];
        b = c0 + dst[k];
    }
}

__global__ void __launch_bounds__(kB) k_edge_keys(const int3* __restrict__ tri, int nt, int ns,
                                                  unsigned long long* __restrict__ key, int* __restrict__ slot) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x) {
        int a, b;
        slot_ends(tri, nt, s, a, b);
        key[s] = ((unsigned long long)(unsigned)min(a, b) << 32) | (unsigned)max(a, b);
        slot[s] = s;
    }
}

// sorted position p -> edge p / 2; each key exactly twice, once each way
__global__ void __launch_bounds__(kB) k_pair(const unsigned long long* __restrict__ key, const int* __restrict__ slot,
                                             const int3* __restrict__ tri, int nt, int ns, int* __restrict__ hslot,
                                             int* __restrict__ bad) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < ns; p += gridDim.x * blockDim.x) {
        const int q = p ^ 1;
        bool ok = q < ns && key[q] == key[p];
        if (ok && (p & 1) == 0 && p + 2 < ns && key[p + 2] == key[p]) ok = false;
        const int s = slot[p];
        int a, b, a2 = 0, b2 = 0;
        slot_ends(tri, nt, s, a, b);
        if (ok) slot_ends(tri, nt, slot[q], a2, b2);
        if (!ok || a == b || (a < b) == (a2 < b2)) {  // twins run opposite ways
            atomicAdd(bad, 1);
            continue;
        }
        hslot[s] = 2 * (p >> 1) + (a < b ? 0 : 1);
    }
}

__global__ void __launch_bounds__(kB) k_fill_mesh(DMesh m, const int3* __restrict__ tri, int nt, int ns,
                                                  const int* __restrict__ hslot, int* __restrict__ bad) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x) {
        int a, b;
        slot_ends(tri, nt, s, a, b);
        const int h = hslot[s];
        if ((unsigned)h >= (unsigned)m.nhe) {
            atomicAdd(bad, 1);
            continue;
        }
        m.he_v[h] = a;
        if (s < 3 * nt) {
            const int t = s / 3, k = s % 3;
            const int3 tt = tri[t];
            m.he_next[h] = hslot[3 * t + (k + 1) % 3];
            m.he_t[h] = t;
            m.he_op[h] = k == 0 ? tt.z : (k == 1 ? tt.x : tt.y);
            if (k == 0) m.t_he[t] = h;
        } else {
            const int o = s - 3 * nt;
            m.he_next[h] = hslot[3 * nt + (o & ~3) + ((o + 1) & 3)];
            m.he_t[h] = -1;
            m.he_op[h] = -1;
        }
        atomicMin(&m.v_he[a], h);
    }
}

__global__ void __launch_bounds__(kB) k_vinit(DMesh m, int ncorner, float2* __restrict__ target) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        m.v_he[v] = 0x7fffffff;
        m.v_pidx[v] = v;
        if (v < ncorner) {
            const float lo = -10000000.f, hi = 10000000.f;  // Common_Triangulation.h:12-13
            const int c = v & 3;
            const float2 p = make_float2(c == 1 || c == 2 ? hi : lo, c >= 2 ? hi : lo);
            m.pos[v] = p;
            target[v] = p;
        }
    }
}

template <typename K, typename V>
int radix_pairs(const K* kin, K* kout, const V* vin, V* vout, int n, int bits, hipStream_t s) {
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, n, 0, bits, s) != hipSuccess)
        return MCL_ERR_HIP;
    void* tmp = nullptr;
    if (hipMalloc(&tmp, bytes ? bytes : 1) != hipSuccess) return MCL_ERR_HIP;
    const bool ok = hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, n, 0, bits, s) == hipSuccess &&
                    hipStreamSynchronize(s) == hipSuccess;
    (void)hipFree(tmp);
    return ok ? MCL_OK : MCL_ERR_HIP;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    T* get() { return (T*)p; }
    int alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 1) == hipSuccess ? MCL_OK : MCL_ERR_HIP; }
};

int read_int(const int* d, hipStream_t s, int* out) {
    if (hipMemcpyAsync(out, d, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return MCL_ERR_HIP;
    return MCL_OK;
}

}  // namespace

// Fills m (allocated) with a valid triangulation of the points at their row
// positions.  Box (m.periodic = 0): ninst instances of counts[i] points, xy
// back to back; nv = N + 4 ninst, nhe = 6N + 10 ninst, nt = 2N + 2 ninst
// (N the total), every instance's corners first.  Torus: ninst = 1, nv = N,
// nhe = 6N, nt = 2N, points in [0,1)^2, N >= 16.  target receives every
// vertex's own position (corners included) and perm (may be null) the input
// index of each point vertex in order.
int gbuild(DMesh& m, const float2* xy, const int* counts, int ninst, float2* target, int* perm, hipStream_t s) {
    const bool torus = m.periodic != 0;
    if (ninst <= 0 || (torus && ninst != 1)) return MCL_ERR_INPUT;
    std::vector<Inst> ih(ninst);
    std::vector<int> pstart(ninst + 1, 0);
    long long N = 0;
    for (int i = 0; i < ninst; i++) {
        if (counts[i] < 0) return MCL_ERR_INPUT;
        pstart[i] = (int)N;
        N += counts[i];
        if (N >= (1LL << 30)) return MCL_ERR_CAPACITY;
    }
    pstart[ninst] = (int)N;
    const int base = torus ? 0 : 4 * ninst;
    if (N <= 0 || (long long)m.nv != N + base || (long long)m.nhe != 6 * N + (torus ? 0 : 10LL * ninst) ||
        (long long)m.nt != 2 * N + (torus ? 0 : 2LL * ninst))
        return MCL_ERR_INPUT;
    if (torus && N < 16) return MCL_ERR_INPUT;
    const int n = (int)N;
    // rows and triangle offsets per instance
    int nrows = 0, tsum = 0, maxcount = 0, maxfan = 1;
    std::vector<int> rowinst;
    for (int i = 0; i < ninst; i++) {
        Inst& I = ih[i];
        I = Inst{};
        I.pstart = pstart[i];
        I.n = counts[i];
        I.R = 0;
        while ((long long)(I.R + 1) * (I.R + 1) <= I.n) I.R++;
        I.q = I.R ? I.n / I.R : 0;
        I.extra = I.R ? I.n % I.R : 0;
        I.rowbase = nrows;
        I.tbase = tsum;
        nrows += I.R;
        tsum += 2 * I.n + (torus ? 0 : 2);
        rowinst.insert(rowinst.end(), I.R, i);
        maxcount = std::max(maxcount, I.n);
        if (I.R) maxfan = std::max(maxfan, I.count(0) + I.count(I.R - 1) + 2 * I.R + 2);
        if (torus && (I.R < 3 || I.q < 3)) return MCL_ERR_INPUT;
    }
    int rc = MCL_OK;
    DevBuf d_inst, d_pstart, d_rowinst, mm;
    if ((rc = d_inst.alloc(sizeof(Inst) * ninst)) || (rc = d_pstart.alloc(sizeof(int) * (ninst + 1))) ||
        (rc = d_rowinst.alloc(sizeof(int) * std::max(nrows, 1))) || (rc = mm.alloc(sizeof(uint32_t) * 4 * ninst)))
        return rc;
    MCL_HIP(hipMemcpyAsync(d_pstart.p, pstart.data(), sizeof(int) * (ninst + 1), hipMemcpyHostToDevice, s));
    if (nrows) MCL_HIP(hipMemcpyAsync(d_rowinst.p, rowinst.data(), sizeof(int) * nrows, hipMemcpyHostToDevice, s));
    const Ens e{d_inst.get<Inst>(), d_pstart.get<int>(), d_rowinst.get<int>(), ninst, base, torus ? 1 : 0, nrows};
    // bounding boxes, then every instance's grid and row range
    {
        std::vector<uint32_t> box(4 * (size_t)ninst);
        for (int i = 0; i < ninst; i++) {
            box[4 * i] = box[4 * i + 1] = 0xffffffffu;
            box[4 * i + 2] = box[4 * i + 3] = 0u;
        }
        MCL_HIP(hipMemcpyAsync(mm.p, box.data(), sizeof(uint32_t) * 4 * ninst, hipMemcpyHostToDevice, s));
        const int bx = std::max(1, std::min(512, (maxcount + kB * 16 - 1) / (kB * 16)));
        hipLaunchKernelGGL(k_minmax, dim3(bx, std::min(ninst, 65535)), dim3(kB), 0, s, xy, e, mm.get<uint32_t>());
        MCL_HIP(hipGetLastError());
        MCL_HIP(hipMemcpyAsync(box.data(), mm.p, sizeof(uint32_t) * 4 * ninst, hipMemcpyDeviceToHost, s));
        MCL_HIP(hipStreamSynchronize(s));
        const uint32_t G = 1u << 16;
        for (int i = 0; i < ninst; i++) {
            Inst& I = ih[i];
            if (I.n == 0) continue;
            const float lo[2] = {funord(box[4 * i]), funord(box[4 * i + 1])};
            const float hi[2] = {funord(box[4 * i + 2]), funord(box[4 * i + 3])};
            for (int k = 0; k < 2; k++) {
                if (!(lo[k] > -1e7f && hi[k] < 1e7f)) return MCL_ERR_INPUT;  // outside the corners (or NaN)
                if (torus && !(lo[k] >= 0.0f && hi[k] < 1.0f)) return MCL_ERR_INPUT;
            }
            // the host build's grid (hilbert_order in mcl_build.hip)
            I.lox = lo[0];
            I.loy = lo[1];
            I.sx = (hi[0] > lo[0]) ? (G - 1) / ((double)hi[0] - lo[0]) : 0.0;
            I.sy = (hi[1] > lo[1]) ? (G - 1) / ((double)hi[1] - lo[1]) : 0.0;
            I.X0 = lo[0];
            I.X1 = hi[0] > lo[0] ? hi[0] : lo[0] + 1.0f;
        }
        MCL_HIP(hipMemcpyAsync(d_inst.p, ih.data(), sizeof(Inst) * ninst, hipMemcpyHostToDevice, s));
    }
    const int ibits = 32 + bit_width(ninst - 1);
    // 1. vertex order: (instance, Hilbert key), stable
    DevBuf hk, hk2, hi, hi2;
    if ((rc = hk.alloc(8 * (size_t)n)) || (rc = hk2.alloc(8 * (size_t)n)) || (rc = hi.alloc(4 * (size_t)n)) ||
        (rc = hi2.alloc(4 * (size_t)n)))
        return rc;
    hipLaunchKernelGGL(k_hilbert_keys, dim3(nblocks(n)), dim3(kB), 0, s, xy, e, hk.get<unsigned long long>(),
                       hi.get<int>());
    GB_TRY(radix_pairs(hk.get<unsigned long long>(), hk2.get<unsigned long long>(), hi.get<int>(), hi2.get<int>(), n,
                       ibits, s));
    const int* order = hi2.get<int>();
    if (perm) MCL_HIP(hipMemcpyAsync(perm, order, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_vinit, dim3(nblocks(m.nv)), dim3(kB), 0, s, m, base, target);
    // 2. rows by (instance, y, x) rank: an x pass, then a stable (instance, y)
    //    pass; then the rows' points by x
    DevBuf xk, xk2, vid, vid2, dup;
    unsigned long long* yk = hk.get<unsigned long long>();  // the Hilbert keys are spent
    unsigned long long* yk2 = hk2.get<unsigned long long>();
    if ((rc = xk.alloc(4 * (size_t)n)) || (rc = xk2.alloc(4 * (size_t)n)) || (rc = vid.alloc(4 * (size_t)n)) ||
        (rc = vid2.alloc(4 * (size_t)n)) || (rc = dup.alloc(8)))
        return rc;
    hipLaunchKernelGGL(k_place, dim3(nblocks(n)), dim3(kB), 0, s, xy, order, e, target, xk.get<uint32_t>(),
                       vid.get<int>());
    GB_TRY(radix_pairs(xk.get<uint32_t>(), xk2.get<uint32_t>(), vid.get<int>(), vid2.get<int>(), n, 32, s));
    hipLaunchKernelGGL(k_iy_keys, dim3(nblocks(n)), dim3(kB), 0, s, (const int*)vid2.get<int>(), e,
                       (const float2*)target, yk);
    GB_TRY(radix_pairs(yk, yk2, vid2.get<int>(), vid.get<int>(), n, ibits, s));
    MCL_HIP(hipMemsetAsync(dup.p, 0, 8, s));
    hipLaunchKernelGGL(k_row_keys, dim3(nblocks(n)), dim3(kB), 0, s, (const int*)vid.get<int>(), e,
                       (const float2*)target, yk, vid2.get<int>(), dup.get<int>());
    int ndup = 0;
    GB_TRY(read_int(dup.get<int>(), s, &ndup));
    if (ndup) return MCL_ERR_DUPLICATE;
    GB_TRY(radix_pairs(yk, yk2, vid2.get<int>(), vid.get<int>(), n, 32 + bit_width(nrows - 1), s));
    const int* lat = vid.get<int>();  // local row r of instance i at pstart[i] + start(r)
    DevBuf ry, gap;
    if ((rc = ry.alloc(sizeof(float) * std::max(nrows, 1))) || (rc = gap.alloc(sizeof(float) * ninst))) return rc;
    hipLaunchKernelGGL(k_row_mean, dim3(nblocks(nrows)), dim3(kB), 0, s, lat, e, (const float2*)target, ry.get<float>());
    hipLaunchKernelGGL(k_row_fix, dim3(nblocks(ninst)), dim3(kB), 0, s, e, ry.get<float>(), gap.get<float>());
    // row positions and triangles, first with the bowed hull
    const int nt = m.nt;
    DevBuf tri, bad;
    if ((rc = tri.alloc(sizeof(int3) * (size_t)nt)) || (rc = bad.alloc(8))) return rc;
    // a triangle left unwritten would read (0, 0, 0): degenerate, so reported
    MCL_HIP(hipMemsetAsync(tri.p, 0, sizeof(int3) * (size_t)nt, s));
    for (int bulge = 1; bulge >= 0; bulge--) {
        if (torus) {
            hipLaunchKernelGGL(k_row_pos_torus, dim3(nblocks(nrows)), dim3(kB), 0, s, lat, e,
                               (const float*)ry.get<float>(), (const float2*)target, m.pos);
        } else {
            MCL_HIP(hipMemsetAsync(dup.p, 0, 8, s));
            hipLaunchKernelGGL(k_row_pos, dim3(nblocks(nrows)), dim3(kB), 0, s, lat, e, (const float*)ry.get<float>(),
                               (const float*)gap.get<float>(), (const float2*)target, m.pos, dup.get<int>(), bulge);
            GB_TRY(read_int(dup.get<int>(), s, &ndup));
            if (ndup) return MCL_ERR_INPUT;
        }
        hipLaunchKernelGGL(k_strips, dim3(nblocks(nrows)), dim3(kB), 0, s, lat, e, (const float2*)m.pos,
                           tri.get<int3>());
        if (!torus)
            hipLaunchKernelGGL(k_fans, dim3(std::min(nblocks(maxfan), 64), std::min(ninst, 65535)), dim3(kB), 0, s, lat,
                               e, tri.get<int3>());
        MCL_HIP(hipMemsetAsync(bad.p, 0, 8, s));
        hipLaunchKernelGGL(k_tri_check, dim3(nblocks(nt)), dim3(kB), 0, s, (const int3*)tri.get<int3>(), nt,
                           (const float2*)m.pos, m.periodic, bad.get<int>());
        MCL_HIP(hipGetLastError());
        int nbad = 0;
        GB_TRY(read_int(bad.get<int>(), s, &nbad));
        if (nbad == 0) break;
        if (torus) return MCL_ERR_INPUT;  // rows too sparse for the torus rows (a row gap or x gap >= 1/2)
        if (bulge == 0) return MCL_ERR_TOPOLOGY;
    }
    // 3. half-edges: pair the twins by one sort of the undirected edge keys
    const int ns = 3 * nt + (torus ? 0 : 4 * ninst);  // = nhe
    DevBuf ek, ek2, es, es2, hs;
    if ((rc = ek.alloc(8 * (size_t)ns)) || (rc = ek2.alloc(8 * (size_t)ns)) || (rc = es.alloc(4 * (size_t)ns)) ||
        (rc = es2.alloc(4 * (size_t)ns)) || (rc = hs.alloc(4 * (size_t)ns)))
        return rc;
    hipLaunchKernelGGL(k_edge_keys, dim3(nblocks(ns)), dim3(kB), 0, s, (const int3*)tri.get<int3>(), nt, ns,
                       ek.get<unsigned long long>(), es.get<int>());
    GB_TRY(radix_pairs(ek.get<unsigned long long>(), ek2.get<unsigned long long>(), es.get<int>(), es2.get<int>(), ns,
                       32 + bit_width(m.nv - 1), s));
    MCL_HIP(hipMemsetAsync(bad.p, 0, 8, s));
    hipLaunchKernelGGL(k_pair, dim3(nblocks(ns)), dim3(kB), 0, s, (const unsigned long long*)ek2.get<unsigned long long>(),
                       (const int*)es2.get<int>(), (const int3*)tri.get<int3>(), nt, ns, hs.get<int>(), bad.get<int>());
    int nbad = 0;
    GB_TRY(read_int(bad.get<int>(), s, &nbad));
    if (nbad) return MCL_ERR_TOPOLOGY;
    hipLaunchKernelGGL(k_fill_mesh, dim3(nblocks(ns)), dim3(kB), 0, s, m, (const int3*)tri.get<int3>(), nt, ns,
                       (const int*)hs.get<int>(), bad.get<int>());
    MCL_HIP(hipGetLastError());
    GB_TRY(read_int(bad.get<int>(), s, &nbad));
    return nbad ? MCL_ERR_TOPOLOGY : MCL_OK;
}

}  // namespace mcl
