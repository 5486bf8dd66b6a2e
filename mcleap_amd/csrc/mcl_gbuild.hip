// mcl_gbuild.hip -- the exact Delaunay triangulation of a point set built on
// the GPU (in place of HostTriangulation::addDelaunayPoints and the upload,
// Host_Triangulation.h:296-300, Device_Triangulation.h:232-306).
//
//  1. Vertex order: the host build's Hilbert keys, one stable radix sort, so
//     vertex 4 + k is the same point as in mcl_hmesh_build.
//  2. A valid triangulation of the points at "row" positions.  Rows by
//     (y, x) rank: R = floor(sqrt n) rows of n / R or n / R + 1 points.  Each
//     row lies on a horizontal line (the mean y of its points, raised by one
//     ulp where needed so that rows strictly increase).  Its points keep their
//     own x, except the first and last, which sit at the global min and max x
//     (ties raised by one ulp).  Consecutive rows are joined by zipper strips,
//     and the rectangle they fill is joined to the 4 corners by one fan per
//     side.  Every triangle is strictly CCW by construction.
//  3. Half-edges from the triangle list: one radix sort of the undirected
//     edge keys pairs every half-edge with its twin.
//  4. (mcl_tri_build in mcl_tri.hip) the kinetic step moves every vertex from
//     its row position to its point -- at most about one row height -- and
//     ends on the exact Delaunay triangulation.
#include <hipcub/hipcub.hpp>

#include "mcl_geom.hpp"
#include "mcl_internal.hpp"

#define GB_TRY(x)                  \
    do {                           \
        const int rc_ = (x);       \
        if (rc_ != MCL_OK) return rc_; \
    } while (0)

namespace mcl {
namespace {

constexpr int kB = 256;

int nblocks(long long n) {
    long long b = (n + kB - 1) / kB;
    return (int)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
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

__global__ void __launch_bounds__(kB) k_minmax(const float2* __restrict__ xy, int n, uint32_t* __restrict__ mm) {
    uint32_t lx = 0xffffffffu, ly = 0xffffffffu, hx = 0, hy = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float2 p = xy[i];
        lx = min(lx, ford(p.x));
        ly = min(ly, ford(p.y));
        hx = max(hx, ford(p.x));
        hy = max(hy, ford(p.y));
    }
    atomicMin(&mm[0], lx);
    atomicMin(&mm[1], ly);
    atomicMax(&mm[2], hx);
    atomicMax(&mm[3], hy);
}

// the host build's key: 16-bit grid coordinates in double, Hilbert index
__global__ void __launch_bounds__(kB) k_hilbert_keys(const float2* __restrict__ xy, int n, float lox, float loy,
                                                     double sx, double sy, uint32_t* __restrict__ key,
                                                     int* __restrict__ idx) {
#pragma clang fp contract(off)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float2 p = xy[i];
        const uint32_t gx = (uint32_t)(((double)p.x - lox) * sx);
        const uint32_t gy = (uint32_t)(((double)p.y - loy) * sy);
        key[i] = (uint32_t)hilbert_d(1u << 16, gx, gy);
        idx[i] = i;
    }
}

// vertex base + k = point order[k] (base 4 in a box, 0 on the torus):
// target positions and (y, x) keys
__global__ void __launch_bounds__(kB) k_place(const float2* __restrict__ xy, const int* __restrict__ order, int n,
                                              int base, float2* __restrict__ target,
                                              unsigned long long* __restrict__ yx, int* __restrict__ vid) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const float2 p = xy[order[k]];
        target[base + k] = p;
        yx[k] = ((unsigned long long)ford(p.y) << 32) | ford(p.x);
        vid[k] = base + k;
    }
}

// rows: rows 0 .. extra-1 hold q + 1 points, the others q
struct Rows {
    int R, q, extra;
    __host__ __device__ int start(int r) const {
        return r < extra ? r * (q + 1) : extra * (q + 1) + (r - extra) * q;
    }
    __host__ __device__ int count(int r) const { return r < extra ? q + 1 : q; }
    __host__ __device__ int of(int j) const {
        const int big = extra * (q + 1);
        return j < big ? j / (q + 1) : extra + (j - big) / q;
    }
};

// adjacent equal (y, x) keys: duplicate points; then the (row, x) keys
__global__ void __launch_bounds__(kB) k_row_keys(const unsigned long long* __restrict__ yx, const int* __restrict__ vid,
                                                 int n, Rows rows, unsigned long long* __restrict__ rx,
                                                 int* __restrict__ rvid, int* __restrict__ dup) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const unsigned long long k = yx[j];
        if (j + 1 < n && yx[j + 1] == k) atomicAdd(dup, 1);
        rx[j] = ((unsigned long long)rows.of(j) << 32) | (k & 0xffffffffull);
        rvid[j] = vid[j];
    }
}

// one thread per row: the mean y of its points (double)
__global__ void __launch_bounds__(kB) k_row_mean(const int* __restrict__ lat, Rows rows,
                                                 const float2* __restrict__ target, float* __restrict__ ry) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows.R; r += gridDim.x * blockDim.x) {
        const int s = rows.start(r), m = rows.count(r);
        double acc = 0.0;
        for (int c = 0; c < m; c++) acc += (double)target[lat[s + c]].y;
        ry[r] = (float)(acc / m);
    }
}

// one thread: row lines strictly increasing; ry[R] = the smallest gap
__global__ void k_row_fix(Rows rows, float* __restrict__ ry) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float gap = 3.0e38f;
    for (int r = 1; r < rows.R; r++) {
        if (!(ry[r] > ry[r - 1])) ry[r] = nextafterf(ry[r - 1], 3.0e38f);
        gap = fminf(gap, ry[r] - ry[r - 1]);
    }
    ry[rows.R] = gap;
}

// one thread per row: row positions.  First point at X0, last at X1, the
// others at their own x raised to one ulp above their left neighbour; a row
// where that reaches X1 is spread evenly instead.  With bulge set (R >= 3)
// the hull gets no three collinear points: the first and last points of the
// inner rows move out along a parabola in r (zero at the end rows), and the
// bottom and top rows bow out by a parabola in x of height a quarter of
// their gap to the next row.  The caller checks every triangle and rebuilds without bulge if
// one is not strictly CCW.
__global__ void __launch_bounds__(kB) k_row_pos(const int* __restrict__ lat, Rows rows, const float* __restrict__ ry,
                                                float X0, float X1, const float2* __restrict__ target,
                                                float2* __restrict__ pos, int* __restrict__ err, int bulge) {
#pragma clang fp contract(off)
    const int R = rows.R;
    const bool bow = bulge && R >= 3;
    const float W = X1 - X0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows.R; r += gridDim.x * blockDim.x) {
        const int s = rows.start(r), m = rows.count(r);
        const float y = ry[r];
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
                const float x = X0 + (X1 - X0) * ((float)c / (float)(m - 1));
                if (!(x > prev && x < X1)) atomicAdd(err, 1);  // x range too narrow for the row
                pos[lat[s + c]] = make_float2(x, y);
                prev = x;
            }
        }
        pos[lat[s + m - 1]] = make_float2(X1, y);
        if (bow) {
            if (r > 0 && r + 1 < R) {  // side chains
                // height A: a quarter of the smallest row gap, or more where
                // float resolution at |x| would straighten the parabola, but
                // never so much that a chain step leans out by a row gap
                const float ulp = nextafterf(fmaxf(fabsf(X0), fabsf(X1)), 3.0e38f) - fmaxf(fabsf(X0), fabsf(X1));
                const float A = fminf(0.0625f * ry[R] * (float)(R - 1),
                                      fmaxf(0.25f * ry[R], 0.25f * ulp * (float)R * (float)R));
                const float b = (float)r * (float)(R - 1 - r) / (0.25f * (float)(R - 1) * (float)(R - 1));
                const float dx = A * b;
                pos[lat[s]].x = X0 - dx;
                pos[lat[s + m - 1]].x = X1 + dx;
            } else {  // bottom row down, top row up
                const float gap = r == 0 ? ry[1] - ry[0] : ry[R - 1] - ry[R - 2];
                const float sgn = r == 0 ? -1.0f : 1.0f;
                for (int c = 1; c + 1 < m; c++) {
                    const float x = pos[lat[s + c]].x;
                    const float g = (x - X0) * (X1 - x) / (0.25f * W * W);
                    pos[lat[s + c]].y = y + sgn * 0.25f * gap * g;
                }
            }
        }
    }
}

// torus rows: every point at its own x raised to one ulp above its left
// neighbour (a row is a closed loop, so no point is pinned); a row where
// that reaches 1 is spread evenly instead
__global__ void __launch_bounds__(kB) k_row_pos_torus(const int* __restrict__ lat, Rows rows,
                                                      const float* __restrict__ ry, const float2* __restrict__ target,
                                                      float2* __restrict__ pos) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows.R; r += gridDim.x * blockDim.x) {
        const int s = rows.start(r), m = rows.count(r);
        const float y = ry[r];
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

// torus strips: row r and row r+1 (mod R) bound an annulus, triangulated by
// a zipper that starts at the rung between the first points of both rows
// and runs once around (m_r + m_r+1 triangles, from triangle
// rows.start(r) + rows.start(r+1) - m0); x positions are unrolled past 1
__global__ void __launch_bounds__(kB) k_strips_torus(const int* __restrict__ lat, Rows rows,
                                                     const float2* __restrict__ pos, int3* __restrict__ tri) {
    const int R = rows.R, m0 = rows.count(0);
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < R; r += gridDim.x * blockDim.x) {
        const int rt = r + 1 == R ? 0 : r + 1;
        const int sb = rows.start(r), mb = rows.count(r), st = rows.start(rt), mt = rows.count(rt);
        int o = rows.start(r) + rows.start(r + 1) - m0;
        auto bx = [&](int i) { return i < mb ? pos[lat[sb + i]].x : pos[lat[sb + i - mb]].x + 1.0f; };
        auto tx = [&](int j) { return j < mt ? pos[lat[st + j]].x : pos[lat[st + j - mt]].x + 1.0f; };
        auto bv = [&](int i) { return lat[sb + (i < mb ? i : i - mb)]; };
        auto tv = [&](int j) { return lat[st + (j < mt ? j : j - mt)]; };
        int i = 0, j = 0;
        while (i < mb || j < mt) {
            const bool bottom = j == mt || (i < mb && bx(i + 1) <= tx(j + 1));
            if (bottom) {
                tri[o++] = make_int3(bv(i), bv(i + 1), tv(j));
                i++;
            } else {
                tri[o++] = make_int3(bv(i), tv(j + 1), tv(j));
                j++;
            }
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

// strip r between rows r and r+1 starts at triangle rows.start(r) +
// rows.start(r+1) - m0 - 2r; each is a zipper that advances along the row
// whose next point has the smaller x
__global__ void __launch_bounds__(kB) k_strips(const int* __restrict__ lat, Rows rows, const float2* __restrict__ pos,
                                               int3* __restrict__ tri) {
    const int m0 = rows.count(0);
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r + 1 < rows.R; r += gridDim.x * blockDim.x) {
        const int sb = rows.start(r), mb = rows.count(r), st = rows.start(r + 1), mt = rows.count(r + 1);
        int o = rows.start(r) + rows.start(r + 1) - m0 - 2 * r;
        int i = 0, j = 0;
        while (i < mb - 1 || j < mt - 1) {
            const bool bottom = j == mt - 1 || (i < mb - 1 && pos[lat[sb + i + 1]].x <= pos[lat[st + j + 1]].x);
            if (bottom) {
                tri[o++] = make_int3(lat[sb + i], lat[sb + i + 1], lat[st + j]);
                i++;
            } else {
                tri[o++] = make_int3(lat[sb + i], lat[st + j + 1], lat[st + j]);
                j++;
            }
        }
    }
}

// fans to the corners (0 BL, 1 BR, 2 TR, 3 TL) and the 4 corner triangles
__global__ void __launch_bounds__(kB) k_fans(const int* __restrict__ lat, Rows rows, int n, int3* __restrict__ tri) {
    const int R = rows.R, m0 = rows.count(0), mt = rows.count(R - 1), s0 = 0, sT = rows.start(R - 1);
    const int nstrip = rows.start(R - 1) + rows.start(R) - m0 - 2 * (R - 1);
    const int ob = nstrip, orr = ob + (m0 - 1), ot = orr + (R - 1), ol = ot + (mt - 1), oc = ol + (R - 1);
    auto first = [&](int r) { return lat[rows.start(r)]; };
    auto last = [&](int r) { return lat[rows.start(r) + rows.count(r) - 1]; };
    const int total = oc + 4;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < total - nstrip; k += gridDim.x * blockDim.x) {
        const int o = nstrip + k;
        int3 t;
        if (o < orr) {
            const int i = o - ob;
            t = make_int3(lat[s0 + i + 1], lat[s0 + i], 0);
        } else if (o < ot) {
            const int r = o - orr;
            t = make_int3(last(r + 1), last(r), 1);
        } else if (o < ol) {
            const int i = o - ot;
            t = make_int3(lat[sT + i], lat[sT + i + 1], 2);
        } else if (o < oc) {
            const int r = o - ol;
            t = make_int3(first(r), first(r + 1), 3);
        } else {
            const int c = o - oc;
            if (c == 0) t = make_int3(0, 1, last(0));
            else if (c == 1) t = make_int3(1, 2, last(R - 1));
            else if (c == 2) t = make_int3(2, 3, first(R - 1));
            else t = make_int3(3, 0, first(0));
        }
        tri[o] = t;
    }
    (void)n;
}

// half-edge slots: 3 per triangle (corner k: org tri[k], dst tri[k+1]), then
// the outer face BR->BL, BL->TL, TL->TR, TR->BR
__device__ __forceinline__ void slot_ends(const int3* tri, int nt, int s, int& a, int& b) {
    // (s >= 3 nt only in a box mesh: its outer face)
    if (s < 3 * nt) {
        const int3 t = tri[s / 3];
        const int k = s % 3;
        a = k == 0 ? t.x : (k == 1 ? t.y : t.z);
        b = k == 0 ? t.y : (k == 1 ? t.z : t.x);
    } else {
        const int o = s - 3 * nt;
        const int org[4] = {1, 0, 3, 2}, dst[4] = {0, 3, 2, 1};
        a = org[o];
        b = dst[o];
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
            m.he_next[h] = hslot[3 * nt + (o + 1) % 4];
            m.he_t[h] = -1;
            m.he_op[h] = -1;
        }
        atomicMin(&m.v_he[a], h);
    }
}

__global__ void __launch_bounds__(kB) k_vinit(DMesh m, float2* __restrict__ target) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        m.v_he[v] = 0x7fffffff;
        m.v_pidx[v] = v;
        if (v < 4 && !m.periodic) {
            const float lo = -10000000.f, hi = 10000000.f;  // Common_Triangulation.h:12-13
            const float2 c = make_float2(v == 1 || v == 2 ? hi : lo, v >= 2 ? hi : lo);
            m.pos[v] = c;
            target[v] = c;
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

}  // namespace

// Fills m with a valid triangulation of the points at their row positions.
// Box (m.periodic = 0): nv = n + 4, nhe = 6n + 10, nt = 2n + 2, corners
// first.  Torus: nv = n, nhe = 6n, nt = 2n, points in [0,1)^2, n >= 16.
// target receives every vertex's own position (corners included) and perm
// (may be null) the input index of vertex base + k.
int gbuild(DMesh& m, const float2* xy, int n, float2* target, int* perm, hipStream_t s) {
    const bool torus = m.periodic != 0;
    const int base = torus ? 0 : 4;
    if (n <= 0 || m.nv != n + base || m.nhe != 6 * n + (torus ? 0 : 10) || m.nt != 2 * n + (torus ? 0 : 2))
        return MCL_ERR_INPUT;
    if (torus && n < 16) return MCL_ERR_INPUT;
    int rc = MCL_OK;
    DevBuf mm, hk, hk2, hi, hi2, yx, yx2, vid, vid2, dup, ry;
    if ((rc = mm.alloc(sizeof(uint32_t) * 8))) return rc;
    uint32_t init[8] = {0xffffffffu, 0xffffffffu, 0u, 0u, 0u, 0u, 0u, 0u};
    MCL_HIP(hipMemcpyAsync(mm.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_minmax, dim3(nblocks(n)), dim3(kB), 0, s, xy, n, mm.get<uint32_t>());
    uint32_t h[8];
    MCL_HIP(hipMemcpyAsync(h, mm.p, sizeof(h), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    const float lo[2] = {funord(h[0]), funord(h[1])}, hi_[2] = {funord(h[2]), funord(h[3])};
    for (int k = 0; k < 2; k++) {
        if (!(lo[k] > -1e7f && hi_[k] < 1e7f)) return MCL_ERR_INPUT;  // outside the corners (or NaN)
        if (torus && !(lo[k] >= 0.0f && hi_[k] < 1.0f)) return MCL_ERR_INPUT;
    }
    // 1. vertex order (the host build's keys and ties)
    const uint32_t G = 1u << 16;
    const double sx = (hi_[0] > lo[0]) ? (G - 1) / ((double)hi_[0] - lo[0]) : 0.0;
    const double sy = (hi_[1] > lo[1]) ? (G - 1) / ((double)hi_[1] - lo[1]) : 0.0;
    if ((rc = hk.alloc(4 * (size_t)n)) || (rc = hk2.alloc(4 * (size_t)n)) || (rc = hi.alloc(4 * (size_t)n)) ||
        (rc = hi2.alloc(4 * (size_t)n)))
        return rc;
    hipLaunchKernelGGL(k_hilbert_keys, dim3(nblocks(n)), dim3(kB), 0, s, xy, n, lo[0], lo[1], sx, sy,
                       hk.get<uint32_t>(), hi.get<int>());
    GB_TRY(radix_pairs(hk.get<uint32_t>(), hk2.get<uint32_t>(), hi.get<int>(), hi2.get<int>(), n, 32, s));
    const int* order = hi2.get<int>();
    if (perm) MCL_HIP(hipMemcpyAsync(perm, order, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_vinit, dim3(nblocks(m.nv)), dim3(kB), 0, s, m, target);
    // 2. rows by (y, x) rank
    if ((rc = yx.alloc(8 * (size_t)n)) || (rc = yx2.alloc(8 * (size_t)n)) || (rc = vid.alloc(4 * (size_t)n)) ||
        (rc = vid2.alloc(4 * (size_t)n)) || (rc = dup.alloc(8)))
        return rc;
    hipLaunchKernelGGL(k_place, dim3(nblocks(n)), dim3(kB), 0, s, xy, order, n, base, target,
                       yx.get<unsigned long long>(), vid.get<int>());
    GB_TRY(radix_pairs(yx.get<unsigned long long>(), yx2.get<unsigned long long>(), vid.get<int>(), vid2.get<int>(),
                        n, 64, s));
    Rows rows;
    rows.R = 1;
    while ((long long)(rows.R + 1) * (rows.R + 1) <= n) rows.R++;
    rows.q = n / rows.R;
    rows.extra = n % rows.R;
    if (torus && (rows.R < 3 || rows.q < 3)) return MCL_ERR_INPUT;
    MCL_HIP(hipMemsetAsync(dup.p, 0, 8, s));
    hipLaunchKernelGGL(k_row_keys, dim3(nblocks(n)), dim3(kB), 0, s, yx2.get<unsigned long long>(), vid2.get<int>(), n,
                       rows, yx.get<unsigned long long>(), vid.get<int>(), dup.get<int>());
    int ndup = 0;
    MCL_HIP(hipMemcpyAsync(&ndup, dup.p, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    if (ndup) return MCL_ERR_DUPLICATE;
    GB_TRY(radix_pairs(yx.get<unsigned long long>(), yx2.get<unsigned long long>(), vid.get<int>(), vid2.get<int>(),
                        n, 64, s));
    const int* lat = vid2.get<int>();  // row-major: row r at rows.start(r)
    if ((rc = ry.alloc(sizeof(float) * (rows.R + 1)))) return rc;
    hipLaunchKernelGGL(k_row_mean, dim3(nblocks(rows.R)), dim3(kB), 0, s, lat, rows, (const float2*)target,
                       ry.get<float>());
    hipLaunchKernelGGL(k_row_fix, dim3(1), dim3(64), 0, s, rows, ry.get<float>());
    float X0 = lo[0], X1 = hi_[0];
    if (!(X1 > X0)) X1 = X0 + 1.0f;
    // row positions and triangles, first with the bulged hull
    const int nt = m.nt;
    DevBuf tri, ek, ek2, es, es2, hs, bad;
    if ((rc = tri.alloc(sizeof(int3) * (size_t)nt)) || (rc = bad.alloc(8))) return rc;
    for (int bulge = 1; bulge >= 0; bulge--) {
        if (torus) {
            hipLaunchKernelGGL(k_row_pos_torus, dim3(nblocks(rows.R)), dim3(kB), 0, s, lat, rows,
                               (const float*)ry.get<float>(), (const float2*)target, m.pos);
            hipLaunchKernelGGL(k_strips_torus, dim3(nblocks(rows.R)), dim3(kB), 0, s, lat, rows, (const float2*)m.pos,
                               tri.get<int3>());
        } else {
            MCL_HIP(hipMemsetAsync(dup.p, 0, 8, s));
            hipLaunchKernelGGL(k_row_pos, dim3(nblocks(rows.R)), dim3(kB), 0, s, lat, rows,
                               (const float*)ry.get<float>(), X0, X1, (const float2*)target, m.pos, dup.get<int>(),
                               bulge);
            MCL_HIP(hipMemcpyAsync(&ndup, dup.p, sizeof(int), hipMemcpyDeviceToHost, s));
            MCL_HIP(hipStreamSynchronize(s));
            if (ndup) return MCL_ERR_INPUT;
            hipLaunchKernelGGL(k_strips, dim3(nblocks(rows.R)), dim3(kB), 0, s, lat, rows, (const float2*)m.pos,
                               tri.get<int3>());
            hipLaunchKernelGGL(k_fans, dim3(nblocks(2 * (n + rows.R) + 8)), dim3(kB), 0, s, lat, rows, n,
                               tri.get<int3>());
        }
        MCL_HIP(hipMemsetAsync(bad.p, 0, 8, s));
        hipLaunchKernelGGL(k_tri_check, dim3(nblocks(nt)), dim3(kB), 0, s, (const int3*)tri.get<int3>(), nt,
                           (const float2*)m.pos, m.periodic, bad.get<int>());
        MCL_HIP(hipGetLastError());
        int nbad = 0;
        MCL_HIP(hipMemcpyAsync(&nbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
        MCL_HIP(hipStreamSynchronize(s));
        if (nbad == 0) break;
        if (torus) return MCL_ERR_INPUT;  // rows too sparse for the torus rows (a row gap or x gap >= 1/2)
        if (bulge == 0) return MCL_ERR_TOPOLOGY;
    }
    // 3. half-edges: pair the twins by one sort of the undirected edge keys
    const int ns = 3 * nt + (torus ? 0 : 4);  // = nhe
    if ((rc = ek.alloc(8 * (size_t)ns)) || (rc = ek2.alloc(8 * (size_t)ns)) || (rc = es.alloc(4 * (size_t)ns)) ||
        (rc = es2.alloc(4 * (size_t)ns)) || (rc = hs.alloc(4 * (size_t)ns)))
        return rc;
    hipLaunchKernelGGL(k_edge_keys, dim3(nblocks(ns)), dim3(kB), 0, s, (const int3*)tri.get<int3>(), nt, ns,
                       ek.get<unsigned long long>(), es.get<int>());
    GB_TRY(radix_pairs(ek.get<unsigned long long>(), ek2.get<unsigned long long>(), es.get<int>(), es2.get<int>(), ns,
                        64, s));
    MCL_HIP(hipMemsetAsync(bad.p, 0, 8, s));
    hipLaunchKernelGGL(k_pair, dim3(nblocks(ns)), dim3(kB), 0, s, (const unsigned long long*)ek2.get<unsigned long long>(),
                       (const int*)es2.get<int>(), (const int3*)tri.get<int3>(), nt, ns, hs.get<int>(), bad.get<int>());
    int nbad = 0;
    MCL_HIP(hipMemcpyAsync(&nbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    if (nbad) return MCL_ERR_TOPOLOGY;
    hipLaunchKernelGGL(k_fill_mesh, dim3(nblocks(ns)), dim3(kB), 0, s, m, (const int3*)tri.get<int3>(), nt, ns,
                       (const int*)hs.get<int>(), bad.get<int>());
    MCL_HIP(hipGetLastError());
    MCL_HIP(hipMemcpyAsync(&nbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    return nbad ? MCL_ERR_TOPOLOGY : MCL_OK;
}

}  // namespace mcl
