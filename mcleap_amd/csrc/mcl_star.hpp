// mcl_star.hpp -- kinetic step on the torus by star recomputation.
//
// After the move, each vertex recomputes its star (its Delaunay neighbours in
// CCW order) from candidates taken in the previous mesh, then the stars are
// checked for consistency and the half-edge arrays are rebuilt from them.
//
// Star of v from a candidate set C: under the inversion p -> (p - v)/|p - v|^2
// circles through v become lines, so u is a Delaunay neighbour of v (an empty
// circle through v and u exists) exactly when u' is a vertex of the convex hull
// of C' (v inside).  The hull is wrapped counter-clockwise from the nearest
// candidate (the farthest inverted point, always a hull vertex); the turn test
// orient(a', b', c') has the sign of incircle(a, b, c, v) (divide the rows of
// the lifted determinant by |p - v|^2), so every decision is the library's
// exact incircle with the symbolic tie rule.  A group of lanes (a half wave
// or a whole one) handles one vertex.  With at most one candidate per lane the
// hull is found by elimination in angle order and certified with the filter
// (k_star); otherwise, or when the certification fails, each lane finds, for
// the candidates it owns (two at most), the most clockwise other candidate as
// seen from it (its hull successor when it is a hull vertex), and the star is
// traced through those successors from the nearest candidate.
//
// If every triangle (v, s_i, s_i+1) of every star also appears in the stars of
// its other two vertices, the stars form a triangulation whose every edge is
// locally Delaunay with respect to its two triangles, i.e. the (unique)
// Delaunay triangulation.  Vertices whose star is inconsistent are recomputed
// from the 2-rings of both the previous mesh and the stars found so far (a new
// neighbour more than two edges away in the previous mesh is two edges away
// in the new stars of its other neighbours); if that does not settle it, the
// caller falls back to the flip-based step.
#pragma once
#include "mcl_geom.hpp"
#include "mcl_internal.hpp"

namespace mcl {

constexpr int kStarCap = 24;    // stored neighbours per vertex
constexpr int kCornerCap = 128; // stored fan entries per corner (points clamped onto a side all reach a corner)
constexpr int kCornerWindow = 3;  // fan entries on each side of a vertex that a corner adds to its candidates

// Star rows: vertex v >= ncorner at v * kStarCap; the corners' (longer) fans
// after all vertex rows, at nv * kStarCap + c * kCornerCap.
__host__ __device__ __forceinline__ size_t star_row(int v, int nv, int ncorner) {
    return v < ncorner ? (size_t)nv * kStarCap + (size_t)v * kCornerCap : (size_t)v * kStarCap;
}
__host__ __device__ __forceinline__ size_t star_rows_size(int nv, int ncorner) {
    return (size_t)nv * kStarCap + (size_t)ncorner * kCornerCap;
}
constexpr int kCandMax = 128;   // candidates per vertex: two per lane
constexpr int kHashSlots = 256; // per-wave LDS hash set for candidate de-duplication
constexpr int kStarPasses = 12;  // star computations per step before the flip-based fallback
constexpr int kStarAux = 24;    // star-step counters
constexpr int kStarMinVertices = 1 << 16;  // smaller meshes take the flip-based step
constexpr int kStarDecidePasses = 8;  // checks that list only the losing side of each conflict
constexpr int kStarSyncFrom = 3;  // retry passes queued before the host reads the counts

// stars of the current mesh: CCW rotation from v.he (h <- next(next(h))^1).
// A corner's star is an open fan [A, x1, ..., B] whose ends are the two
// corners next to it on the outer square (the outer face lies between B and
// A): walked clockwise, h <- next(h^1), which passes the outer face through
// its 4-cycle, and reversed so that it starts after the outer half-edge.
__global__ void __launch_bounds__(256) k_stars_from_mesh(DMesh m, int* __restrict__ deg, int* __restrict__ nbr,
                                                         int* __restrict__ fail) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.ncorner; v += gridDim.x * blockDim.x) {
        // clockwise from the outgoing half-edge that borders the outer face (v -> B)
        int h = m.v_he[v], guard = 0;
        while (m.he_t[h] != -1 && guard++ < 2 * kCornerCap) h = m.he_next[h ^ 1];
        int d = 0;
        const int h0 = h;
        bool okf = m.he_t[h] == -1;
        int* row = nbr + star_row(v, m.nv, m.ncorner);
        do {
            if (d == kCornerCap) { okf = false; break; }
            row[d++] = m.he_v[h ^ 1];  // B, xk, ..., x1, A
            h = m.he_next[h ^ 1];
        } while (okf && h != h0);
        if (okf && h == h0) {
            for (int i = 0; i < d / 2; i++) {  // reverse: A, x1, ..., xk, B
                const int a = row[i], b = row[d - 1 - i];
                row[i] = b;
                row[d - 1 - i] = a;
            }
            deg[v] = d;
        } else {
            deg[v] = -1;
            atomicAdd(fail, 1);
        }
    }
    for (int v = m.ncorner + blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        const int h0 = m.v_he[v];
        int h = h0, d = 0;
        do {
            if (d == kStarCap) break;
            nbr[(size_t)v * kStarCap + d++] = m.he_v[h ^ 1];
            h = m.he_next[m.he_next[h]] ^ 1;
        } while (h != h0);
        if (h != h0) {
            deg[v] = -1;
            atomicAdd(fail, 1);
        } else {
            deg[v] = d;
        }
    }
}

// inserts c into the LDS hash set (kH slots, a power of two) and, when it is
// new, appends it to the candidate list (insertion order, so the list is in
// ring order; at most kC entries are stored, the count keeps going)
template <int kH, int kC>
__device__ __forceinline__ void cand_insert(int* tab, int* list, int* count, int c) {
    constexpr int kBits = kH == 128 ? 7 : 8;
    static_assert((1 << kBits) == kH, "hash table size");
    unsigned h = ((unsigned)c * 2654435761u) >> (32 - kBits);
    for (int probe = 0; probe < kH; probe++) {
        const int old = atomicCAS(&tab[h], -1, c);
        if (old == -1) {
            const int k = atomicAdd(count, 1);
            if (k < kC) list[k] = c;
            return;
        }
        if (old == c) return;
        h = (h + 1) & (kH - 1);
    }
    atomicAdd(count, kC + 1);  // table full: reported as too many candidates
}

// inserts the star of u (from one star array) into the candidate set of v;
// of a corner's fan only the entries within kCornerWindow + 1 of v
template <int kH, int kC>
__device__ __forceinline__ void insert_star(int* tab, int* list, int* count, const int* __restrict__ deg,
                                            const int* __restrict__ nbr, int u, int v, int nv, int ncorner) {
    const int du = deg[u];
    const int* row = nbr + star_row(u, nv, ncorner);
    if (u < ncorner) {
        if (du < 1 || du > kCornerCap) return;
        int p = -1;
        for (int k = 0; k < du; k++)
            if (row[k] == v) p = k;
        if (p < 0) return;
        for (int k = -kCornerWindow - 1; k <= kCornerWindow + 1; k++) {
            const int c = row[((p + k) % du + du) % du];
            if (c != v && (unsigned)c < (unsigned)nv) cand_insert<kH, kC>(tab, list, count, c);
        }
        return;
    }
    for (int k = 0; k < du && k < kStarCap; k++) {
        const int c = row[k];
        if (c != v && (unsigned)c < (unsigned)nv) cand_insert<kH, kC>(tab, list, count, c);
    }
}

// position of x in a star row of du entries, -1 if absent; the row is read as
// 16-byte vectors (rows are kStarCap = 24 or kCornerCap = 128 ints, 16-byte
// aligned)
__device__ __forceinline__ int row_find(const int* __restrict__ row, int du, int x) {
    const int4* r = reinterpret_cast<const int4*>(row);
    int pos = -1;
    for (int k = 0; k < du; k += 4) {
        const int4 q = r[k >> 2];
        if (q.w == x && k + 3 < du) pos = k + 3;
        if (q.z == x && k + 2 < du) pos = k + 2;
        if (q.y == x && k + 1 < du) pos = k + 1;
        if (q.x == x) pos = k;
    }
    return pos;
}
// position of x in the star of vertex u >= ncorner
__device__ __forceinline__ int star_find(const int* __restrict__ nbr, int u, int du, int x) {
    return row_find(nbr + (size_t)u * kStarCap, du, x);
}

constexpr float kStarIccBound = 16.0f * 5.9604644775390625e-08f;  // the scan's incircle filter, x permanent

constexpr int kStarMaxDepth = 5;  // deepest ring of the previous mesh searched
// candidate sets of k_star
constexpr int kCandHash = 0;   // rings de-duplicated in an LDS hash set (any depth)
constexpr int kCandRing2 = 1;  // the 2-ring read off the previous (consistent) stars
constexpr int kCandRaw = 2;    // old and new first rings with their new stars, copies kept

// The FP32 stage of incircle(a, b, c, v) from precomputed rows (x, y, lift)
// of a - v, b - v, c - v: the differences rounded to FP32 from their exact
// values (star_diff) and lift = x * x + y * y in FP32, the operations and
// bound of the filter inside incircle(), so a decided sign is the exact sign.
// Returns +-1, or 0 when the filter cannot decide.
__device__ __forceinline__ int star_incircle_f32(float4 A, float4 B, float4 C) {
#pragma clang fp contract(off)
    const float e = 5.9604644775390625e-08f;
    const float bdxcdy = B.x * C.y, cdxbdy = C.x * B.y;
    const float cdxady = C.x * A.y, adxcdy = A.x * C.y;
    const float adxbdy = A.x * B.y, bdxady = B.x * A.y;
    const float det = A.z * (bdxcdy - cdxbdy) + B.z * (cdxady - adxcdy) + C.z * (adxbdy - bdxady);
    const float perm = (fabsf(bdxcdy) + fabsf(cdxbdy)) * A.z + (fabsf(cdxady) + fabsf(adxcdy)) * B.z +
                       (fabsf(adxbdy) + fabsf(bdxady)) * C.z;
    const float bound = (10.0f + 96.0f * e) * e * perm * 1.0001f;
    const bool sure = (perm > 7.9e-31f) & (perm < 1e37f) & (fabsf(det) > bound);
    return sure ? (det > 0.0f ? 1 : -1) : 0;
}

// The FP32 stage of orient(v, a, b) from the rows of a - v and b - v (the
// bound of the orient filter, whose differences carry the same relative
// error as these rows).  Returns +-1, or 0 when the filter cannot decide.
__device__ __forceinline__ int star_orient_f32(float4 A, float4 B) {
#pragma clang fp contract(off)
    const float e32 = 5.9604644775390625e-08f;
    const float dl = A.x * B.y, dr = A.y * B.x, det = dl - dr;
    const float ds = fabsf(dl) + fabsf(dr);
    const float bound = (3.0f + 16.0f * e32) * e32 * ds * 1.0001f;
    if (ds > 7.9e-31f && (det > bound || -det > bound)) return det > 0 ? 1 : -1;
    return 0;
}

// Lane masks of one group of kL lanes (a half wave or a whole one) for the
// hull elimination: the group's bits of a wave ballot, the nearest set bit
// cyclically before and after a lane (the lane itself when none is set).
template <int kL> struct HullMask;
template <> struct HullMask<32> {
    using type = unsigned;
    static __device__ __forceinline__ unsigned group(unsigned long long b, int gbase) { return (unsigned)(b >> gbase); }
    static __device__ __forceinline__ int popc(unsigned m) { return __popc(m); }
    static __device__ __forceinline__ int lowest(unsigned m) { return __ffs(m) - 1; }
    static __device__ __forceinline__ int before(unsigned m, unsigned lo, int lane) {
        const unsigned b = m & lo;
        return m ? (b ? 31 - __clz(b) : 31 - __clz(m)) : lane;
    }
    static __device__ __forceinline__ int after(unsigned m, unsigned hi, int lane) {
        const unsigned a = m & hi;
        return m ? (a ? __ffs(a) - 1 : __ffs(m) - 1) : lane;
    }
};
template <> struct HullMask<64> {
    using type = unsigned long long;
    static __device__ __forceinline__ unsigned long long group(unsigned long long b, int) { return b; }
    static __device__ __forceinline__ int popc(unsigned long long m) { return __popcll(m); }
    static __device__ __forceinline__ int lowest(unsigned long long m) { return __ffsll((long long)m) - 1; }
    static __device__ __forceinline__ int before(unsigned long long m, unsigned long long lo, int lane) {
        const unsigned long long b = m & lo;
        return m ? (b ? 63 - __clzll(b) : 63 - __clzll(m)) : lane;
    }
    static __device__ __forceinline__ int after(unsigned long long m, unsigned long long hi, int lane) {
        const unsigned long long a = m & hi;
        return m ? (a ? __ffsll((long long)a) - 1 : __ffsll((long long)m) - 1) : lane;
    }
};

// b - a (minimum image when periodic) rounded to FP32: the FP32 subtraction
// rounds the exact difference once, and is that value whenever no image shift
// can be taken (|b - a| < 0.4999 is certainly below 0.5); otherwise the exact
// expansion, shifted, is rounded (through FP64).  Either way the row carries a
// relative error of at most about one FP32 rounding, which the filters'
// bounds allow for.
__device__ __forceinline__ float star_diff(float b, float a, bool periodic) {
#pragma clang fp contract(off)
    const float d = b - a;
    if (!periodic || fabsf(d) < 0.4999f) return d;
    return (float)dexp_value(diff_exact(b, a, 0, true));
}

// incircle(a, b, c, v) from the filter rows A, B, C of a - v, b - v, c - v in
// the successor scan's evaluation order (minors of the x, y columns, expanded
// along the lifts with FMAs): det has the sign of orient(a', b', c') of the
// inverted points, and is that sign exactly when sure (kStarIccBound covers
// this order: about 11 eps of the permanent from input, product, difference
// and FMA roundings).
__device__ __forceinline__ bool star_icc_sure(float4 A, float4 B, float4 C, float& det) {
    const float axby = A.x * B.y, bxay = B.x * A.y;
    const float bxcy = B.x * C.y, cxby = C.x * B.y;
    const float cxay = C.x * A.y, axcy = A.x * C.y;
    const float m1 = bxcy - cxby, m2 = cxay - axcy, m3 = axby - bxay;
    const float s1 = fabsf(bxcy) + fabsf(cxby), s2 = fabsf(cxay) + fabsf(axcy), s3 = fabsf(axby) + fabsf(bxay);
    det = fmaf(A.z, m1, fmaf(B.z, m2, C.z * m3));
    const float perm = fmaf(A.z, s1, fmaf(B.z, s2, C.z * s3));
    return (perm > 7.9e-31f) & (fabsf(det) > kStarIccBound * perm);
}

// Orders the LDS accesses of one wave: k_star's per-vertex LDS data belong to
// a group of 32 or 64 lanes, which lies inside one wave, and the LDS
// operations of a wave execute in order; the fences keep the compiler from
// moving accesses across.  (A block barrier would also make the block's
// other waves, working on other vertices, wait for each other.)
__device__ __forceinline__ void group_sync() { wave_sync(); }

// Recomputes the stars of the listed vertices (all vertices when vlist is
// null) with the current positions.  Candidates: the 2-ring of v in the star
// arrays deg_a/nbr_a (the previous mesh).  With deg_b/nbr_b (the stars found
// so far in this step): v's neighbours in both, expanded through deg_b, or,
// when v's star in deg_b failed (v moved out of its 2-ring), the depth-ring
// of v in the previous stars plus the 2-ring through deg_b.  Processes list
// positions 0..nlist-1 (the count read from *nlist_dev when given), or the
// *nrows positions in rows when given (vertices vlist[row], or row itself
// without a list).  Writes deg_out/nbr_out at row idx of the list (at row v
// without a list); a vertex whose star cannot be computed gets deg_out = -1.
// With heavy given, a vertex that failed for too many candidates or an
// undecided test is appended to heavy (heavy[0] the count, the rows after it;
// counted in fail[10]) for the full kernel; other failures are counted in
// fail[0] and, by reason, in fail[4] (too few or too many candidates),
// fail[5] (undecided test), fail[6] (hull not closed) or fail[7] (v not
// inside).  acc: the positions at the start of the step (of vertices found
// on one point, all but the smallest id go back there, counted as
// duplicates when that vertex is at its own target tgt, else as stalled;
// null: they just fail).  reset (may be null): a counter this launch zeroes
// for a later kernel.  give_up >= 0: a list longer than that is not
// processed; the list count *nlist_dev is set to 0 and *gave_up to 1.
// hint (may be null; retries): a further first-ring candidate per vertex
// (-1: none), left by the consistency check.
//
// kL lanes handle one vertex (64: one per wave, up to 128 candidates; 32: two
// per wave, up to 64 candidates, which holds nearly every 2-ring and halves
// the instructions per vertex).  The loop over vertices advances 256 / kL
// vertices per block iteration; a vertex's LDS data stay inside its wave, so
// its stages are ordered by wave-level syncs (group_sync), and each wave
// moves on without waiting for the block's others.  kFast: FP32 filter only,
// a test it cannot decide fails the
// vertex (reason 5) for a later pass with the full cascade; this variant
// carries none of the exact-arithmetic code and keeps a small register
// footprint.
constexpr int kStarBlock = 128;  // threads per block of k_star (two waves; 256 and 64 measured slower)
template <bool kF32, bool kFast, int kL, int kCand>
__global__ void __launch_bounds__(kStarBlock) k_star(DMesh m, const int* __restrict__ vlist, int nlist,
                                              const int* __restrict__ nlist_dev, const int* __restrict__ rows,
                                              const int* __restrict__ nrows, int depth,
                                              const int* __restrict__ deg_a, const int* __restrict__ nbr_a,
                                              const int* __restrict__ deg_b, const int* __restrict__ nbr_b,
                                              int* __restrict__ deg_out, int* __restrict__ nbr_out,
                                              int* __restrict__ fail, int* __restrict__ heavy,
                                              const float2* __restrict__ acc, const float2* __restrict__ tgt,
                                              unsigned long long* __restrict__ stats, int* __restrict__ reset,
                                              int give_up, int* __restrict__ gave_up, const int* __restrict__ hint) {
    static_assert(kL == 32 || kL == 64, "lanes per vertex");
    static_assert(kCand != kCandRaw || kL == 64, "the raw gather takes a first ring of up to 49 entries, one per lane");
    constexpr int kV = kStarBlock / kL;             // vertices per block iteration
    constexpr int kC = kL == 64 ? kCandMax : 64;    // candidates per vertex
    constexpr int kH = 2 * kC;                      // hash slots per vertex
    // the successor scan (below) is compiled into the full kernel and the
    // hash-gather filter kernel; the others hand what the hull elimination
    // leaves to a later kernel
    constexpr bool kScan = !(kFast && kCand != kCandHash);
    __shared__ int s_tab[kV][kH];
    __shared__ int s_cand[kV][kC];
    __shared__ float2 s_pos[kV][kFast ? 1 : kC];  // the candidates' positions (exact predicates)
    __shared__ float4 s_row[kV][kC];  // (x, y, lift) of candidate - v, candidate id (bits)
    __shared__ int s_succ[kV][kC];
    __shared__ __align__(16) int s_key[kV][kL];  // angle keys of the hull elimination
    __shared__ int s_cnt[kV];
    const int g = threadIdx.x / kL, lane = threadIdx.x % kL;
    const int gbase = (threadIdx.x & 63) - lane;  // first wave lane of this vertex's group
    const unsigned long long gmask = kL == 64 ? ~0ull : (0xffffffffull << gbase);
    int n = rows ? *nrows : (vlist ? (nlist_dev ? *nlist_dev : nlist) : m.nv);
    if (give_up >= 0 && n > give_up) {
        // every block reads the count before or after block 0 clears it, and
        // either way takes it as 0
        n = 0;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            *const_cast<int*>(nlist_dev) = 0;
            *gave_up = 1;
        }
    }
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0;
    const bool periodic = m.periodic != 0;
    const int max_depth = depth < 2 ? 2 : (depth > kStarMaxDepth ? kStarMaxDepth : depth);
    unsigned loc[4] = {0, 0, 0, 0};
    unsigned stalled[2] = {0, 0};  // stalled, duplicates
    int first, end, stride;
    xcd_split(n, kV, first, end, stride);
    for (int base = first; base < end; base += stride) {
        const int idx0 = base + g < end ? (rows ? rows[base + g] : base + g) : 0;  // output row
        const int v0 = base + g < end ? (vlist ? vlist[idx0] : idx0) : 0;
        // corners get no star here (theirs are derived from their neighbours')
        const bool active = base + g < end && v0 >= m.ncorner;
        const int idx = active ? idx0 : 0;
        const int v = active ? v0 : 0;
        int* tab = s_tab[g];
        int* list = s_cand[g];
        int* count = &s_cnt[g];
        int kc;
        if (kCand == kCandRing2) {
            // ---- candidates straight from the previous (consistent) stars:
            // v's star, then for each neighbour u_i the part of u_i's star
            // behind it.  In u_i's star v is followed by u_i-1 and preceded by
            // u_i+1, whose predecessor there is the vertex opposite edge
            // (u_i, u_i+1), listed by u_i+1 as the first entry behind it; so
            // u_i contributes the entries from two after v to three before v.
            // A vertex can still come twice (two non-consecutive neighbours);
            // the hull step compares candidates by id, so that is harmless.
            // A corner's fan adds only the kCornerWindow entries on either side
            // of that range (its far entries are hull points far from v).
            const int d = active ? deg_a[v] : 0;
            int u = -1, du = 0, p = -1, cnt = 0;
            const int* urow = nbr_a;
            if (lane < d) {
                u = nbr_a[(size_t)v * kStarCap + lane];
                du = deg_a[u];
                urow = nbr_a + star_row(u, m.nv, m.ncorner);
                if (u < m.ncorner) {
                    p = (du >= 1 && du <= kCornerCap) ? row_find(urow, du, v) : -1;
                    cnt = p >= 0 && du > 4 ? min(du - 4, 2 * kCornerWindow) : 0;
                } else {
                    p = row_find(urow, du, v);
                    cnt = p >= 0 ? (du > 4 ? du - 4 : 0) : du;
                }
            }
            int x = cnt;  // inclusive prefix over the group's lanes
            for (int off = 1; off < kL; off <<= 1) {
                const int y = __shfl_up(x, off);
                if (lane >= off) x += y;
            }
            kc = d + __shfl(x, gbase + kL - 1);
            if (lane < d && lane < kC) list[lane] = u;
            int at = d + x - cnt;
            const bool fan = u >= 0 && u < m.ncorner;
            for (int k = 0; k < cnt; k++, at++) {
                int q = k;
                if (fan) {
                    q = k < kCornerWindow ? p + 2 + k : p - 3 - (k - kCornerWindow);
                    q = (q % du + du) % du;
                } else if (p >= 0) {
                    q = p + 2 + k;
                    if (q >= du) q -= du;
                }
                if (at < kC) list[at] = urow[q];
            }
            group_sync();
        } else if (kCand == kCandRaw) {
            // ---- v's old and new first rings (and the check's hint), each
            // with its star in this step's stars (a corner: the entries of its
            // previous fan around v): the hash-set candidates, gathered without
            // atomics.  Repeats are dropped twice, in the first ring and in the
            // gathered list, by an LDS table indexed by a hash of the id: every
            // entry writes its index to its slot, and after a barrier an entry
            // whose slot holds another entry with the same id drops out (a slot
            // shared by two ids only leaves repeats, which the hull elimination
            // drops).  A vertex whose own star failed (it left its 2-ring) is
            // left to the full kernel, which searches deeper rings.  (One wave
            // per vertex: at most 49 first-ring entries, one per lane.)
            constexpr int kBits = kH == 128 ? 7 : 8;
            static_assert((1 << kBits) == kH, "hash table size");
            const int da = active ? min(deg_a[v], kStarCap) : 0;
            const int db0 = active ? deg_b[v] : 0;
            const bool widen = active && db0 < 0;
            const int db = widen ? 0 : min(db0, kStarCap);
            int u = -1;
            if (lane < da) u = nbr_a[(size_t)v * kStarCap + lane];
            else if (lane < da + db) u = nbr_b[(size_t)v * kStarCap + lane - da];
            else if (lane == da + db && hint && active) u = hint[v];
            if ((unsigned)u >= (unsigned)m.nv || u == v) u = -1;
            const unsigned hu = ((unsigned)u * 2654435761u) >> (32 - kBits);
            if (u >= 0) tab[hu] = lane;
            group_sync();
            {
                const int w = u >= 0 ? tab[hu] : lane;
                const int uw = __shfl(u, gbase + w);
                if (w != lane && uw == u) u = -1;
            }
            int du = 0, p = -1, cnt = 0;
            const int* urow = nbr_b;
            int e[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // the row's first 8 entries, loaded at once
            const bool fan = u >= 0 && u < m.ncorner;
            if (fan) {
                du = deg_a[u];
                urow = nbr_a + star_row(u, m.nv, m.ncorner);
                p = (du >= 1 && du <= kCornerCap) ? row_find(urow, du, v) : -1;
                cnt = 1 + (p >= 0 ? 2 * kCornerWindow + 3 : 0);
            } else if (u >= 0) {
                urow = nbr_b + (size_t)u * kStarCap;
                const int4 q0 = reinterpret_cast<const int4*>(urow)[0];
                const int4 q1 = reinterpret_cast<const int4*>(urow)[1];
                du = min(deg_b[u], kStarCap);
                e[0] = q0.x; e[1] = q0.y; e[2] = q0.z; e[3] = q0.w;
                e[4] = q1.x; e[5] = q1.y; e[6] = q1.z; e[7] = q1.w;
                cnt = 1 + (du > 0 ? du : 0);
            }
            int x = cnt;  // inclusive prefix over the group's lanes
            for (int off = 1; off < kL; off <<= 1) {
                const int y = __shfl_up(x, off);
                if (lane >= off) x += y;
            }
            const int raw = __shfl(x, gbase + kL - 1);
            int at = x - cnt;
            if (cnt > 0 && at < kC) list[at] = u;
            if (!fan && du <= 8) {
#pragma unroll
                for (int k = 0; k < 8; k++)
                    if (k < du && at + 1 + k < kC) list[at + 1 + k] = e[k];
            } else {
                for (int k = 1; k < cnt; k++) {
                    const int q = fan ? ((p + k - kCornerWindow - 2) % du + du) % du : k - 1;
                    if (at + k < kC) list[at + k] = urow[q];
                }
            }
            group_sync();
            // the gathered entries, two per lane: ids in range other than v,
            // one copy of each
            constexpr int kPer = kC / kL;
            int c[kPer];
            unsigned hc[kPer];
#pragma unroll
            for (int h = 0; h < kPer; h++) {
                const int j = lane + kL * h;
                c[h] = (j < raw && j < kC && !widen) ? list[j] : -1;
                if ((unsigned)c[h] >= (unsigned)m.nv || c[h] == v) c[h] = -1;
                hc[h] = ((unsigned)c[h] * 2654435761u) >> (32 - kBits);
                if (c[h] >= 0) tab[hc[h]] = j;
            }
            group_sync();
            unsigned long long keep[kPer];
#pragma unroll
            for (int h = 0; h < kPer; h++) {
                const int j = lane + kL * h;
                if (c[h] >= 0) {
                    const int w = tab[hc[h]];
                    if (w != j && list[w] == c[h]) c[h] = -1;
                }
                keep[h] = __ballot(c[h] >= 0) & gmask;
            }
            group_sync();
            int before = 0;
#pragma unroll
            for (int h = 0; h < kPer; h++) {
                const unsigned long long mine = keep[h] & ((1ull << (threadIdx.x & 63)) - 1ull);
                if (c[h] >= 0) list[before + __popcll(mine)] = c[h];
                before += __popcll(keep[h]);
            }
            // more gathered entries than the list holds, or a widening vertex:
            // full kernel (no candidate count, so that no row is read from
            // entries the squeeze did not write)
            kc = (raw > kC || widen) ? 0 : before;
            group_sync();
        } else {
            // ---- candidates, de-duplicated in LDS, listed ring by ring
            for (int k = 0; k < kH / kL; k++) tab[lane + kL * k] = -1;
            if (lane == 0) *count = 0;
            group_sync();
            const int db = (active && deg_b) ? deg_b[v] : 0;
            // a vertex whose star failed (it left its 2-ring) widens through the
            // previous stars; a valid but inconsistent one expands its first ring
            // (old and new neighbours) through the new stars only
            const bool widen = active && deg_b && db < 0;
            const int my_depth = widen ? max_depth : 2;
            // first ring, from the previous mesh and from this step's stars
            if (active) {
                const int da = deg_a[v];
                for (int i = lane; i < da && i < kStarCap; i += kL) {
                    const int ni = nbr_a[(size_t)v * kStarCap + i];
                    if ((unsigned)ni < (unsigned)m.nv) cand_insert<kH, kC>(tab, list, count, ni);
                }
                for (int i = lane; i < db && i < kStarCap; i += kL) {
                    const int ni = nbr_b[(size_t)v * kStarCap + i];
                    if ((unsigned)ni < (unsigned)m.nv) cand_insert<kH, kC>(tab, list, count, ni);
                }
                // a vertex whose star lists v while v's did not (left by a check)
                if (hint && lane == 0) {
                    const int hv = hint[v];
                    if ((unsigned)hv < (unsigned)m.nv && hv != v) cand_insert<kH, kC>(tab, list, count, hv);
                }
            }
            group_sync();
            // further rings: expand the entries the previous ring added
            int lo = 0, hi = *count;
            for (int dd = 2; dd <= max_depth; dd++) {
                if (active && dd <= my_depth) {
                    const int top = hi < kC ? hi : kC;
                    for (int i = lo + lane; i < top; i += kL) {
                        const int u = list[i];
                        // (this step's rows of the corners are not built yet: their previous ones)
                        if (!deg_b || widen || u < m.ncorner)
                            insert_star<kH, kC>(tab, list, count, deg_a, nbr_a, u, v, m.nv, m.ncorner);
                        if (deg_b && dd == 2 && u >= m.ncorner)
                            insert_star<kH, kC>(tab, list, count, deg_b, nbr_b, u, v, m.nv, m.ncorner);
                    }
                }
                group_sync();
                lo = hi;
                hi = *count;
            }
            kc = hi;
        }
        // positions of the candidates and their filter rows relative to v; a
        // candidate on top of v (filter kernels: found here)
        bool same = false;
        {
#pragma clang fp contract(off)
            const float2 pv = m.pos[v];
            for (int j = lane; active && j < kc && j < kC; j += kL) {
                const float2 pj = m.pos[list[j]];
                if (kFast) same |= pj.x == pv.x && pj.y == pv.y;
                else s_pos[g][j] = pj;
                const float x = star_diff(pj.x, pv.x, periodic);
                const float y = star_diff(pj.y, pv.y, periodic);
                s_row[g][j] = make_float4(x, y, x * x + y * y, __int_as_float(list[j]));
            }
        }
        group_sync();
        bool ok = active && kc >= 3 && kc <= kC;
        int why = ok ? 0 : 4;  // failure reason slot in fail[] (0: none)
        const float2 pv = m.pos[v];
        // ---- start: the nearest candidate (largest inverted radius), always
        // a hull vertex, where the scan starts its trace and the star is
        // written from (kernels without the scan write it from the first
        // corner in angle order: the star's rotation is arbitrary).  The full
        // kernel takes exact squared distances; the filter kernel the FP32
        // lift of the filter rows, which can pick a near-tie that is not on
        // the hull: the trace then does not close and the vertex goes to the
        // full kernel.  A candidate on top of v makes the inversion
        // meaningless.
        int best = kCandMax;
        int same_lo = 0x7fffffff;  // smallest id of a candidate on top of v (full kernel)
        if (kFast && !kScan) {
        } else if (kFast) {
            float dist = 3.4e38f;
            for (int j = lane; ok && j < kc; j += kL) {
                const float4 rj = s_row[g][j];
                if (rj.z < dist) { dist = rj.z; best = j; }
            }
            for (int off = kL / 2; off >= 1; off >>= 1) {
                const float od = __shfl_xor(dist, off);
                const int ob = __shfl_xor(best, off);
                if (od < dist || (od == dist && ob < best)) { dist = od; best = ob; }
            }
        } else {
            double dist = 1e300;
            for (int j = lane; ok && j < kc; j += kL) {
                const float2 pj = s_pos[g][j];
                if (pj.x == pv.x && pj.y == pv.y) {
                    same = true;
                    same_lo = min(same_lo, s_cand[g][j]);
                }
                const double dx = dexp_value(diff_exact(pj.x, pv.x, 0, periodic));
                const double dy = dexp_value(diff_exact(pj.y, pv.y, 0, periodic));
                const double dd = dx * dx + dy * dy;
                if (dd < dist) { dist = dd; best = j; }
            }
            for (int off = kL / 2; off >= 1; off >>= 1) {
                const double od = __shfl_xor(dist, off);
                const int ob = __shfl_xor(best, off);
                if (od < dist || (od == dist && ob < best)) { dist = od; best = ob; }
            }
        }
        if (!kFast)
            for (int off = kL / 2; off >= 1; off >>= 1) same_lo = min(same_lo, __shfl_xor(same_lo, off));
        if ((__ballot(same) & gmask) && ok) {
            // v landed exactly on another vertex (points clamped into the same
            // corner of the box): of the vertices on that point the smallest
            // id keeps it; the others take their previous positions for this
            // step, and every one of them is computed again in the next pass
            ok = false;
            why = 4;
            if (!kFast && acc && lane == 0 && same_lo < v) {  // (the filter kernel hands it to the full kernel)
                const float2 pu = m.pos[same_lo], tu = tgt[same_lo];
                m.pos[v] = acc[v];
                stalled[(pu.x == tu.x && pu.y == tu.y) ? 1 : 0]++;
            }
        }
        const int start = best;
        if (kScan && start >= kc && ok) { ok = false; why = 4; }
        const int id_start = (kScan && ok) ? __float_as_int(s_row[g][start].w) : -1;
        // ---- the hull by elimination (filter kernel, at most 32 candidates).
        // Lane r takes the candidate of rank r in angle around v (the
        // inversion keeps directions from v).  Rounds drop every remaining
        // candidate whose inverted point is not a strictly convex corner of
        // the chain of the remaining ones: it then lies in the triangle of the
        // origin and its two chain neighbours, inside the hull.  The rounds
        // use plain FP32 on inverted points and prove nothing.  What they leave
        // is certified with the filter: every corner strictly convex
        // (incircle), every wedge (v, h_i, h_i+1) strictly CCW (orient) with
        // one wedge crossing the ray y = 0, x > 0 (one turn around v), and
        // every other candidate strictly inside the triangle of the origin and
        // the corners of its wedge.  The corners are then the hull of the
        // inverted candidates and every other candidate lies strictly inside
        // it: the star the successor scan below traces, written from the same
        // start.  A vertex that fails any of this takes the scan.
        bool hull_done = false;
        int h_pos = -1, h_deg = 0, h_id = -1;
        if (kFast || kL == 64) {
            using Mask = typename HullMask<kL>::type;
            const bool try_hull = ok && kc <= kL;
            const int kc_h = try_hull ? kc : 0;
            // own candidate's key: a pseudo-angle in [-1, 3) as a monotone
            // int, the slot in its low bits (unique keys)
            constexpr int kSlotBits = kL == 64 ? 6 : 5;
            int key = 0x7fffffff;
            if (lane < kc_h) {
                const float4 R = s_row[g][lane];
                const float q = R.y * __builtin_amdgcn_rcpf(fabsf(R.x) + fabsf(R.y));
                const int b = __float_as_int(R.x >= 0.0f ? q : 2.0f - q);
                key = ((b >= 0 ? b : b ^ 0x7fffffff) & ~(kL - 1)) | lane;
            }
            static_assert((1 << kSlotBits) == kL, "slot bits");
            s_key[g][lane] = key;
            group_sync();
            int kc_w = kc_h;
            if (kL == 32) kc_w = max(kc_w, __shfl_xor(kc_w, 32));
            kc_w = __builtin_amdgcn_readfirstlane(kc_w);
            int rank = 0;
            for (int i = 0; i < kc_w; i += 4) {
                const int4 k4 = *reinterpret_cast<const int4*>(&s_key[g][i]);
                rank += (k4.x < key) + (k4.y < key) + (k4.z < key) + (k4.w < key);
            }
            if (lane >= kc_h) rank = lane;
            // forward permute: lane gbase + rank receives this lane's slot
            const int slot = __builtin_amdgcn_ds_permute((gbase + rank) << 2, lane);
            const bool in = lane < kc_h;
            float4 R = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float ix = 0.0f, iy = 0.0f;
            if (in) {
                R = s_row[g][slot];
                const float rz = __builtin_amdgcn_rcpf(R.z);
                ix = R.x * rz;
                iy = R.y * rz;
            }
            const int myid = in ? __float_as_int(R.w) : -1;
            // a second copy of a candidate (same id, next rank) starts dropped
            const int pid = __shfl(myid, gbase + (lane > 0 ? lane - 1 : 0));
            bool alive = in && !(lane > 0 && pid == myid);
            const Mask lo_mask = (Mask(1) << lane) - Mask(1), hi_mask = ~((Mask(2) << lane) - Mask(1));
            for (int round = 0; round < 64; round++) {
                const Mask M = HullMask<kL>::group(__ballot(alive), gbase);
                const int pr = HullMask<kL>::before(M, lo_mask, lane);
                const int nr = HullMask<kL>::after(M, hi_mask, lane);
                const float px = __shfl(ix, gbase + pr), py = __shfl(iy, gbase + pr);
                const float nx = __shfl(ix, gbase + nr), ny = __shfl(iy, gbase + nr);
                // reflex (or straight, or not a number): this point lies left
                // of prev -> next, on the side of the origin
                const bool drop = alive && !((nx - px) * (iy - py) - (ny - py) * (ix - px) < 0.0f);
                if (!__ballot(drop)) break;
                alive = alive && !drop;
            }
            // certification, one filter test (and one or two orientations) per lane
            const Mask M = HullMask<kL>::group(__ballot(alive), gbase);
            const int h = HullMask<kL>::popc(M);
            const int pr = HullMask<kL>::before(M, lo_mask, lane);
            const int nr = HullMask<kL>::after(M, hi_mask, lane);
            const int ps = __shfl(slot, gbase + pr), ns = __shfl(slot, gbase + nr);
            bool good = !in, turn = false;
            unsigned tests = 0;
            if (in && h >= 3) {
                const float4 P = s_row[g][ps], N = s_row[g][ns];
                float det;
                if (alive) {
                    // a corner: strictly convex, and (v, this, next) strictly CCW
                    good = star_icc_sure(P, R, N, det) && det > 0.0f && star_orient_f32(R, N) > 0;
                    turn = R.y < 0.0f && N.y >= 0.0f;
                    tests = 1;
                } else if (myid == __float_as_int(P.w) || myid == __float_as_int(N.w)) {
                    good = true;  // a second copy of a corner
                } else {
                    // strictly inside the triangle (origin, P', N')
                    good = star_icc_sure(P, N, R, det) && det > 0.0f && star_orient_f32(P, R) > 0 &&
                           star_orient_f32(R, N) > 0;
                    tests = 1;
                }
            }
            const Mask gbad = HullMask<kL>::group(__ballot(!good), gbase);
            const Mask gturn = HullMask<kL>::group(__ballot(alive && turn), gbase);
            const Mask gstart = kScan ? HullMask<kL>::group(__ballot(alive && slot == start), gbase) : M;
            hull_done = try_hull && h >= 3 && h <= kStarCap && gbad == 0 && HullMask<kL>::popc(gturn) == 1 &&
                        gstart != 0;
            if (hull_done) {
                // place in the star, counted from the start
                const int rs = HullMask<kL>::lowest(gstart);
                int pos = HullMask<kL>::popc(M & lo_mask) - HullMask<kL>::popc(M & ((Mask(1) << rs) - Mask(1)));
                if (pos < 0) pos += h;
                h_pos = alive ? pos : -1;
                h_deg = h;
                h_id = myid;
                loc[0] += tests;
            }
            // filter-kernel vertices that took the scan (one count per wave);
            // the wave-per-vertex kernels also count the vertices with more
            // candidates than lanes (fail[22]) and uncertified hulls (fail[23])
            if (kFast) {
                const unsigned long long fb = __ballot(lane == 0 && try_hull && !hull_done);
                if (fb && (threadIdx.x & 63) == 0) atomicAdd(fail + 21, (int)__popcll(fb));
                if (kL == 64 && lane == 0 && ok && !try_hull) atomicAdd(fail + 22, 1);
                if (kL == 64 && lane == 0 && try_hull && !hull_done) atomicAdd(fail + 23, 1);
            }
        }
        // the first pass (2-rings) leaves such a vertex to the full kernel:
        // without the scan below it keeps fewer registers, more waves per SIMD
        if (!kScan && ok && !hull_done) {
            ok = false;
            why = 5;
        }
        // ---- successors: for each owned candidate j, the most clockwise other
        // candidate as seen from j' (b beats a when orient(j', a', b') < 0,
        // i.e. incircle(j, a, b, v) < 0); -1 when a test was undecided
        if (!kScan) {
        } else if (kFast) {
            // Branch-free, with wave-uniform loops: both vertices of a wave
            // run to the larger candidate count, entries past a vertex's own
            // count (and the lane's own candidate) are masked out, and the
            // winner is kept by selects.
            const int kc_g = (ok && !hull_done) ? kc : 0;
            int kc_w = kc_g;
            if (kL == 32) kc_w = max(kc_w, __shfl_xor(kc_w, 32));
            kc_w = __builtin_amdgcn_readfirstlane(kc_w);
            for (int jj = 0; jj * kL < kc_w; jj++) {
                const int j = lane + jj * kL;
                const bool own = j < kc_g;
                const float4 A = s_row[g][own ? j : 0];
                const int idj = __float_as_int(A.w);
                // the first best: slot 0, or slot 1 when slot 0 is j itself
                // (two leading copies of j are left to the full kernel)
                const float4 r0 = s_row[g][0], r1 = s_row[g][1];
                const bool take1 = __float_as_int(r0.w) == idj;
                int a = take1 ? 1 : 0;
                float4 B = take1 ? r1 : r0;
                bool undecided = __float_as_int(B.w) == idj;
                // the minor and permanent part that depend on A and B only
                float m3 = A.x * B.y - B.x * A.y;
                float s3 = fabsf(A.x * B.y) + fabsf(B.x * A.y);
                for (int b = 0; b < kc_w; b++) {
                    const float4 C = s_row[g][b];
                    const int idb = __float_as_int(C.w);
                    // incircle(j, a, b, v) from the rows, FMA form; the bound
                    // 16 eps * permanent covers this evaluation order (about
                    // 11 eps from input, product, difference and FMA roundings)
                    const float bxcy = B.x * C.y, cxby = C.x * B.y;
                    const float cxay = C.x * A.y, axcy = A.x * C.y;
                    const float m1 = bxcy - cxby, m2 = cxay - axcy;
                    const float s1 = fabsf(bxcy) + fabsf(cxby), s2 = fabsf(cxay) + fabsf(axcy);
                    const float det = fmaf(A.z, m1, fmaf(B.z, m2, C.z * m3));
                    const float perm = fmaf(A.z, s1, fmaf(B.z, s2, C.z * s3));
                    const bool sure = (perm > 7.9e-31f) & (fabsf(det) > kStarIccBound * perm);
                    const bool use = (b < kc_g) & (idb != idj) & (idb != __float_as_int(B.w));
                    undecided |= use & !sure;
                    const bool win = use & sure & (det < 0.0f);
                    a = win ? b : a;
                    B.x = win ? C.x : B.x;
                    B.y = win ? C.y : B.y;
                    B.z = win ? C.z : B.z;
                    B.w = win ? C.w : B.w;
                    m3 = win ? -m2 : m3;
                    s3 = win ? s2 : s3;
                }
                if (own && kc_g > 2) loc[0] += (unsigned)(kc_g - 2);  // filter tests (duplicates aside)
                // successor, offset by kC when it is the start (same id)
                if (own) {
                    const bool to_start = __float_as_int(B.w) == id_start;
                    s_succ[g][j] = undecided ? -1 : (to_start ? a + kC : a);
                }
            }
        } else if (ok && !hull_done) {
            for (int j = lane; j < kc; j += kL) {
                const float4 rj = s_row[g][j];
                const int idj = __float_as_int(rj.w);
                int a = -1;
                float4 ra = rj;
                bool undecided = false;
                unsigned fast = 0;
                for (int b = 0; b < kc; b++) {
                    const float4 rb = s_row[g][b];
                    const int idb = __float_as_int(rb.w);
                    if (idb == idj) continue;
                    if (a < 0) {
                        a = b;
                        ra = rb;
                        continue;
                    }
                    if (idb == __float_as_int(ra.w)) continue;
                    int sg = kF32 ? star_incircle_f32(rj, ra, rb) : 0;
                    if (sg != 0) {
                        fast++;
                    } else {
                        int st = 0;
                        sg = incircle<kF32, false>(s_pos[g][j], s_pos[g][a], s_pos[g][b], pv, periodic, &st);
                        bump(loc, st == kStageExactBig ? 3 : st);
                        // 0 after the tie rule: the four points are collinear
                        // (e.g. points clamped onto one side of the box); such a
                        // candidate is not more clockwise than the one held
                        if (sg == kNeedBig) undecided = true;
                    }
                    if (sg < 0) { a = b; ra = rb; }
                }
                loc[0] += fast;
                const bool to_start = a >= 0 && __float_as_int(ra.w) == id_start;
                s_succ[g][j] = (undecided || a < 0) ? -1 : (to_start ? a + kC : a);
            }
        }
        group_sync();
        // ---- trace the hull from the start through the successors
        int my_slot = -1, d = 0;
        if (ok && !hull_done) {
            int cur = start;
            bool closed = false;
            for (;;) {
                if (d == kStarCap) break;
                if (lane == d) my_slot = cur;
                d++;
                const int nx = s_succ[g][cur];
                if (nx < 0) { why = 5; break; }
                if (nx >= kC) { closed = true; break; }  // back at the start
                cur = nx;
            }
            if (!closed || d < 3) {
                ok = false;
                if (!why) why = 6;
            }
        }
        // every triangle (v, s_i, s_i+1) strictly CCW (v inside the hull)
        if (ok && !hull_done) {
            const int nl = (lane + 1 < d) ? lane + 1 : 0;
            const int nslot = __shfl(my_slot, gbase + nl);
            bool bad = false, undecided = false;
            if (lane < d) {
                int o = kF32 ? star_orient_f32(s_row[g][my_slot], s_row[g][nslot]) : 0;
                if (o == 0 && !kFast) {
                    int st;
                    o = orient<kF32>(pv, s_pos[g][my_slot], s_pos[g][nslot], periodic, &st);
                    bad = o <= 0;
                } else {
                    undecided = o == 0;
                    bad = o < 0;
                }
            }
            if (__ballot(bad) & gmask) { ok = false; why = 7; }
            else if (__ballot(undecided) & gmask) { ok = false; why = 5; }
        }
        if (active) {
            const size_t row = (size_t)idx;
            if (hull_done) {
                if (h_pos >= 0) nbr_out[row * kStarCap + h_pos] = h_id;
                if (lane == 0) deg_out[row] = h_deg;
            } else if (ok) {
                if (lane < d) nbr_out[row * kStarCap + lane] = s_cand[g][my_slot];
                if (lane == 0) deg_out[row] = d;
            } else if (lane == 0) {
                deg_out[row] = -1;
                if (heavy && (why == 4 || why == 5)) {
                    // for the full kernel: more candidates than kC, or a test
                    // the filter could not decide
                    heavy[1 + atomicAdd(heavy, 1)] = idx;
                    atomicAdd(fail + 10, 1);
                } else {
                    atomicAdd(fail, 1);
                    atomicAdd(fail + why, 1);
                }
            }
        }
    }
    flush_stats(loc, stats, kStIncircle0);
    const unsigned st1[1] = {stalled[0]}, st2[1] = {stalled[1]};
    flush_stats(st1, stats, kStStalled);
    flush_stats(st2, stats, kStRelDup);
}


// Which of two triangles on the same side of edge (v, u) -- v's (v, u, w)
// and u's (v, u, w') -- is not Delaunay: 1 when w' lies inside
// circle(v, u, w) (v's triangle loses), 0 when it does not (u's loses), -1
// when the FP32 filter cannot tell.  Rows relative to w' as in the star
// kernel (exact minimum-image differences rounded to FP32), the FP32 stage of
// incircle(v, u, w, w').
__device__ __forceinline__ int star_conflict(const float2* __restrict__ pos, int v, int u, int w, int wp,
                                             bool periodic) {
#pragma clang fp contract(off)
    const float2 q = pos[wp];
    const int ids[3] = {v, u, w};
    float4 R[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float2 p = pos[ids[k]];
        const float x = (float)dexp_value(diff_exact(p.x, q.x, 0, periodic));
        const float y = (float)dexp_value(diff_exact(p.y, q.y, 0, periodic));
        R[k] = make_float4(x, y, x * x + y * y, 0.0f);
    }
    const int sg = star_incircle_f32(R[0], R[1], R[2]);
    return sg > 0 ? 1 : (sg < 0 ? 0 : -1);
}

// adds x (one value per thread, every thread of the wave calling) to *dst
// with one atomic per wave: a counter that every thread of a large launch
// adds to would otherwise serialise the launch on its address
__device__ __forceinline__ void wave_add(long long x, unsigned long long* dst) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(dst, (unsigned long long)x);
}


// true when v's star (degree d >= 3) must be recomputed.  A triangle (v, u, w)
// of v's star is consistent when w directly precedes v in u's star.  Without
// `decide` any inconsistent triangle marks v; with it, v is marked only when
// v is missing from u's star or v's triangle loses against u's (v, u, w')
// (star_conflict), the other vertex being marked by its own check.  The
// triangles go in batches of 8 with every load of a batch issued up front:
// the neighbours' degrees and first 12 entries (all but a few stars); a
// neighbour with more entries is searched one vector at a time.  q0, q1: v's
// first 8 entries, loaded by the caller with its degree.  *missing: a
// neighbour whose star lacks v (-1 if none found).
__device__ __forceinline__ bool star_inconsistent(const int* __restrict__ deg, const int* __restrict__ nbr, int v,
                                                  int d, int4 q0, int4 q1, bool decide,
                                                  const float2* __restrict__ pos, bool periodic, int ncorner,
                                                  int* missing) {
    const int* row = nbr + (size_t)v * kStarCap;
    const int4* rv = reinterpret_cast<const int4*>(row);
    *missing = -1;
    bool bad = false;
    const int first = q0.x;
#pragma unroll 1
    for (int i0 = 0; i0 < d && !bad; i0 += 8) {
        int4 p0 = q0, p1 = q1;
        if (i0 > 0) {
            p0 = rv[i0 >> 2];
            p1 = rv[(i0 >> 2) + 1];
        }
        const int r[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
        // the entry after the batch (the apex of its last triangle)
        const int after = i0 + 8 < d ? row[i0 + 8] : first;
        int du[8];
        int4 e0[8], e1[8], e2[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            du[k] = 0;
            e0[k] = e1[k] = e2[k] = make_int4(-1, -1, -1, -1);
            if (i0 + k < d && r[k] >= ncorner) {
                const int4* ru = reinterpret_cast<const int4*>(nbr + (size_t)r[k] * kStarCap);
                du[k] = deg[r[k]];
                e0[k] = ru[0];
                e1[k] = ru[1];
                e2[k] = ru[2];
            }
        }
        unsigned conf = 0;  // triangles to decide by star_conflict
        int pv[8];          // per triangle: the entry before v in u's star
#pragma unroll
        for (int k = 0; k < 8; k++) {
            pv[k] = -1;
            const int i = i0 + k;
            if (i >= d || r[k] < ncorner) continue;  // corner rows are derived later, not looked up
            const int w = i + 1 < d ? (k < 7 ? r[k + 1] : after) : first;
            const int n = du[k];
            if (n < 3) {
                bad = true;
                continue;
            }
            int pred = -1;
            if (n <= 12) {
                const int e[12] = {e0[k].x, e0[k].y, e0[k].z, e0[k].w, e1[k].x, e1[k].y,
                                   e1[k].z, e1[k].w, e2[k].x, e2[k].y, e2[k].z, e2[k].w};
                int last = e[0];
#pragma unroll
                for (int c = 1; c < 12; c++) last = (c == n - 1) ? e[c] : last;
#pragma unroll
                for (int c = 0; c < 12; c++)
                    if (c < n && e[c] == v) pred = c > 0 ? e[c - 1] : last;
            } else {
                const int j = star_find(nbr, r[k], n, v);
                pred = j < 0 ? -1 : nbr[(size_t)r[k] * kStarCap + (j == 0 ? n - 1 : j - 1)];
            }
            pv[k] = pred;
            if (pred == w) continue;
            if (pred < 0 && *missing < 0) *missing = r[k];
            if (!decide || pred < 0) bad = true;
            else conf |= 1u << k;
        }
        // the conflicts, one at a time (the test appears once in the code)
#pragma unroll 1
        while (conf && !bad) {
            const int i = __ffs(conf) - 1;
            conf &= conf - 1;
            int u = 0, w = 0, p = -1;
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (k == i) {
                    u = r[k];
                    w = i0 + k + 1 < d ? (k < 7 ? r[k + 1] : after) : first;
                    p = pv[k];
                }
            if (star_conflict(pos, v, u, w, p, periodic) != 0) bad = true;
        }
    }
    return bad;
}

// k_star_check with one thread per vertex, for the first check of a pass
// (every vertex): the listed vertices go through an LDS buffer written out
// once per block (append_reserve_nb: the block's waves never wait for each
// other before the end), where k_star_check's lanes per vertex make the
// short lists of the later checks faster.  Checks the *nlist listed vertices (all when vlist is
// null).  Vertices that fail (or whose star failed) are appended to `list`:
// of a conflict only the side that loses the incircle test
// (star_inconsistent with `decide`), each vertex at most twice per tile (v,
// and a neighbour whose star lacks v, with hint), which the append buffer
// holds.  Without a list the degree sum of the valid stars is added to
// *sum_deg.  reset (may be null): a counter this launch zeroes for a later
// kernel.
__global__ void __launch_bounds__(256) k_star_check_all(int nv, const int* __restrict__ vlist, const int* __restrict__ nlist,
                                                    const int* __restrict__ deg, const int* __restrict__ nbr,
                                                    const float2* __restrict__ pos, int periodic, int ncorner,
                                                    int* __restrict__ mark, const int* __restrict__ stamp_base,
                                                    int stamp_off, int* __restrict__ list, int* __restrict__ count,
                                                    long long* __restrict__ sum_deg, int* __restrict__ reset,
                                                    int* __restrict__ hint) {
    // (a quarter of the LDS of the default buffer, for more waves per SIMD:
    // in the first check some 8% of the vertices fail)
    __shared__ AppendBuf<2048> ab;
    const int stamp = *stamp_base + stamp_off;
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0;  // a counter zeroed for a later kernel
    append_init(ab);
    long long my_sum = 0;
    const int n = vlist ? *nlist : nv;
    int first, end, stride;
    xcd_split(n, blockDim.x, first, end, stride);
    for (int base = first; base < end; base += stride) {
        const int idx = base + threadIdx.x;
        const int v0 = idx < end ? (vlist ? vlist[idx] : idx) : 0;
        const bool active = idx < end && v0 >= ncorner;  // corner stars are derived, not checked
        const int v = active ? v0 : 0;
        const int* row = nbr + (size_t)v * kStarCap;
        // v's degree and first 8 entries, loaded together
        const int4 q0 = reinterpret_cast<const int4*>(row)[0], q1 = reinterpret_cast<const int4*>(row)[1];
        const int d = active ? deg[v] : 3;
        bool bad = active && d < 3;
        int missing = -1;
        if (active && !bad) {
            if (!vlist) my_sum += d;
            bad = star_inconsistent(deg, nbr, v, d, q0, q1, true, pos, periodic != 0, ncorner, &missing);
        }
        // claim v once per stamp, then append it.  With hint given, a
        // neighbour u whose star lacks v is listed too, with v left as a
        // candidate for it: when v has moved into u's star from beyond u's
        // 2-ring, and u's other neighbours (recomputed without v as well)
        // agree with u, no other check lists u, and u's 2-ring in the new
        // stars does not reach v either.
        bool claim_v = false, claim_m = false;
        if (bad) {
            claim_v = atomicExch(&mark[v], stamp) != stamp;
            if (hint && missing >= ncorner) {
                hint[missing] = v;
                claim_m = atomicExch(&mark[missing], stamp) != stamp;
            }
        }
        const AppendSlot sl = append_reserve_nb(ab, (claim_v ? 1 : 0) + (claim_m ? 1 : 0), count);
        if (claim_v) append_put(ab, sl, 0, v, list);
        if (claim_m) append_put(ab, sl, claim_v ? 1 : 0, missing, list);
    }
    append_finish_nb(ab, count, list);
    wave_add(my_sum, (unsigned long long*)sum_deg);
}

// Checks the *nlist listed vertices (all when vlist is null), kStarCheckLanes
// lanes per vertex: lane i takes the triangles (v, u_i, u_i+1) of v's star
// for i = lane, lane + 8, ...  A triangle is consistent when u_i+1 directly
// precedes v in u_i's star (corners are skipped: their rows are derived
// later).  A vertex fails when its star failed (degree < 3), a neighbour's
// star failed, a neighbour's star lacks v, or, with mark_nbrs set, on any
// inconsistent triangle; without it (decide) a triangle whose neighbour's
// star holds another vertex w' before v is decided by one FP32 incircle
// test (star_conflict): only the side whose triangle is not Delaunay fails
// (undecided: v fails), the other side failing in its own check.  Failed
// vertices are appended to `list`, with mark_nbrs also their star
// neighbours, each once per stamp.  With hint given, the first neighbour u
// (in v's row) whose star lacks v is listed too, with hint[u] = v: when v has
// moved into u's star from beyond u's 2-ring, and u's other neighbours
// (recomputed without v as well) agree with u, no other check lists u, and
// u's 2-ring in the new stars does not reach v either.  Without a list the
// degree sum of the valid stars is added to *sum_deg.  reset (may be null):
// a counter this launch zeroes for a later kernel.
//
// Each lane loads its neighbour's degree and first 12 entries at once (all
// but a few stars), so no vertex's check waits on one load after another.
constexpr int kStarCheckLanes = 8;
constexpr int kStarCheckBuf = 2048;  // append buffer ints per block
__global__ void __launch_bounds__(256) k_star_check(int nv, const int* __restrict__ vlist, const int* __restrict__ nlist,
                                                    const int* __restrict__ deg, const int* __restrict__ nbr,
                                                    const float2* __restrict__ pos, int periodic, int ncorner,
                                                    int mark_nbrs, int* __restrict__ mark, const int* __restrict__ stamp_base,
                                                    int stamp_off, int* __restrict__ list, int* __restrict__ count,
                                                    long long* __restrict__ sum_deg, int* __restrict__ reset,
                                                    int* __restrict__ hint) {
    constexpr int kCL = kStarCheckLanes, kCV = 256 / kCL;
    __shared__ AppendBuf<kStarCheckBuf> ab;
    const int stamp = *stamp_base + stamp_off;
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0;  // a counter zeroed for a later kernel
    append_init(ab);
    const int g = threadIdx.x / kCL, lane = threadIdx.x % kCL;
    const int gbase = (threadIdx.x & 63) - lane;
    const unsigned long long gmask = ((1ull << kCL) - 1ull) << gbase;
    const bool decide = !mark_nbrs;
    long long my_sum = 0;
    const int n = vlist ? *nlist : nv;
    int first, end, stride;
    xcd_split(n, kCV, first, end, stride);
    for (int base = first; base < end; base += stride) {
        const int idx = base + g;
        const int v0 = idx < end ? (vlist ? vlist[idx] : idx) : 0;
        const bool active = idx < end && v0 >= ncorner;  // corner stars are derived, not checked
        const int v = active ? v0 : 0;
        const int* row = nbr + (size_t)v * kStarCap;
        // v's degree with its first entries (u_lane, u_lane+1)
        const int d = active ? deg[v] : 0;
        const int u_first = row[lane], w_first = row[lane + 1];
        const int r0 = __shfl(u_first, gbase);
        if (!vlist && lane == 0 && active && d >= 3) my_sum += d;
        bool bad = false;
        int miss = kStarCap;  // smallest i whose neighbour's star lacks v
        for (int i = lane; active && d >= 3 && i < d; i += kCL) {
            const int u = i < kCL ? u_first : row[i];
            const int w = i + 1 < d ? (i < kCL ? w_first : row[i + 1]) : r0;
            if (u < ncorner) continue;
            const int* urow = nbr + (size_t)u * kStarCap;
            const int4 q0 = reinterpret_cast<const int4*>(urow)[0];
            const int4 q1 = reinterpret_cast<const int4*>(urow)[1];
            const int4 q2 = reinterpret_cast<const int4*>(urow)[2];
            const int du = deg[u];
            if (du < 3) {
                bad = true;
                continue;
            }
            int pred = -1;
            if (du <= 12) {
                const int e[12] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
                int last = e[0];
#pragma unroll
                for (int k = 1; k < 12; k++) last = (k == du - 1) ? e[k] : last;
#pragma unroll
                for (int k = 0; k < 12; k++)
                    if (k < du && e[k] == v) pred = k > 0 ? e[k - 1] : last;
            } else {
                const int j = star_find(nbr, u, du, v);
                pred = j < 0 ? -1 : urow[j == 0 ? du - 1 : j - 1];
            }
            if (pred == w) continue;
            if (pred < 0) {
                miss = min(miss, i);
                bad = true;
            } else if (!decide || star_conflict(pos, v, u, w, pred, periodic != 0) != 0) {
                bad = true;
            }
        }
        // the group's verdict, and the first neighbour lacking v
        bad = active && ((d < 3) || (__ballot(bad) & gmask) != 0);
#pragma unroll
        for (int off = 1; off < kCL; off <<= 1) miss = min(miss, __shfl_xor(miss, off));
        const int missing = (bad && miss < kStarCap) ? row[miss] : -1;
        // claim v (lane 0), its neighbours (their lanes, with mark_nbrs) and
        // the neighbour missing v (lane 0) once per stamp, then append them
        int items[2 + (kStarCap + kCL - 1) / kCL];
        int nclaim = 0;
        if (bad) {
            if (lane == 0 && atomicExch(&mark[v], stamp) != stamp) items[nclaim++] = v;
            for (int i = lane; mark_nbrs && i < d; i += kCL) {
                const int u = row[i];
                if (u >= ncorner && atomicExch(&mark[u], stamp) != stamp) items[nclaim++] = u;
            }
            if (lane == 0 && hint && missing >= ncorner) {
                hint[missing] = v;
                if (atomicExch(&mark[missing], stamp) != stamp) items[nclaim++] = missing;
            }
        }
        const AppendSlot sl = append_reserve_nb(ab, nclaim, count);
        for (int k = 0; k < nclaim; k++) append_put(ab, sl, k, items[k], list);
    }
    append_finish_nb(ab, count, list);
    wave_add(my_sum, (unsigned long long*)sum_deg);
}

// Copies the recomputed stars (rows by list position, *nlist_dev of them) to
// their vertices, counts them into *retried (and the pass into the int after it),
// adds the change of the valid-star degree sum to *sum_deg, and lists for
// re-checking each recomputed vertex with its previous star neighbours (the
// only vertices whose consistency the new stars can change: a vertex that
// has a recomputed vertex in its star without being in that vertex's
// previous star was inconsistent, hence recomputed itself).  reset (may be
// null): a counter this launch zeroes for a later kernel.
__global__ void __launch_bounds__(256) k_star_scatter(const int* __restrict__ vlist, const int* __restrict__ nlist_dev,
                                                      const int* __restrict__ tdeg, const int* __restrict__ tnbr,
                                                      int* __restrict__ deg, int* __restrict__ nbr,
                                                      int* __restrict__ mark, const int* __restrict__ stamp_base,
                                                      int stamp_off, int* __restrict__ rlist,
                                                      int* __restrict__ rcount, long long* __restrict__ sum_deg,
                                                      long long* __restrict__ retried, int* __restrict__ reset) {
    __shared__ AppendBuf<kListBuf> ab;  // (what does not fit goes straight to rlist)
    const int stamp = *stamp_base + stamp_off;
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0;  // a counter zeroed for a later kernel
    append_init(ab);
    const int nlist = *nlist_dev;
    if (blockIdx.x == 0 && threadIdx.x == 0 && nlist) {
        atomicAdd((unsigned long long*)retried, (unsigned long long)nlist);
        atomicAdd((int*)(retried + 1), 1);  // passes with a non-empty list
    }
    long long delta = 0;
    for (int base = blockIdx.x * blockDim.x; base < nlist; base += gridDim.x * blockDim.x) {
        const int idx = base + threadIdx.x;
        const bool active = idx < nlist;
        const int v = active ? vlist[idx] : 0;
        const int od = active ? deg[v] : 0;
        int* row = nbr + (size_t)v * kStarCap;
        // claims: v, then the previous entries, the first 8 (nearly all) from
        // two 16-byte loads with their 8 exchanges issued together
        unsigned claimed = 0;
        int nclaim = 0;
        const int4 q0 = reinterpret_cast<const int4*>(row)[0], q1 = reinterpret_cast<const int4*>(row)[1];
        const int e[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
        if (active) {
            if (atomicExch(&mark[v], stamp) != stamp) { claimed |= 1u; nclaim++; }
            int got[8];
#pragma unroll
            for (int i = 0; i < 8; i++) got[i] = i < od ? atomicExch(&mark[e[i]], stamp) : stamp;
#pragma unroll
            for (int i = 0; i < 8; i++)
                if (got[i] != stamp) { claimed |= 2u << i; nclaim++; }
            for (int i = 8; i < od; i++)
                if (atomicExch(&mark[row[i]], stamp) != stamp) {
                    claimed |= 2u << i;
                    nclaim++;
                }
        }
        const AppendSlot sl = append_reserve_nb(ab, nclaim, rcount);
        int at = 0;
        if (claimed & 1u) append_put(ab, sl, at++, v, rlist);
        for (int i = 0; i < od && i < kStarCap; i++)
            if (claimed & (2u << i)) append_put(ab, sl, at++, row[i], rlist);
        if (active) {
            const int nd = tdeg[idx];
            delta += (nd >= 3 ? nd : 0) - (od >= 3 ? od : 0);
            // the new row: 16-byte copies of its first 8 entries (entries past
            // the degree are never read), the rest one by one
            const int4* tr = reinterpret_cast<const int4*>(tnbr + (size_t)idx * kStarCap);
            int4* rr = reinterpret_cast<int4*>(row);
            if (nd > 0) {
                const int4 t0 = tr[0], t1 = tr[1];
                rr[0] = t0;
                rr[1] = t1;
            }
            for (int k = 8; k < nd; k++) row[k] = tnbr[(size_t)idx * kStarCap + k];
            deg[v] = nd;
        }
    }
    append_finish_nb(ab, rcount, rlist);
    wave_add(delta, (unsigned long long*)sum_deg);
}

// the mark stamps of one star step: *base + 1 .. *base + kStampSpan - 1
// (the base advanced by the step's prologue)
constexpr int kStampSpan = 64;

// Corner fans from their neighbours' stars.  For a non-corner x with corner
// c at position i of its star, the triangles (x, c, s_i+1) and (x, s_i-1, c)
// put, in c's CCW fan, s_i+1 right before x and s_i-1 right after it.  Each
// x adds (x, before, after) to c's bucket (kCornerCap entries at most).
__global__ void __launch_bounds__(256) k_corner_collect(int nv, int ncorner, const int* __restrict__ deg,
                                                        const int* __restrict__ nbr, int* __restrict__ bucket,
                                                        int* __restrict__ bcount) {
    for (int x = ncorner + blockIdx.x * blockDim.x + threadIdx.x; x < nv; x += gridDim.x * blockDim.x) {
        const int d = deg[x];
        const int* row = nbr + (size_t)x * kStarCap;
        for (int i = 0; i < d; i++) {
            const int c = row[i];
            if (c >= ncorner) continue;
            const int k = atomicAdd(&bcount[c], 1);
            if (k < kCornerCap) {
                int* e = bucket + ((size_t)c * kCornerCap + k) * 3;
                e[0] = x;
                e[1] = row[i + 1 == d ? 0 : i + 1];  // before x in c's fan
                e[2] = row[i == 0 ? d - 1 : i - 1];  // after x
            }
        }
    }
}

// One thread per corner: chain the bucket into the fan [A, x1, ..., xk, B]
// (A, B corners; another corner may sit inside the fan when two corners are
// joined by a diagonal) and write it as the corner's star.  Every link must
// agree from both sides (x's successor y has x as predecessor); a corner
// whose entries do not chain into one fan of at most kCornerCap is counted in
// *fail.  Adds the corner degrees to *sum_deg.  A corner that no vertex
// lists keeps its previous fan (pdeg, pnbr) if that fan holds corners only:
// the corners of an instance with no points.
__global__ void k_corner_fans(int nv, int ncorner, const int* __restrict__ bucket, const int* __restrict__ bcount,
                              const int* __restrict__ pdeg, const int* __restrict__ pnbr, int* __restrict__ deg,
                              int* __restrict__ nbr, int* __restrict__ fail, long long* __restrict__ sum_deg) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncorner) return;
    const int k = bcount[c];
    const int* b = bucket + (size_t)c * kCornerCap * 3;  // (x, before, after) per entry
    int* row = nbr + star_row(c, nv, ncorner);
    if (k == 0) {
        const int d = pdeg[c];
        const int* prow = pnbr + star_row(c, nv, ncorner);
        bool ok = d >= 2 && d <= kCornerCap;
        for (int i = 0; ok && i < d; i++) ok = prow[i] < ncorner;
        if (ok) {
            for (int i = 0; i < d; i++) row[i] = prow[i];
            deg[c] = d;
            atomicAdd((unsigned long long*)sum_deg, (unsigned long long)d);
        } else {
            deg[c] = -1;
            atomicAdd(fail, 1);
        }
        return;
    }
    bool ok = k <= kCornerCap;
    // the first entry: its predecessor is a corner that follows no entry
    int first = -1;
    for (int i = 0; ok && i < k; i++) {
        const int before = b[i * 3 + 1];
        if (before >= ncorner) continue;
        bool follows = false;
        for (int j = 0; j < k; j++) follows |= b[j * 3 + 2] == before;
        if (follows) continue;
        if (first >= 0) ok = false;  // two fan starts
        first = i;
    }
    ok = ok && first >= 0;
    int d = 0;
    unsigned long long used_lo = 0ull, used_hi = 0ull;  // entries placed (k <= 128)
    auto used = [&](int j) { return ((j < 64 ? used_lo >> j : used_hi >> (j - 64)) & 1ull) != 0; };
    if (ok) {
        row[d++] = b[first * 3 + 1];  // A
        int cur = first;
        while (ok) {
            if (d == kCornerCap || used(cur)) { ok = false; break; }
            if (cur < 64) used_lo |= 1ull << cur;
            else used_hi |= 1ull << (cur - 64);
            const int x = b[cur * 3], after = b[cur * 3 + 2];
            row[d++] = x;
            int nx = -1;
            if (after >= ncorner) {
                for (int j = 0; j < k; j++)
                    if (b[j * 3] == after) nx = j;
                if (nx < 0 || b[nx * 3 + 1] != x) { ok = false; break; }  // links must agree
            } else {
                if (d == kCornerCap) { ok = false; break; }
                row[d++] = after;  // a corner: the end B, or the fan goes on past it
                for (int j = 0; j < k; j++)
                    if (!used(j) && b[j * 3 + 1] == after) nx = j;
                if (nx < 0) break;
            }
            cur = nx;
        }
        ok = ok && __popcll(used_lo) + __popcll(used_hi) == k;
    }
    if (ok) {
        deg[c] = d;
        atomicAdd((unsigned long long*)sum_deg, (unsigned long long)d);
    } else {
        deg[c] = -1;
        atomicAdd(fail, 1);
    }
}

// per vertex: edges it owns (neighbours with a larger id) and triangles it
// owns (both other vertices larger)
__global__ void __launch_bounds__(256) k_star_counts(int nv, int ncorner, const int* __restrict__ deg,
                                                     const int* __restrict__ nbr, int* __restrict__ ecnt,
                                                     int* __restrict__ tcnt) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
        const int d = deg[v];
        const int* row = nbr + star_row(v, nv, ncorner);
        int ne = 0, nt = 0;
        const int first = row[0];
        int u = first;
        for (int i = 0; i < d; i++) {
            const int w = (i + 1 == d) ? first : row[i + 1];
            ne += u > v;
            // a corner's fan does not close: (B, A) borders the outer face
            nt += (u > v && w > v && !(v < ncorner && i + 1 == d));
            u = w;
        }
        ecnt[v] = ne;
        tcnt[v] = nt;
    }
}

// Half-edge numbering of the rebuilt mesh: edge {o, x} (o < x) is edge
// eoff[o] + (rank of x among o's larger neighbours, in star order), its
// half-edge o -> x even and x -> o odd; triangle (o, x, y) with o its
// smallest vertex is toff[o] + (rank among o's owned triangles).
//
// Pass 1: the id of every outgoing half-edge of v (into hid, one row per
// vertex like the stars), with its origin and opposite vertex, and v.he.
__global__ void __launch_bounds__(256) k_star_hid(DMesh m, const int* __restrict__ deg, const int* __restrict__ nbr,
                                                  const int* __restrict__ eoff, int* __restrict__ hid) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        const int d = deg[v];
        const size_t rv = star_row(v, m.nv, m.ncorner);
        const int* row = nbr + rv;
        const int first = row[0];
        const int ev = eoff[v];
        int rank = 0, u = first;
        for (int i = 0; i < d; i++) {
            const int w = (i + 1 == d) ? first : row[i + 1];
            int h;
            if (u > v) {
                h = 2 * (ev + rank++);
            } else {  // owned by u: rank of v among u's larger neighbours
                const int du = deg[u];
                const int4* r = reinterpret_cast<const int4*>(nbr + star_row(u, m.nv, m.ncorner));
                int ru = 0;
                bool seen = false;
                for (int k = 0; k < du && !seen; k += 4) {
                    const int4 q = r[k >> 2];
                    const int e[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        if (seen || k + c >= du) continue;
                        if (e[c] == v) seen = true;
                        else ru += e[c] > u;
                    }
                }
                h = 2 * (eoff[u] + ru) + 1;
            }
            hid[rv + i] = h;
            if ((unsigned)h < (unsigned)m.nhe) {
                const bool outer = v < m.ncorner && i + 1 == d;  // corner -> B: the outer face
                m.he_v[h] = v;
                m.he_op[h] = outer ? -1 : w;
                if (outer) m.he_t[h] = -1;
                if (i == 0) m.v_he[v] = h;
            }
            u = w;
        }
    }
}

// Pass 2: next of every incoming half-edge of v (the triangle (v, u_i, u_i+1)
// has u_i+1 -> v followed by v -> u_i), and the triangles v owns: t.he and
// the face of their two half-edges at v.
__global__ void __launch_bounds__(256) k_star_link(DMesh m, const int* __restrict__ deg, const int* __restrict__ nbr,
                                                   const int* __restrict__ hid, const int* __restrict__ toff) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        const int d = deg[v];
        const size_t rv = star_row(v, m.nv, m.ncorner);
        const int* row = nbr + rv;
        const int* hr = hid + rv;
        const int first = row[0], h_first = hr[0];
        int t = toff[v];
        int u = first, hu = h_first;
        for (int i = 0; i < d; i++) {
            const int w = (i + 1 == d) ? first : row[i + 1];
            const int hw = (i + 1 == d) ? h_first : hr[i + 1];
            const int hin = hw ^ 1;  // w -> v (at a corner's fan end: A -> corner, outer face)
            if ((unsigned)hin < (unsigned)m.nhe) m.he_next[hin] = hu;
            if (u > v && w > v && !(v < m.ncorner && i + 1 == d)) {
                if ((unsigned)t < (unsigned)m.nt && (unsigned)hu < (unsigned)m.nhe && (unsigned)hin < (unsigned)m.nhe) {
                    m.t_he[t] = hu;
                    m.he_t[hu] = t;
                    m.he_t[hin] = t;
                }
                t++;
            }
            u = w;
            hu = hw;
        }
    }
}

// Pass 3: the face of the third half-edge of every triangle
__global__ void __launch_bounds__(256) k_star_face3(DMesh m) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < m.nt; t += gridDim.x * blockDim.x) {
        const int h = m.t_he[t];
        if ((unsigned)h >= (unsigned)m.nhe) continue;
        const int h2 = m.he_next[h];
        if ((unsigned)h2 < (unsigned)m.nhe) m.he_t[h2] = t;
    }
}

// One-ring CSR straight from the stars (offsets: the scan of the star
// degrees), in the order of the half-edge walk (clockwise from v.he, which
// the rebuild sets to v -> first star entry): first entry, then the rest of
// the star backwards.  The rows of a wave's 64 consecutive vertices are one contiguous range of
// the output: a range of at most kRingWave ints is gathered in LDS and
// written with contiguous stores (rows stored thread by thread write partial
// cache lines, several times the bytes of the list); a longer one (corner
// fans) directly.
__global__ void __launch_bounds__(256) k_star_ring_csr(int nv, int ncorner, const int* __restrict__ deg,
                                                       const int* __restrict__ nbr, const int* __restrict__ off,
                                                       int* __restrict__ out, int cap) {
    __shared__ int s_row[4][kRingWave];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // each wave stages its own rows (wave-level syncs)
    for (int bb = blockIdx.x * blockDim.x; bb < nv; bb += gridDim.x * blockDim.x) {
        const int v = bb + threadIdx.x;
        const int wb = bb + (threadIdx.x & ~63);
        int w0 = 0, wn = 0;
        if (wb < nv) {
            w0 = off[wb];
            wn = off[min(wb + 64, nv)] - w0;
        }
        const bool staged = wn <= kRingWave;
        if (v < nv) {
            const int* row = nbr + star_row(v, nv, ncorner);
            // the row's first 8 entries with the degree (rows are 16-byte aligned)
            const int4 q0 = reinterpret_cast<const int4*>(row)[0], q1 = reinterpret_cast<const int4*>(row)[1];
            const int d = deg[v];
            const int k0 = off[v];
            if (d <= 8) {
                // entry j goes to place 0 (j = 0) or d - j of the ring
                const int e[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (j < d) {
                        const int k = k0 + (j == 0 ? 0 : d - j);
                        if (staged) s_row[wv][k - w0] = e[j];
                        else if (k < cap) out[k] = e[j];
                    }
            } else {
                for (int i = 0; i < d; i++) {
                    const int k = k0 + i;
                    const int x = row[i == 0 ? 0 : d - i];
                    if (staged) s_row[wv][k - w0] = x;
                    else if (k < cap) out[k] = x;
                }
            }
        }
        wave_sync();
        if (staged)
            for (int k = lane; k < wn; k += 64)
                if (w0 + k < cap) out[w0 + k] = s_row[wv][k];
        wave_sync();
    }
}

}  // namespace mcl
