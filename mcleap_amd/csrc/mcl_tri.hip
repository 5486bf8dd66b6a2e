// mcl_tri.hip -- device triangulation: SoA half-edge mesh, kinetic moves,
// validity checks, Lawson flip rounds, B/C untangling, one-ring gathers.
//
// Flip rounds (replacing flip_delaunay_kernel + delonize2,
// kernels.cuh:274-300, Device_Triangulation.h:324-351):
//   test kernel  -- for each candidate edge: exact incircle of its quad;
//                   an illegal edge claims both triangles with a 64-bit
//                   atomicMax of (round | hash | edge).  The round sits in
//                   the high bits, so claims never need resetting.
//   flip kernel  -- an edge that holds both claims flips (f2to2 semantics,
//                   Common_Triangulation.h:69-127, on SoA) and queues the
//                   four edges of its quad for the next round; a loser
//                   re-queues itself.  Winners own disjoint triangle pairs,
//                   so no flip reads or writes another winner's records.
// Queues are deduplicated with per-edge round stamps.  Counters live in
// device memory in round-indexed slots (round % 3), each slot cleared by the
// kernel two steps before its next use, so no memset is needed between
// launches (safe under graph capture).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mcl_geom.hpp"
#include "mcl_internal.hpp"
#include "mcl_scan.hpp"

using namespace mcl;

namespace mcl {

constexpr int kThreads = 256;

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon et al., SC'11) keyed by the seed, counter = (id, step)
struct U4 { uint32_t x, y, z, w; };
__device__ __forceinline__ U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
        uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += W0; k1 += W1;
    }
    return {c0, c1, c2, c3};
}
// two standard normals by Box-Muller from one Philox draw
__device__ __forceinline__ float2 gauss2(uint32_t id, unsigned long long step, unsigned long long seed) {
    U4 r = philox(id, (uint32_t)step, (uint32_t)(step >> 32), 0x6d636c70u, (uint32_t)seed, (uint32_t)(seed >> 32));
    float u1 = ((r.x >> 8) + 1u) * 5.9604644775390625e-08f;  // (0, 1]
    float u2 = (r.y >> 8) * 5.9604644775390625e-08f;         // [0, 1)
    float rad = sqrtf(-2.0f * logf(u1));
    float s, c;
    sincospif(2.0f * u2, &s, &c);
    return make_float2(rad * c, rad * s);
}

__device__ __forceinline__ bool is_corner(float2 p) { return fabsf(p.x) >= 1.0e6f || fabsf(p.y) >= 1.0e6f; }

__device__ __forceinline__ float wrap01(float x) {
    float w = x - floorf(x);
    return (w >= 1.0f || w < 0.0f) ? 0.0f : w;
}
__device__ __forceinline__ float clamp_box(float x) {
    const float hi = 0.99999994039535522f;  // 1 - 2^-24
    return fminf(fmaxf(x, 0.0f), hi) + 0.0f;
}

// ---------------------------------------------------- wave-aggregated append
// slot in a list whose length is *ctr, for every lane that calls this at the
// same point: the lanes active here (the exec mask) take consecutive slots
// and one lane does a single atomic for the whole wave
__device__ __forceinline__ int agg_inc(int* ctr) {
    const unsigned long long active = __ballot(1);
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)active) - 1;
    const int rank = __popcll(active & ((1ull << lane) - 1ull));
    int base = 0;
    if (lane == leader) base = atomicAdd(ctr, __popcll(active));
    base = __shfl(base, leader);
    return base + rank;
}

// ---------------------------------------------------- statistics helpers
constexpr int kStatShards = 64;
// per-thread counters reduced in LDS, one global atomic per block and slot
template <int N>
__device__ __forceinline__ void flush_stats(const unsigned (&loc)[N], unsigned long long* stats, int slot0) {
    __shared__ unsigned acc[16];
    __syncthreads();
    if (threadIdx.x < N) acc[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; k++)
        if (loc[k]) atomicAdd(&acc[k], loc[k]);
    __syncthreads();
    // kStatShards copies of every counter, picked by block: thousands of
    // blocks flushing the same counter would otherwise serialise on it
    if (threadIdx.x < N && acc[threadIdx.x])
        atomicAdd(&stats[(size_t)(slot0 + threadIdx.x) * kStatShards + (blockIdx.x & (kStatShards - 1))],
                  (unsigned long long)acc[threadIdx.x]);
}


// loc[st]++ with a run-time st would put loc in scratch memory (a memory
// round trip per predicate); constant indices keep the counters in registers
template <int N>
__device__ __forceinline__ void bump(unsigned (&loc)[N], int st) {
#pragma unroll
    for (int k = 0; k < N; k++) loc[k] += (st == k) ? 1u : 0u;
}

__device__ __forceinline__ void tri_verts(const DMesh& m, int t, int& a, int& b, int& c) {
    int h = m.t_he[t];
    int2 vv = reinterpret_cast<const int2*>(m.he_v)[h >> 1];
    a = (h & 1) ? vv.y : vv.x;
    b = (h & 1) ? vv.x : vv.y;
    c = m.he_op[h];
}

__device__ __forceinline__ unsigned long long claim_key(int round, int e) {
    uint32_t h = (uint32_t)e * 0x9E3779B1u ^ (uint32_t)round * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return ((unsigned long long)(uint32_t)round << 36) | ((unsigned long long)(h & 0xFFu) << 28) |
           (unsigned long long)(uint32_t)e;
}

MCL_HD void flip_soa(const DMesh& m, int e, int& he2, int& he3, int& he4, int& he5) {
    const int he0 = 2 * e, he1 = he0 + 1;
    const int2 nx = reinterpret_cast<const int2*>(m.he_next)[e];
    he2 = nx.x;
    he4 = nx.y;
    he3 = m.he_next[he2];
    he5 = m.he_next[he4];
    const int2 tt = reinterpret_cast<const int2*>(m.he_t)[e];
    const int2 vv = reinterpret_cast<const int2*>(m.he_v)[e];
    const int2 oo = reinterpret_cast<const int2*>(m.he_op)[e];
    const int t0 = tt.x, t1 = tt.y, v0 = vv.x, v1 = vv.y, v2 = oo.x, v3 = oo.y;
    m.t_he[t0] = he0;
    m.t_he[t1] = he1;
    m.v_he[v1] = he2;
    m.v_he[v0] = he4;
    // t0 = (v2 -> v3 -> v1): he0, he5, he2 ;  t1 = (v3 -> v2 -> v0): he1, he3, he4
    reinterpret_cast<int2*>(m.he_next)[e] = make_int2(he5, he3);
    m.he_next[he5] = he2;
    m.he_next[he2] = he0;
    m.he_next[he3] = he4;
    m.he_next[he4] = he1;
    reinterpret_cast<int2*>(m.he_v)[e] = make_int2(v2, v3);
    m.he_t[he5] = t0;
    m.he_t[he3] = t1;
    reinterpret_cast<int2*>(m.he_op)[e] = make_int2(v1, v0);
    m.he_op[he5] = v2;
    m.he_op[he2] = v3;
    m.he_op[he3] = v3;
    m.he_op[he4] = v2;
}

// -------------------------------------------------------------- movement
constexpr int kMotionTarget = 1000;      // internal: the displacement array holds absolute targets
constexpr int kMotionTargetHalf = 1001;  // internal: half way from the current position to those targets
// synthetic displacement of every non-corner vertex; stores the accepted
// position (acc), the target (tgt) and moves the vertex to its target
// (pdev: the parameters in device memory, read at run time, as a captured
// graph needs; null: the by-value pv)
// What a star step's head needs done before its passes, done by block 0 of
// the move instead of separate launches: zero the statistics shards, the
// star counters and the heavy-list count, and advance the check stamps.
// (Empty for the flip-based step.)
struct StepPrologue {
    unsigned long long* stats;
    int nstats;
    int* aux;
    int naux;
    int* heavy;
    int* stamp;
    int stamp_span;
    int* hint;  // per vertex, set to -1 (the checks' candidate hints hold for one step)
};

__global__ void __launch_bounds__(kThreads) k_move_motion(DMesh m, float2* __restrict__ acc, float2* __restrict__ tgt,
                                                          int* __restrict__ v_mv, int* __restrict__ v_fexp, int stamp,
                                                          mcl_step_params_t pv, const mcl_step_params_t* __restrict__ pdev,
                                                          const float2* __restrict__ disp, StepPrologue pr) {
    const mcl_step_params_t p = pdev ? *pdev : pv;
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < pr.nstats; i += blockDim.x) pr.stats[i] = 0;
        for (int i = threadIdx.x; i < pr.naux; i += blockDim.x) pr.aux[i] = 0;
        if (threadIdx.x == 0) {
            if (pr.heavy) *pr.heavy = 0;
            if (pr.stamp) *pr.stamp += pr.stamp_span;
        }
    }
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        if (pr.hint) pr.hint[v] = -1;
        float2 q = m.pos[v];
        acc[v] = q;
        if (is_corner(q)) { tgt[v] = q; continue; }
        float2 t;
        if (disp && p.motion == kMotionTarget) {  // absolute targets (the device build)
            t = disp[v];
        } else if (disp && p.motion == kMotionTargetHalf) {
#pragma clang fp contract(off)
            const float2 d = disp[v];
            t = (d.x == q.x && d.y == q.y) ? q : make_float2(q.x + 0.5f * (d.x - q.x), q.y + 0.5f * (d.y - q.y));
        } else if (disp) {  // caller displacement (movePoints semantics); the torus wraps
            const float2 d = disp[v];
            t = make_float2(q.x + d.x, q.y + d.y);
            if (m.periodic) t = make_float2(wrap01(t.x), wrap01(t.y));
        } else if (p.motion == MCL_MOTION_SHEAR) {
            float2 g = gauss2((uint32_t)v, p.step, p.seed);
            float dx = p.amp * sinpif(2.0f * q.y) + p.sigma * g.x;
            float dy = p.amp * cospif(2.0f * q.x) + p.sigma * g.y;
            t = make_float2(wrap01(q.x + dx), wrap01(q.y + dy));
        } else {
            float2 g = gauss2((uint32_t)v, p.step, p.seed);
            t = make_float2(clamp_box(q.x + p.sigma * g.x), clamp_box(q.y + p.sigma * g.y));
        }
        tgt[v] = t;
        m.pos[v] = t;
        v_mv[v] = stamp;
        v_fexp[v] = 0;
    }
}

// movePoints(d): kernels.cuh:38-43 (corners stay fixed)
__global__ void k_move_disp(DMesh m, const float2* __restrict__ d) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        float2 q = m.pos[v];
        if (is_corner(q)) continue;
        float2 dd = d[v];
        m.pos[v] = make_float2(q.x + dd.x, q.y + dd.y);
    }
}
// integrateMovePoints(dt, v): kernels.cuh:45-50
__global__ void k_integrate(DMesh m, const float* __restrict__ dt, const float2* __restrict__ vel) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        float2 q = m.pos[v];
        if (is_corner(q)) continue;
        float2 w = vel[v];
        float t = dt[v];
        m.pos[v] = make_float2(q.x + w.x * t, q.y + w.y * t);
    }
}

// ---------------------------------------------------- validity (orientation)
// every triangle: strictly CCW?  non-CCW triangles are appended to `bad`
template <bool kF32>
__global__ void __launch_bounds__(kThreads) k_check_all(DMesh m, int* __restrict__ bad, int* __restrict__ cnt, int slot,
                                                        unsigned long long* __restrict__ stats) {
    __shared__ AppendBuf<kListBuf> ab;
    append_init(ab);
    unsigned loc[4] = {0, 0, 0, 0};
    int first, end, stride;
    xcd_split(m.nt, blockDim.x, first, end, stride);
    for (int base = first; base < end; base += stride) {
        const int t = base + threadIdx.x;
        bool is_bad = false;
        if (t < end) {
            int a, b, c, st;
            tri_verts(m, t, a, b, c);
            const int s = orient<kF32>(m.pos[a], m.pos[b], m.pos[c], m.periodic, &st);
            bump(loc, st);
            is_bad = s <= 0;
        }
        const int at = append_reserve(ab, is_bad ? 1 : 0);
        if (is_bad) ab.buf[at] = t;
        append_tile_end(ab, (int)blockDim.x, &cnt[slot], bad);
    }
    append_finish(ab, &cnt[slot], bad);
    flush_stats(loc, stats, kStOrient0);
}

// Taking back moves, phase 1 (positions and topology only read, so every
// decision is race-free).  For each bad triangle:
//  - made by a repair flip of this iteration: undo that flip;
//  - otherwise: take back the one vertex, among those that moved in this
//    iteration and are not back at their accepted position, with the largest
//    displacement.
// Every bad triangle has an action: an original triangle whose vertices are
// all at accepted positions is valid (the iteration's starting state is).
struct RevertArgs {
    const float2* acc;
    const int* v_mv;
    int mv_stamp;
    const int* t_rep;
    const int* t_rep_e;
    int rep_stamp;
    int* e_undo;
    int* undo_list;
    int* v_rev;
    int* rev_list;
    int* stuck;
};
__device__ __forceinline__ void revert_mark_body(const DMesh& m, const int* __restrict__ bad, int n, const RevertArgs& A,
                                                 int undo_stamp, int rev_stamp, int* c_undo, int* c_rev, int* c_stuck,
                                                 int i0, int stride) {
    const float2* __restrict__ acc = A.acc;
    const int* __restrict__ v_mv = A.v_mv;
    const int mv_stamp = A.mv_stamp;
    for (int i = i0; i < n; i += stride) {
        const int t = bad[i];
        if (A.t_rep[t] == A.rep_stamp) {
            const int e = A.t_rep_e[t];
            if (atomicExch(&A.e_undo[e], undo_stamp) != undo_stamp) A.undo_list[agg_inc(c_undo)] = e;
            continue;
        }
        int vs[3];
        tri_verts(m, t, vs[0], vs[1], vs[2]);
        // the vertex that moved most relative to the other two (a common
        // drift of all three does not invert a triangle)
        float2 dsp[3];
        for (int k = 0; k < 3; k++) {
            const float2 q = m.pos[vs[k]], a = acc[vs[k]];
            float dx = q.x - a.x, dy = q.y - a.y;
            if (m.periodic) { dx = min_image_f(dx); dy = min_image_f(dy); }
            dsp[k] = make_float2(dx, dy);
        }
        int best = -1;
        float bd = -1.0f;
        for (int k = 0; k < 3; k++) {
            const int v = vs[k];
            if (v_mv[v] != mv_stamp) continue;
            const float2 q = m.pos[v], a = acc[v];
            if (q.x == a.x && q.y == a.y) continue;
            const float rx = dsp[k].x - 0.5f * (dsp[(k + 1) % 3].x + dsp[(k + 2) % 3].x);
            const float ry = dsp[k].y - 0.5f * (dsp[(k + 1) % 3].y + dsp[(k + 2) % 3].y);
            const float d = rx * rx + ry * ry;
            if (d > bd) { bd = d; best = v; }
        }
        if (best < 0) {
            A.stuck[agg_inc(c_stuck)] = t;
            continue;
        }
        if (atomicExch(&A.v_rev[best], rev_stamp) != rev_stamp) A.rev_list[agg_inc(c_rev)] = best;
    }
}
__global__ void __launch_bounds__(kThreads) k_revert_mark(DMesh m, const int* __restrict__ bad, int* __restrict__ cnt,
                                                          int slot_bad, RevertArgs A, int undo_stamp, int rev_stamp) {
    revert_mark_body(m, bad, cnt[slot_bad], A, undo_stamp, rev_stamp, &cnt[kCnRev + 1], &cnt[kCnRev], &cnt[kCnMisc],
                     blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// all triangles around v strictly CCW with v at p (neighbours where they are now)
template <bool kF32>
__device__ __forceinline__ bool star_valid(const DMesh& m, int v, float2 p) {
    const int h0 = m.v_he[v];
    int h = h0, guard = 0, st;
    do {
        if (m.he_t[h] >= 0) {
            const float2 a = m.pos[m.he_v[h ^ 1]], b = m.pos[m.he_op[h]];
            if (orient<kF32>(p, a, b, m.periodic, &st) <= 0) return false;
        }
        h = m.he_next[h ^ 1];
    } while (h != h0 && ++guard < 64);
    return true;
}

__device__ __forceinline__ float2 shifted(const DMesh& m, float2 a, float dx, float dy) {
    if (m.periodic) return make_float2(wrap01(a.x + dx), wrap01(a.y + dy));
    return make_float2(a.x + dx, a.y + dy);
}

// phase 2a: move the marked vertices back.  The first time in a step a vertex
// is sent back it tries positions consistent with its neighbours' motion --
// its previous position shifted by their mean displacement, then the
// centroid of its neighbours other than box corners (a corner would pull it
// far out of the point cloud, where it would count as a corner itself) --
// and keeps the first at which its whole star is strictly CCW; otherwise
// (and on any later take-back) it returns to its previous position.  Soft moves happen at most once per vertex and step, so
// the passes still end with every taken-back vertex at its previous position
// in the worst case, which is valid.
template <bool kF32>
__device__ __forceinline__ void revert_apply_body(const DMesh& m, const int* __restrict__ rev_list, int n,
                                                  const float2* __restrict__ acc, int* __restrict__ v_fexp,
                                                  int* __restrict__ v_soft, int soft_stamp, int* __restrict__ all_rev,
                                                  int* c_all, unsigned (&loc)[1], int i0, int stride) {
    for (int i = i0; i < n; i += stride) {
        const int v = rev_list[i];
        const float2 a = acc[v];
        float2 dest = a;
        const bool first = v_soft[v] != soft_stamp;
        if (first) {
            v_soft[v] = soft_stamp;
            all_rev[agg_inc(c_all)] = v;
            // neighbour displacement and neighbour centroid, relative to a
            float sdx = 0.f, sdy = 0.f, scx = 0.f, scy = 0.f;
            int deg = 0, ncen = 0;
            const int h0 = m.v_he[v];
            int h = h0;
            do {
                const int u = m.he_v[h ^ 1];
                const float2 pu = m.pos[u], au = acc[u];
                float dx = pu.x - au.x, dy = pu.y - au.y, cx = pu.x - a.x, cy = pu.y - a.y;
                if (m.periodic) { dx = min_image_f(dx); dy = min_image_f(dy); cx = min_image_f(cx); cy = min_image_f(cy); }
                sdx += dx; sdy += dy;
                if (!is_corner(pu)) { scx += cx; scy += cy; ncen++; }
                deg++;
                h = m.he_next[h ^ 1];
            } while (h != h0 && deg < 64);
            const float inv = 1.0f / (float)deg;
            const float2 c1 = shifted(m, a, sdx * inv, sdy * inv);
            const float icen = ncen ? 1.0f / (float)ncen : 0.0f;
            const float2 c2 = shifted(m, a, scx * icen, scy * icen);
            if (!is_corner(c1) && star_valid<kF32>(m, v, c1)) dest = c1;
            else if (ncen && !is_corner(c2) && star_valid<kF32>(m, v, c2)) dest = c2;
        }
        m.pos[v] = dest;
        v_fexp[v] = min(v_fexp[v] + 1, 64);
        loc[0]++;
    }
}
template <bool kF32>
__global__ void __launch_bounds__(kThreads) k_revert_apply(DMesh m, const int* __restrict__ rev_list,
                                                           const int* __restrict__ cnt, int slot_rev,
                                                           const float2* __restrict__ acc, int* __restrict__ v_fexp,
                                                           int* __restrict__ v_soft, int soft_stamp,
                                                           int* __restrict__ all_rev, int* __restrict__ cnt_w, int slot_all,
                                                           unsigned long long* __restrict__ stats) {
    unsigned loc[1] = {0};
    revert_apply_body<kF32>(m, rev_list, cnt[slot_rev], acc, v_fexp, v_soft, soft_stamp, all_rev, &cnt_w[slot_all], loc,
                            blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
    flush_stats(loc, stats, kStReverts);
}

// phase 2b: undo marked repair flips (a repair's two triangles are never
// touched by another repair of the same iteration, so each undo restores
// them exactly) and queue the restored triangles that are still bad
template <bool kF32>
__device__ __forceinline__ void undo_apply_body(const DMesh& m, const int* __restrict__ undo_list, int n,
                                                int* __restrict__ t_rep, int* __restrict__ bad_out, int* c_out,
                                                unsigned (&loc)[4], int i0, int stride) {
    for (int i = i0; i < n; i += stride) {
        const int e = undo_list[i];
        int h2, h3, h4, h5;
        flip_soa(m, e, h2, h3, h4, h5);
        const int2 tt = reinterpret_cast<const int2*>(m.he_t)[e];
        t_rep[tt.x] = 0;
        t_rep[tt.y] = 0;
        const int ts[2] = {tt.x, tt.y};
        for (int k = 0; k < 2; k++) {
            int a, b, c, st;
            tri_verts(m, ts[k], a, b, c);
            int sg = orient<kF32>(m.pos[a], m.pos[b], m.pos[c], m.periodic, &st);
            bump(loc, st);
            if (sg <= 0) bad_out[agg_inc(c_out)] = ts[k];
        }
    }
}
template <bool kF32>
__global__ void __launch_bounds__(kThreads) k_undo_apply(DMesh m, const int* __restrict__ undo_list, int* __restrict__ cnt,
                                                         int slot_undo, int* __restrict__ t_rep, int* __restrict__ bad_out,
                                                         int slot_out, unsigned long long* __restrict__ stats) {
    unsigned loc[4] = {0, 0, 0, 0};
    undo_apply_body<kF32>(m, undo_list, cnt[slot_undo], t_rep, bad_out, &cnt[slot_out], loc,
                          blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
    flush_stats(loc, stats, kStOrient0);
}

// Repair rounds: keep the triangles of `bad` that are still not strictly CCW
// and queue their edges as candidates of round `round`.
template <bool kF32>
__global__ void __launch_bounds__(kThreads) k_filter_bad(DMesh m, const int* __restrict__ bad, int* __restrict__ cnt,
                                                         int slot_in, int* __restrict__ bad_out, int slot_out,
                                                         int* __restrict__ e_stamp, int round, int* __restrict__ cand,
                                                         unsigned long long* __restrict__ stats) {
    const int n = cnt[slot_in];
    unsigned loc[4] = {0, 0, 0, 0};
    const int slot_c = kCnCand + round % 3;
    // both lists through LDS, one global reservation per flush (a wave-level
    // atomic per list and wave serialises thousands of waves on two counters)
    __shared__ AppendBuf<kListBuf> ab_bad, ab_cand;
    append_init(ab_bad);
    append_init(ab_cand);
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int i = base + threadIdx.x;
        bool keep = false;
        int t = -1, es[3] = {-1, -1, -1};
        if (i < n) {
            t = bad[i];
            int a, b, c, st;
            tri_verts(m, t, a, b, c);
            keep = orient<kF32>(m.pos[a], m.pos[b], m.pos[c], m.periodic, &st) <= 0;
            bump(loc, st);
            if (keep) {
                const int h0 = m.t_he[t], h1 = m.he_next[h0], h2 = m.he_next[h1];
                es[0] = h0 >> 1; es[1] = h1 >> 1; es[2] = h2 >> 1;
                for (int k = 0; k < 3; k++)
                    if (atomicExch(&e_stamp[es[k]], round) == round) es[k] = -1;
            }
        }
        const int at = append_reserve(ab_bad, keep ? 1 : 0);
        if (keep) ab_bad.buf[at] = t;
        int p = append_reserve(ab_cand, (es[0] >= 0) + (es[1] >= 0) + (es[2] >= 0));
        for (int k = 0; k < 3; k++)
            if (es[k] >= 0) ab_cand.buf[p++] = es[k];
        append_tile_end(ab_bad, (int)blockDim.x, &cnt[slot_out], bad_out);
        append_tile_end(ab_cand, 3 * (int)blockDim.x, &cnt[slot_c], cand);
    }
    append_finish(ab_bad, &cnt[slot_out], bad_out);
    append_finish(ab_cand, &cnt[slot_c], cand);
    flush_stats(loc, stats, kStOrient0);
}

// winners of a repair round flip; their two triangles are marked with the
// iteration stamp and the edge (for a possible undo); the quad's vertices are
// recorded so that later iterations re-queue the quad's edges for Lawson
__global__ void __launch_bounds__(kThreads) k_repair_flip(DMesh m, int* __restrict__ cnt, int round,
                                                          const int* __restrict__ wants,
                                                          const unsigned long long* __restrict__ claim,
                                                          int* __restrict__ t_rep, int* __restrict__ t_rep_e, int rep_stamp,
                                                          int* __restrict__ v_touch, int* __restrict__ touched,
                                                          int slot_touched, unsigned long long* __restrict__ stats) {
    const int r3 = round % 3;
    const int n = cnt[kCnWants + r3];
    unsigned loc[1] = {0};
    // the touched list through LDS and the flip count per block (wave-level
    // atomics on one counter serialise thousands of waves)
    __shared__ AppendBuf<kListBuf> ab;
    __shared__ int s_flips;
    if (threadIdx.x == 0) s_flips = 0;
    append_init(ab);
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int i = base + threadIdx.x;
        int qv[4] = {-1, -1, -1, -1};
        if (i < n) {
            const int e = wants[i];
            const int2 tt = reinterpret_cast<const int2*>(m.he_t)[e];
            const unsigned long long key = claim_key(round, e);
            if (claim[tt.x] == key && claim[tt.y] == key) {
                const int2 vv = reinterpret_cast<const int2*>(m.he_v)[e];
                const int2 oo = reinterpret_cast<const int2*>(m.he_op)[e];
                int h2, h3, h4, h5;
                flip_soa(m, e, h2, h3, h4, h5);
                t_rep[tt.x] = rep_stamp; t_rep[tt.y] = rep_stamp;
                t_rep_e[tt.x] = e; t_rep_e[tt.y] = e;
                qv[0] = vv.x; qv[1] = vv.y; qv[2] = oo.x; qv[3] = oo.y;
                for (int k = 0; k < 4; k++)
                    if (atomicExch(&v_touch[qv[k]], rep_stamp) == rep_stamp) qv[k] = -1;
                atomicAdd(&s_flips, 1);
                loc[0]++;
            }
        }
        int p = append_reserve(ab, (qv[0] >= 0) + (qv[1] >= 0) + (qv[2] >= 0) + (qv[3] >= 0));
        for (int k = 0; k < 4; k++)
            if (qv[k] >= 0) ab.buf[p++] = qv[k];
        append_tile_end(ab, 4 * (int)blockDim.x, &cnt[slot_touched], touched);
    }
    append_finish(ab, &cnt[slot_touched], touched);
    if (threadIdx.x == 0 && s_flips) atomicAdd(&cnt[kCnErr], s_flips);
    flush_stats(loc, stats, kStUntangleFlips);
}


// check the triangles around each listed vertex; append non-CCW ones
template <bool kF32>
__device__ __forceinline__ void check_ring_body(const DMesh& m, const int* __restrict__ vlist, int n,
                                                int* __restrict__ t_stamp, int stamp, int* __restrict__ bad, int* c_out,
                                                unsigned (&loc)[4], int i0, int stride) {
    for (int i = i0; i < n; i += stride) {
        int v = vlist[i];
        int h0 = m.v_he[v], h = h0, guard = 0;
        do {
            int t = m.he_t[h];
            if (t >= 0 && atomicExch(&t_stamp[t], stamp) != stamp) {
                int a, b, c, st;
                tri_verts(m, t, a, b, c);
                int s = orient<kF32>(m.pos[a], m.pos[b], m.pos[c], m.periodic, &st);
                bump(loc, st);
                if (s <= 0) bad[agg_inc(c_out)] = t;
            }
            h = m.he_next[h ^ 1];
        } while (h != h0 && ++guard < 4096);
    }
}
template <bool kF32>
__global__ void __launch_bounds__(kThreads) k_check_ring(DMesh m, const int* __restrict__ vlist, const int* __restrict__ cnt_in,
                                                         int slot_in, int* __restrict__ t_stamp, int stamp, int* __restrict__ bad,
                                                         int* __restrict__ cnt_out, int slot_out,
                                                         unsigned long long* __restrict__ stats) {
    unsigned loc[4] = {0, 0, 0, 0};
    check_ring_body<kF32>(m, vlist, cnt_in[slot_in], t_stamp, stamp, bad, &cnt_out[slot_out], loc,
                          blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
    flush_stats(loc, stats, kStOrient0);
}

// Take-back passes finished by one workgroup: the same mark / apply / undo /
// re-check phases as the grid kernels, separated by workgroup barriers, with
// the list lengths in LDS, until no triangle is bad.  Stamps stamp0 + 2p and
// stamp0 + 2p + 1 are used by pass p.  out = {passes, final list slot,
// stuck, undos}.
template <bool kF32>
__global__ void __launch_bounds__(1024) k_revert_tail(DMesh m, int* __restrict__ bad0, int* __restrict__ bad1, int cur0,
                                                      int* __restrict__ cnt, RevertArgs A, int* __restrict__ t_rep,
                                                      int stamp0, int max_passes, int* __restrict__ v_fexp,
                                                      int* __restrict__ v_soft, int soft_stamp, int* __restrict__ all_rev,
                                                      int* __restrict__ t_stamp, unsigned long long* __restrict__ stats,
                                                      int* __restrict__ out) {
    __shared__ int s_nb, s_rev, s_undo, s_stuck, s_next;
    unsigned lrev[1] = {0}, lor[4] = {0, 0, 0, 0};
    int cur = cur0, pass = 0, undos = 0, stuck = 0;
    if (threadIdx.x == 0) s_nb = cnt[kCnBad + cur0];
    __syncthreads();
    for (; pass < max_passes; pass++) {
        const int nb = s_nb;
        if (nb == 0) break;
        int* bad = cur ? bad1 : bad0;
        int* bad_next = cur ? bad0 : bad1;
        __syncthreads();
        if (threadIdx.x == 0) { s_rev = 0; s_undo = 0; s_stuck = 0; s_next = 0; }
        __syncthreads();
        const int st0 = stamp0 + 2 * pass;
        revert_mark_body(m, bad, nb, A, st0, st0, &s_undo, &s_rev, &s_stuck, threadIdx.x, blockDim.x);
        __syncthreads();
        if (s_stuck > 0) { stuck = 1; break; }
        const int nr = s_rev, nu = s_undo;
        undos += nu;
        revert_apply_body<kF32>(m, A.rev_list, nr, A.acc, v_fexp, v_soft, soft_stamp, all_rev, &cnt[kCnRev + 2], lrev,
                                threadIdx.x, blockDim.x);
        __syncthreads();
        undo_apply_body<kF32>(m, A.undo_list, nu, t_rep, bad_next, &s_next, lor, threadIdx.x, blockDim.x);
        __syncthreads();
        check_ring_body<kF32>(m, A.rev_list, nr, t_stamp, st0 + 1, bad_next, &s_next, lor, threadIdx.x, blockDim.x);
        __syncthreads();
        if (threadIdx.x == 0) {
            s_nb = s_next;
            cnt[kCnBad + (1 - cur)] = s_next;
        }
        cur = 1 - cur;
        __syncthreads();
    }
    flush_stats(lrev, stats, kStReverts);
    flush_stats(lor, stats, kStOrient0);
    if (threadIdx.x == 0) {
        out[0] = pass;
        out[1] = cur;
        out[2] = stuck;
        out[3] = undos;
    }
}

// next attempt for vertices short of their target.  fexp == 0: the whole
// remaining displacement.  fexp >= 1 (the previous attempt was taken back):
// 0.9 * 2^-(fexp-1) of the fraction at which the first triangle around the
// vertex would flip orientation with its neighbours where they are now
// (orientation is affine in the fraction, so the crossing is O0 / (O0 - O1)).
// The attempt is checked exactly afterwards, so the estimate needs no care
// beyond staying in (0, 1].
__global__ void __launch_bounds__(kThreads) k_advance(DMesh m, const int* __restrict__ vlist, const int* __restrict__ cnt,
                                                      int slot, float2* __restrict__ acc, float2* __restrict__ tgt,
                                                      int* __restrict__ v_mv, int stamp, const int* __restrict__ v_fexp,
                                                      int fexp_cap, unsigned long long* __restrict__ stats) {
    const int n = cnt[slot];
    unsigned loc[1] = {0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int v = vlist[i];
        const float2 q = m.pos[v], t = tgt[v];
        acc[v] = q;
        const int fe = v_fexp[v];
        if (fe >= fexp_cap) {  // give up: the vertex stays where it is this step
            tgt[v] = q;
            loc[0]++;
            continue;
        }
        float dx = t.x - q.x, dy = t.y - q.y;
        if (m.periodic) { dx = min_image_f(dx); dy = min_image_f(dy); }
        float2 x;
        if (fe == 0) {
            x = t;
        } else {
            double fstar = 1.0;
            const int h0 = m.v_he[v];
            int h = h0, guard = 0;
            do {
                if (m.he_t[h] >= 0) {
                    // triangle (v, a, b) with a = dest(h), b = op(h)
                    const float2 pa = m.pos[m.he_v[h ^ 1]], pb = m.pos[m.he_op[h]];
                    double ax = (double)pa.x - q.x, ay = (double)pa.y - q.y;
                    double bx = (double)pb.x - q.x, by = (double)pb.y - q.y;
                    if (m.periodic) {
                        ax = ax > 0.5 ? ax - 1.0 : (ax < -0.5 ? ax + 1.0 : ax);
                        ay = ay > 0.5 ? ay - 1.0 : (ay < -0.5 ? ay + 1.0 : ay);
                        bx = bx > 0.5 ? bx - 1.0 : (bx < -0.5 ? bx + 1.0 : bx);
                        by = by > 0.5 ? by - 1.0 : (by < -0.5 ? by + 1.0 : by);
                    }
                    const double o0 = ax * by - ay * bx;
                    const double o1 = (ax - dx) * (by - dy) - (ay - dy) * (bx - dx);
                    if (o1 <= 0.0 && o0 > 0.0) fstar = fmin(fstar, o0 / (o0 - o1));
                }
                h = m.he_next[h ^ 1];
            } while (h != h0 && ++guard < 4096);
            const float f = (float)(0.9 * fstar) * ldexpf(1.0f, -(fe - 1));
            if (m.periodic) x = make_float2(wrap01(q.x + f * dx), wrap01(q.y + f * dy));
            else x = make_float2(q.x + f * dx, q.y + f * dy);
        }
        m.pos[v] = x;
        v_mv[v] = stamp;
    }
    flush_stats(loc, stats, kStStalled);
}

// edges whose Delaunay test can change after vertex v moved: the edges
// incident to v and the edges opposite v in its triangles
__global__ void __launch_bounds__(kThreads) k_enqueue_ring(DMesh m, const int* __restrict__ vlist, const int* __restrict__ cnt,
                                                           int slot, const float2* __restrict__ acc, int* __restrict__ e_stamp,
                                                           int estamp, int* __restrict__ cand, int* __restrict__ cnt_c,
                                                           int slot_c, int require_moved) {
    const int n = cnt[slot];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int v = vlist[i];
        float2 q = m.pos[v], a = acc[v];
        if (require_moved && q.x == a.x && q.y == a.y) continue;
        int h0 = m.v_he[v], h = h0, guard = 0;
        do {
            int e1 = h >> 1, e2 = m.he_next[h] >> 1;
            if (atomicExch(&e_stamp[e1], estamp) != estamp) cand[agg_inc(&cnt_c[slot_c])] = e1;
            if (m.he_t[h] >= 0 && atomicExch(&e_stamp[e2], estamp) != estamp) cand[agg_inc(&cnt_c[slot_c])] = e2;
            h = m.he_next[h ^ 1];
        } while (h != h0 && ++guard < 4096);
    }
}

// vertices of this iteration that did not reach their target stay pending;
// a successful partial move doubles the next attempt's fraction
__global__ void __launch_bounds__(kThreads) k_collect_pending(DMesh m, const int* __restrict__ vlist, const int* __restrict__ cnt,
                                                              int slot, const float2* __restrict__ acc,
                                                              const float2* __restrict__ tgt, int* __restrict__ v_fexp,
                                                              int* __restrict__ pend, int* __restrict__ cnt_p, int slot_p) {
    const int n = cnt[slot];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int v = vlist[i];
        float2 q = m.pos[v], t = tgt[v], a = acc[v];
        if (q.x == t.x && q.y == t.y) continue;
        bool moved = !(q.x == a.x && q.y == a.y);
        if (moved) v_fexp[v] = 0;  // after a successful partial move, try the whole rest
        pend[agg_inc(&cnt_p[slot_p])] = v;
    }
}

// ------------------------------------------------------------ relocation
// Moving a vertex to a point its current star cannot reach without
// inverting a triangle: remove it (flips inside its star down to degree 3,
// then the three triangles merge), and insert it again at the target (1->3
// split of the containing triangle, 2->4 if the target lies on an edge).
// This is the remove-and-reinsert move that the reference declares as
// moveFlipflop (Host_Triangulation.h:77, 408) and leaves unimplemented.
// Every flip and split keeps all triangles strictly CCW (checked exactly),
// so the mesh stays valid; Lawson rounds restore the Delaunay property.
constexpr int kMaxStar = 64;

// Remembering stochastic visibility walk (Devillers, Pion, Teillaud, "Walking
// in a triangulation", 2002) from triangle t toward p: at each triangle the
// edges are tried from a random one, the edge just crossed is skipped, and
// the walk crosses the first edge that has p strictly on its far side.  A
// deterministic visibility walk can cycle forever in a triangulation that is
// not Delaunay; the random start makes termination certain with probability
// one.  Returns the containing triangle (and in *on the half-edge of it that
// p lies on, or -1); -1 if the walk leaves the mesh (p outside the
// triangulated domain); -5 if it exceeds kWalkCap triangles (a retry starts
// from another random sequence); -2 if p coincides with a vertex, whose id
// is then in *on.
constexpr int kWalkCap = 2048;
template <bool kF32>
MCL_HD int locate_point(const DMesh& m, float2 p, int t, int* on, unsigned seed) {
    int came = -1;  // half-edge of t we entered through
    unsigned r = seed * 0x9E3779B9u + 0x7F4A7C15u;
    for (int step = 0; step < kWalkCap; step++) {
        const int h0 = m.t_he[t];
        const int h1 = m.he_next[h0], h2 = m.he_next[h1];
        r ^= r << 13; r ^= r >> 17; r ^= r << 5;
        const int k0 = (int)(r % 3u);
        int zero = -1, cross = -1;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int idx = (k0 + k) % 3;  // selects, not an indexed array: stays in registers
            const int h = idx == 0 ? h0 : (idx == 1 ? h1 : h2);
            const int ia = m.he_v[h], ib = m.he_v[h ^ 1];
            const float2 a = m.pos[ia], b = m.pos[ib];
            if (a.x == p.x && a.y == p.y) { *on = ia; return -2; }
            if (b.x == p.x && b.y == p.y) { *on = ib; return -2; }
            if (h == came) continue;  // p is on our side of the edge we crossed
            int st;
            const int o = orient<kF32>(a, b, p, m.periodic, &st);
            if (o < 0) { cross = h; break; }
            if (o == 0) zero = h;
        }
        if (cross < 0) {
            *on = zero;
            return t;
        }
        const int nt = m.he_t[cross ^ 1];
        if (nt < 0) return -1;
        came = cross ^ 1;
        t = nt;
    }
    return -5;
}

// phase 1 (read-only): find where each listed vertex's target lies.
// r_tt[v]: the target triangle, or
//   -3  v is at its target already
//   -2  the target is the position of a vertex w whose own target it is
//       (two vertices with one target: v stays where it is, a duplicate)
//   -6  the target is the position of a vertex w that has not reached its
//       own target yet: v waits for w to leave (retried next round)
//   -5  the walk exceeded its length cap (retried next round)
//   -1  the walk left the mesh (target outside the triangulated domain)
template <bool kF32>
__device__ __forceinline__ void reloc_locate_body(const DMesh& m, const int* __restrict__ vlist, int n,
                                                  const float2* __restrict__ tgt, int* __restrict__ r_tt,
                                                  int* __restrict__ r_tt2, int round, int i0, int stride) {
    for (int i = i0; i < n; i += stride) {
        const int v = vlist[i];
        const float2 t = tgt[v], q = m.pos[v];
        if (t.x == q.x && t.y == q.y) { r_tt[v] = -3; continue; }
        int on = -1;
        int tt = locate_point<kF32>(m, t, m.he_t[m.v_he[v]], &on, (unsigned)v ^ ((unsigned)round << 20));
        if (tt == -2) {
            const float2 tw = tgt[on];
            tt = (tw.x == t.x && tw.y == t.y) ? -2 : -6;
            on = -1;
        }
        r_tt[v] = tt;
        r_tt2[v] = (tt >= 0 && on >= 0) ? m.he_t[on ^ 1] : -1;
    }
}
template <bool kF32>
__global__ void __launch_bounds__(kThreads) k_reloc_locate(DMesh m, const int* __restrict__ vlist,
                                                           const int* __restrict__ cnt, int slot,
                                                           const float2* __restrict__ tgt, int* __restrict__ r_tt,
                                                           int* __restrict__ r_tt2, int round) {
    reloc_locate_body<kF32>(m, vlist, cnt[slot], tgt, r_tt, r_tt2, round, blockIdx.x * blockDim.x + threadIdx.x,
                            gridDim.x * blockDim.x);
}

// phase 2: claim the star triangles and the target triangle(s)
__device__ __forceinline__ void reloc_claim_body(const DMesh& m, const int* __restrict__ vlist, int n, int round,
                                                 int* __restrict__ r_tt, const int* __restrict__ r_tt2,
                                                 unsigned long long* __restrict__ claim, int i0, int stride) {
    for (int i = i0; i < n; i += stride) {
        const int v = vlist[i];
        const int tt = r_tt[v];
        if (tt < 0) continue;
        const unsigned long long key = claim_key(round, v);
        const int h0 = m.v_he[v];
        int h = h0, guard = 0;
        do {
            const int t = m.he_t[h];
            if (t >= 0) atomicMax(&claim[t], key);
            h = m.he_next[h ^ 1];
        } while (h != h0 && ++guard < kMaxStar);
        // the mesh is not modified during this phase, so this is the true
        // degree: a star too large for the apply phase is given up here
        if (h != h0) { r_tt[v] = -4; continue; }
        atomicMax(&claim[tt], key);
        if (r_tt2[v] >= 0) atomicMax(&claim[r_tt2[v]], key);
    }
}
__global__ void __launch_bounds__(kThreads) k_reloc_claim(DMesh m, const int* __restrict__ vlist,
                                                          const int* __restrict__ cnt, int slot, int round,
                                                          int* __restrict__ r_tt, const int* __restrict__ r_tt2,
                                                          unsigned long long* __restrict__ claim) {
    reloc_claim_body(m, vlist, cnt[slot], round, r_tt, r_tt2, claim, blockIdx.x * blockDim.x + threadIdx.x,
                     gridDim.x * blockDim.x);
}

MCL_HD void set_he(const DMesh& m, int h, int nx, int v, int t, int o) {
    m.he_next[h] = nx;
    m.he_v[h] = v;
    m.he_t[h] = t;
    m.he_op[h] = o;
}

// containment of p in triangle X: 1 inside (on_edge = half-edge or -1), 0 outside
template <bool kF32>
MCL_HD int tri_contains(const DMesh& m, int X, float2 p, int* on_edge) {
    const int h0 = m.t_he[X];
    const int hs[3] = {h0, m.he_next[h0], m.he_next[m.he_next[h0]]};
    int zero = -1, st;
    for (int k = 0; k < 3; k++) {
        const int o = orient<kF32>(m.pos[m.he_v[hs[k]]], m.pos[m.he_v[hs[k] ^ 1]], p, m.periodic, &st);
        if (o < 0) return 0;
        if (o == 0) zero = hs[k];
    }
    *on_edge = zero;
    return 1;
}

// Ear test of the removal: the ear (c, a, b) of v's link polygon (CCW order
// c, a, b; strictly convex at a) is a proper ear when no other link vertex
// lies in the closed triangle, i.e. when the diagonal (c, b) runs inside the
// polygon and touches its boundary only at c and b.  h: the spoke v -> a.
template <bool kF32>
MCL_HD bool ear_empty(const DMesh& m, int h, int a, int b, int c, float2 pa, float2 pb,
                                          float2 pc) {
    int st;
    int g = m.he_next[h ^ 1], guard = 0;
    while (g != h && ++guard < kMaxStar) {
        const int x = m.he_v[g ^ 1];
        if (x != a && x != b && x != c) {
            const float2 px = m.pos[x];
            if (orient<kF32>(pc, pa, px, m.periodic, &st) >= 0 && orient<kF32>(pa, pb, px, m.periodic, &st) >= 0 &&
                orient<kF32>(pb, pc, px, m.periodic, &st) >= 0)
                return false;
        }
        g = m.he_next[g ^ 1];
    }
    return true;
}

// Removal and re-insertion of vertex v (the mesh part of a relocation; the
// caller holds every triangle involved): star[0..deg0) are the triangles
// around v, tt (and tt2 when the target lies on an edge) the triangle(s)
// holding the target pt.  Returns 0 when v is at pt, 1 when the removal found
// no proper ear (mesh unchanged but for the flips made), 2 when v had to be
// put back at its old position.  Host and device.
template <bool kF32>
MCL_HD int reloc_core(const DMesh& m, int v, const int* star, int deg0, int tt, int tt2, float2 pt) {
    int deg = deg0;
    const float2 pv = m.pos[v];
    int st;
    // --- remove: flip spokes (v, a) down to degree 3.  Each flip clips
    // the ear (c, a, b) off v's link polygon (CCW order c, a, b) and
    // adds the ear triangle, which must be strictly CCW.  First only ears
    // that keep v strictly inside the rest (both new triangles strictly
    // CCW).  A link with collinear vertices can have none: v on the line
    // (c, b) of every convex ear, as at the centre of a square.  Then any
    // proper ear is clipped; v's own triangles may turn degenerate, but
    // they all disappear in the merge, and the ears of a simple polygon
    // with four or more vertices always include a proper one (Meisters'
    // two-ears theorem), whose last remaining triangle is strictly CCW.
    bool strict = true;
    while (deg > 3) {
        bool flipped = false;
        const int h0 = m.v_he[v];
        int h = h0, guard = 0;
        do {
            const int a = m.he_v[h ^ 1], b = m.he_op[h], c = m.he_op[h ^ 1];
            const float2 pa = m.pos[a], pb = m.pos[b], pc = m.pos[c];
            if (orient<kF32>(pb, pc, pa, m.periodic, &st) > 0 &&
                (strict ? orient<kF32>(pc, pb, pv, m.periodic, &st) > 0
                        : ear_empty<kF32>(m, h, a, b, c, pa, pb, pc))) {
                int h2, h3, h4, h5;
                flip_soa(m, h >> 1, h2, h3, h4, h5);  // keeps v.he an outgoing half-edge of v
                flipped = true;
                break;
            }
            h = m.he_next[h ^ 1];
        } while (h != h0 && ++guard < kMaxStar);
        if (!flipped) {
            if (!strict) break;
            strict = false;
            continue;
        }
        deg--;
    }
    if (deg > 3) return 1;  // no proper ear (a link that is not a simple polygon): mesh still valid
    // --- merge the three triangles around v into one
    int hv[3], av[3], tv[3], gv[3], ov[3];
    {
        int h = m.v_he[v];
        for (int k = 0; k < 3; k++) {
            hv[k] = h;
            av[k] = m.he_v[h ^ 1];
            tv[k] = m.he_t[h];
            gv[k] = m.he_next[h];  // link half-edge a_k -> op(h_k)
            ov[k] = m.he_op[h];
            h = m.he_next[h ^ 1];
        }
    }
    const int T0 = tv[0], F1 = tv[1], F2 = tv[2];
    const int E0 = hv[0] >> 1, E1 = hv[1] >> 1, E2 = hv[2] >> 1;
    const int vsum = av[0] + av[1] + av[2];
    for (int k = 0; k < 3; k++) {
        int j = 0;
        for (int q = 0; q < 3; q++)
            if (av[q] == ov[k]) j = q;
        m.he_next[gv[k]] = gv[j];
        m.he_t[gv[k]] = T0;
        m.he_op[gv[k]] = vsum - av[k] - ov[k];
        m.v_he[av[k]] = gv[k];
    }
    m.t_he[T0] = gv[0];
    // --- the triangle that receives v: search the claimed live triangles
    int cand[kMaxStar + 2], nc = 0;
    for (int k = 0; k < deg0; k++)
        if (star[k] != F1 && star[k] != F2) cand[nc++] = star[k];
    bool tt_known = false, tt2_known = tt2 < 0;
    for (int k = 0; k < nc; k++) { tt_known |= cand[k] == tt; tt2_known |= cand[k] == tt2; }
    if (!tt_known && tt != F1 && tt != F2) cand[nc++] = tt;
    if (!tt2_known && tt2 != F1 && tt2 != F2) cand[nc++] = tt2;
    float2 pins = pt;
    int X = -1, on = -1;
    for (int pass = 0; pass < 2 && X < 0; pass++) {
        for (int k = 0; k < nc && X < 0; k++) {
            int oe = -1;
            if (tri_contains<kF32>(m, cand[k], pins, &oe)) {
                if (oe >= 0) {  // on an edge: the other side must be ours and live
                    const int tw = m.he_t[oe ^ 1];
                    bool ours = false;
                    for (int q = 0; q < nc; q++) ours |= cand[q] == tw;
                    if (!ours) continue;
                }
                X = cand[k];
                on = oe;
            }
        }
        if (X < 0) pins = pv;  // cannot place the target: put v back where it was
    }
    // the old position lies strictly inside the merged star polygon, so X >= 0 here
    const int n0 = 2 * E0, n1 = 2 * E1, n2 = 2 * E2;
    if (on < 0) {
        const int x0 = m.t_he[X], x1 = m.he_next[x0], x2 = m.he_next[x1];
        const int a = m.he_v[x0], b = m.he_v[x1], c = m.he_v[x2];
        set_he(m, x0, n0, a, X, v); set_he(m, n0, n2 ^ 1, b, X, a); set_he(m, n2 ^ 1, x0, v, X, b);
        set_he(m, x1, n1, b, F1, v); set_he(m, n1, n0 ^ 1, c, F1, b); set_he(m, n0 ^ 1, x1, v, F1, c);
        set_he(m, x2, n2, c, F2, v); set_he(m, n2, n1 ^ 1, a, F2, c); set_he(m, n1 ^ 1, x2, v, F2, a);
        m.t_he[X] = x0; m.t_he[F1] = x1; m.t_he[F2] = x2;
        m.v_he[v] = n2 ^ 1;
    } else {
        const int he0 = on, he1 = on ^ 1;
        const int he2 = m.he_next[he0], he3 = m.he_next[he2], he4 = m.he_next[he1], he5 = m.he_next[he4];
        const int t0 = X, t1 = m.he_t[he1];
        const int a = m.he_v[he0], b = m.he_v[he1], c = m.he_v[he3], d = m.he_v[he5];
        set_he(m, he0, n1, a, t0, c); set_he(m, n1, he3, v, t0, a); set_he(m, he3, he0, c, t0, v);
        set_he(m, n0, he2, v, F1, c); set_he(m, he2, n1 ^ 1, b, F1, v); set_he(m, n1 ^ 1, n0, c, F1, b);
        set_he(m, he1, he4, v, t1, d); set_he(m, he4, n2 ^ 1, a, t1, v); set_he(m, n2 ^ 1, he1, d, t1, a);
        set_he(m, n0 ^ 1, n2, b, F2, d); set_he(m, n2, he5, v, F2, b); set_he(m, he5, n0 ^ 1, d, F2, v);
        m.t_he[t0] = he0; m.t_he[F1] = n0; m.t_he[t1] = he1; m.t_he[F2] = n0 ^ 1;
        m.v_he[v] = n0; m.v_he[a] = he0; m.v_he[b] = he2;
    }
    m.pos[v] = pins;
    return (pins.x == pt.x && pins.y == pt.y) ? 0 : 2;
}

// phase 3: winners (all claims held) remove and re-insert their vertex.
// loc: {relocated, stalled, walk left the mesh or retries spent, star above
// kMaxStar (counted by the driver, see reloc_deferred), removal stuck, put
// back, duplicate targets}
constexpr int kRelocRetries = 64;  // waits and walk retries per vertex and step
template <bool kF32>
__device__ __forceinline__ void reloc_apply_body(const DMesh& m, const int* __restrict__ vlist, int n, int round,
                                                 const int* __restrict__ r_tt, const int* __restrict__ r_tt2,
                                                 int* __restrict__ r_wait, const unsigned long long* __restrict__ claim,
                                                 const float2* __restrict__ tgt, int* __restrict__ next_list, int* c_next,
                                                 unsigned (&loc)[7], int i0, int stride) {
    for (int i = i0; i < n; i += stride) {
        const int v = vlist[i];
        const int tt = r_tt[v], tt2 = r_tt2[v];
        if (tt == -3) continue;              // already at its target
        if (tt == -2) { loc[6]++; continue; }  // another vertex holds the same target: stays put
        if (tt == -6 || tt == -5) {          // waits for a vertex to leave its target, or walks again
            if (++r_wait[v] <= kRelocRetries) {
                next_list[agg_inc(c_next)] = v;
            } else {
                loc[1]++;
                loc[2]++;
            }
            continue;
        }
        if (tt == -4) continue;  // star larger than kMaxStar: stays put for this phase (counted by the
                                 // driver, which may relocate it again after Lawson rounds)
        if (tt < 0) { loc[1]++; loc[2]++; continue; }  // target outside the mesh: stays put
        const unsigned long long key = claim_key(round, v);
        int star[kMaxStar];
        int deg = 0;
        bool mine = claim[tt] == key && (tt2 < 0 || claim[tt2] == key);
        {
            const int h0 = m.v_he[v];
            int h = h0;
            do {
                if (deg >= kMaxStar) { mine = false; break; }
                const int t = m.he_t[h];
                if (t < 0 || claim[t] != key) mine = false;
                star[deg++] = t;
                h = m.he_next[h ^ 1];
            } while (h != h0);
        }
        if (!mine) {  // lost a claim: retry next round.  (A loser's walk may
            // cross triangles a winner is rewriting, so its count is not a degree.)
            next_list[agg_inc(c_next)] = v;
            continue;
        }
        const int rc = reloc_core<kF32>(m, v, star, deg, tt, tt2, tgt[v]);
        if (rc == 0) loc[0]++;
        else if (rc == 1) { loc[1]++; loc[4]++; }
        else { loc[1]++; loc[5]++; }
    }
}
template <bool kF32>
__global__ void __launch_bounds__(kThreads) k_reloc_apply(DMesh m, const int* __restrict__ vlist,
                                                          const int* __restrict__ cnt, int slot, int round,
                                                          const int* __restrict__ r_tt, const int* __restrict__ r_tt2,
                                                          int* __restrict__ r_wait,
                                                          const unsigned long long* __restrict__ claim,
                                                          const float2* __restrict__ tgt, int* __restrict__ next_list,
                                                          int* __restrict__ cnt_next, int slot_next,
                                                          unsigned long long* __restrict__ stats) {
    unsigned loc[7] = {0, 0, 0, 0, 0, 0, 0};  // relocated, stalled, the reasons, duplicates
    reloc_apply_body<kF32>(m, vlist, cnt[slot], round, r_tt, r_tt2, r_wait, claim, tgt, next_list,
                           &cnt_next[slot_next], loc, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
    flush_stats(loc, stats, kStRelocated);
    const unsigned grid_loc[1] = {loc[0]};
    flush_stats(grid_loc, stats, kStRelocGrid);
}

// relocation rounds finished by one workgroup (locate / claim / apply with
// workgroup barriers, list lengths in LDS); rounds round0, round0 + 1, ...
// out = {rounds, final list index (0/1), remaining}
template <bool kF32>
__global__ void __launch_bounds__(1024) k_reloc_tail(DMesh m, int* __restrict__ list0, int* __restrict__ list1, int cur0,
                                                     int n0, int round0, int max_rounds, int* __restrict__ r_tt,
                                                     int* __restrict__ r_tt2, int* __restrict__ r_wait,
                                                     unsigned long long* __restrict__ claim,
                                                     const float2* __restrict__ tgt, unsigned long long* __restrict__ stats,
                                                     int* __restrict__ out) {
    __shared__ int s_n, s_next;
    unsigned loc[7] = {0, 0, 0, 0, 0, 0, 0};
    int cur = cur0, r = 0;
    if (threadIdx.x == 0) s_n = n0;
    __syncthreads();
    for (; r < max_rounds; r++) {
        const int n = s_n;
        if (n == 0) break;
        const int* list = cur ? list1 : list0;
        int* next = cur ? list0 : list1;
        const int round = round0 + r;
        __syncthreads();
        if (threadIdx.x == 0) s_next = 0;
        reloc_locate_body<kF32>(m, list, n, tgt, r_tt, r_tt2, round, threadIdx.x, blockDim.x);
        __syncthreads();
        reloc_claim_body(m, list, n, round, r_tt, r_tt2, claim, threadIdx.x, blockDim.x);
        __syncthreads();
        reloc_apply_body<kF32>(m, list, n, round, r_tt, r_tt2, r_wait, claim, tgt, next, &s_next, loc, threadIdx.x,
                               blockDim.x);
        __syncthreads();
        if (threadIdx.x == 0) s_n = s_next;
        cur = 1 - cur;
        __syncthreads();
    }
    flush_stats(loc, stats, kStRelocated);
    if (threadIdx.x == 0) {
        out[0] = r;
        out[1] = cur;
        out[2] = s_n;
    }
}

// the relocation list of a step (the vertices sent back) with their retry
// counts cleared
__global__ void __launch_bounds__(kThreads) k_reloc_init(const int* __restrict__ src, int n, int* __restrict__ dst,
                                                         int* __restrict__ r_wait) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int v = src[i];
        dst[i] = v;
        r_wait[v] = 0;
    }
}

// the vertices of list (n) whose relocation met a star above kMaxStar (r_tt
// -4) into out, counted in *count
__global__ void __launch_bounds__(kThreads) k_reloc_deferred(const int* __restrict__ list, int n,
                                                             const int* __restrict__ r_tt, int* __restrict__ out,
                                                             int* __restrict__ count) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int v = list[i];
        if (r_tt[v] == -4) out[agg_inc(count)] = v;
    }
}
// adds v to statistics slot k (its first shard)
__global__ void k_stat_add(unsigned long long* __restrict__ stats, int k, unsigned long long v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) stats[(size_t)k * kStatShards] += v;
}

// ------------------------------------------------------------ flip rounds
// kind 0 = Lawson (exact incircle), kind 1 = B/C untangle (reference steps).
// kFilter (Lawson on a box mesh): the FP32 filter alone, which keeps the
// kernel small (more waves in flight for the gathers of a round over every
// edge); an edge it cannot decide is queued for the next round, whose kernel
// has the exact stages.
template <int kKind, bool kAll, bool kF32, bool kBig, bool kFilter = false>
__global__ void __launch_bounds__(kThreads) k_round_test(DMesh m, const int* __restrict__ cand, int* __restrict__ cnt,
                                                         int round, int n_all, int* __restrict__ wants,
                                                         unsigned long long* __restrict__ claim, int* __restrict__ e_stamp,
                                                         int* __restrict__ cand_next, unsigned long long* __restrict__ stats,
                                                         const int* __restrict__ t_rep = nullptr, int rep_stamp = 0) {
    const int r3 = round % 3;
    const int n = kAll ? n_all : cnt[kCnCand + r3];
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[kCnCand + (r3 + 2) % 3] = 0;
    __shared__ AppendBuf<kListBuf> ab;
    append_init(ab);
    unsigned loc[4] = {0, 0, 0, 0};
    int first, end, stride;
    xcd_split(n, blockDim.x, first, end, stride);
    for (int base = first; base < end; base += stride) {
        const int i = base + threadIdx.x;
        bool want = false;
        int e = -1;
        do {  // one item; `break` leaves it without a flip request
        if (i >= end) break;
        e = kAll ? i : cand[i];
        const int2 tt = reinterpret_cast<const int2*>(m.he_t)[e];
        if (tt.x < 0 || tt.y < 0) break;
        const int2 vv = reinterpret_cast<const int2*>(m.he_v)[e];
        const int2 oo = reinterpret_cast<const int2*>(m.he_op)[e];
        if (kKind == 0) {
            int st = kStageF32;
            // inCircle(org(2e), op(2e+1), org(2e+1), op(2e)) > 0 (kernels.cuh:287-292)
            int s;
            if (kFilter) {
                s = incircle_f32_plane(m.pos[vv.x], m.pos[oo.y], m.pos[vv.y], m.pos[oo.x]);
                if (s == 0) s = kNeedBig;
            } else {
                s = incircle<kF32, kBig>(m.pos[vv.x], m.pos[oo.y], m.pos[vv.y], m.pos[oo.x], m.periodic, &st);
            }
            if (s == kNeedBig) {
                if (atomicExch(&e_stamp[e], round + 1) != round + 1)
                    cand_next[agg_inc(&cnt[kCnCand + (r3 + 1) % 3])] = e;
                break;
            }
            bump(loc, st);
            want = s > 0;
        } else if (kKind == 2) {
            if (t_rep[tt.x] == rep_stamp || t_rep[tt.y] == rep_stamp) break;
            // repair: one triangle not strictly CCW, the other strictly CCW, and
            // both triangles after the flip strictly CCW (B-step of
            // kernels.cuh:329-331 including its degenerate on-edge case)
            int st;
            float2 p0 = m.pos[vv.x], p1 = m.pos[vv.y], a0 = m.pos[oo.x], a1 = m.pos[oo.y];
            bool up0 = orient<kF32>(p0, p1, a0, m.periodic, &st) > 0; bump(loc, st);
            bool up1 = orient<kF32>(p1, p0, a1, m.periodic, &st) > 0; bump(loc, st);
            if (up0 != up1) {
                bool n0 = orient<kF32>(a0, a1, p1, m.periodic, &st) > 0; bump(loc, st);
                bool n1 = n0 && orient<kF32>(a1, a0, p0, m.periodic, &st) > 0;
                want = n0 && n1;
            }
        } else {
            // isCreased / B-step / C-step (kernels.cuh:308-338) with exact orientation;
            // t0 = (org, dst, op0), t1 = (dst, org, op1)
            int st;
            float2 p0 = m.pos[vv.x], p1 = m.pos[vv.y], a0 = m.pos[oo.x], a1 = m.pos[oo.y];
            int s0 = orient<kF32>(p0, p1, a0, m.periodic, &st); bump(loc, st);
            int s1 = orient<kF32>(p1, p0, a1, m.periodic, &st); bump(loc, st);
            bool ccw0 = s0 >= 0, ccw1 = s1 >= 0;
            if (ccw0 != ccw1) {
                // upright triangle U, inverted triangle I
                float2 u0 = ccw0 ? p0 : p1, u1 = ccw0 ? p1 : p0, ua = ccw0 ? a0 : a1;
                float2 op_inv = ccw0 ? a1 : a0;
                float2 i0 = ccw0 ? p1 : p0, i1 = ccw0 ? p0 : p1, ia = ccw0 ? a1 : a0;
                float2 op_up = ua;
                // B: op of the inverted triangle strictly inside the upright one
                bool inB = orient<kF32>(u0, u1, op_inv, m.periodic, &st) > 0 &&
                           orient<kF32>(u1, ua, op_inv, m.periodic, &st) > 0 &&
                           orient<kF32>(ua, u0, op_inv, m.periodic, &st) > 0;
                bool inC = false;
                if (!inB)  // C: op of the upright triangle strictly inside the inverted one
                    inC = orient<kF32>(i0, i1, op_up, m.periodic, &st) < 0 &&
                          orient<kF32>(i1, ia, op_up, m.periodic, &st) < 0 &&
                          orient<kF32>(ia, i0, op_up, m.periodic, &st) < 0;
                want = inB || inC;
            }
        }
        if (want) {
            unsigned long long key = claim_key(round, e);
            atomicMax(&claim[tt.x], key);
            atomicMax(&claim[tt.y], key);
        }
        } while (false);
        const int at = append_reserve(ab, want ? 1 : 0);
        if (want) ab.buf[at] = e;
        append_tile_end(ab, (int)blockDim.x, &cnt[kCnWants + r3], wants);
    }
    append_finish(ab, &cnt[kCnWants + r3], wants);
    flush_stats(loc, stats, kKind == 0 ? kStIncircle0 : kStOrient0);
}

__device__ __forceinline__ void enqueue_edge(int e, int stamp, int* e_stamp, int* cand_next, int* cnt, int slot) {
    if (atomicExch(&e_stamp[e], stamp) != stamp) cand_next[agg_inc(&cnt[slot])] = e;
}

__global__ void __launch_bounds__(kThreads) k_round_flip(DMesh m, int* __restrict__ cnt, int round,
                                                         const int* __restrict__ wants,
                                                         const unsigned long long* __restrict__ claim,
                                                         int* __restrict__ e_stamp, int* __restrict__ cand_next,
                                                         unsigned long long* __restrict__ stats, int stat_slot) {
    const int r3 = round % 3;
    const int n = cnt[kCnWants + r3];
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[kCnWants + (r3 + 1) % 3] = 0;
    const int slot = kCnCand + (r3 + 1) % 3, stamp = round + 1;
    __shared__ AppendBuf<kListBuf> ab;
    append_init(ab);
    unsigned loc[1] = {0};
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int i = base + threadIdx.x;
        // up to four edges to queue for the next round (the quad's, after a
        // flip; the edge itself, after a lost claim), each once per round
        int q0 = -1, q1 = -1, q2 = -1, q3 = -1;
        if (i < n) {
            const int e = wants[i];
            const int2 tt = reinterpret_cast<const int2*>(m.he_t)[e];
            const unsigned long long key = claim_key(round, e);
            if (claim[tt.x] == key && claim[tt.y] == key) {
                int h2, h3, h4, h5;
                flip_soa(m, e, h2, h3, h4, h5);
                loc[0]++;
                q0 = h2 >> 1; q1 = h3 >> 1; q2 = h4 >> 1; q3 = h5 >> 1;
            } else {
                q0 = e;
            }
            if (q0 >= 0 && atomicExch(&e_stamp[q0], stamp) == stamp) q0 = -1;
            if (q1 >= 0 && atomicExch(&e_stamp[q1], stamp) == stamp) q1 = -1;
            if (q2 >= 0 && atomicExch(&e_stamp[q2], stamp) == stamp) q2 = -1;
            if (q3 >= 0 && atomicExch(&e_stamp[q3], stamp) == stamp) q3 = -1;
        }
        const int nq = (q0 >= 0) + (q1 >= 0) + (q2 >= 0) + (q3 >= 0);
        int p = append_reserve(ab, nq);
        if (q0 >= 0) ab.buf[p++] = q0;
        if (q1 >= 0) ab.buf[p++] = q1;
        if (q2 >= 0) ab.buf[p++] = q2;
        if (q3 >= 0) ab.buf[p++] = q3;
        append_tile_end(ab, 4 * (int)blockDim.x, &cnt[slot], cand_next);
    }
    append_finish(ab, &cnt[slot], cand_next);
    flush_stats(loc, stats, stat_slot);
}

// all flip rounds of a small remaining queue inside one workgroup; uses the
// big exact stage (the workgroup's few waves bound its scratch memory)
template <int kKind>
__global__ void __launch_bounds__(1024) k_round_tail(DMesh m, int* __restrict__ cnt, int round0, int max_rounds,
                                                      int* __restrict__ cand0, int* __restrict__ cand1,
                                                      int* __restrict__ wants, unsigned long long* __restrict__ claim,
                                                      int* __restrict__ e_stamp, unsigned long long* __restrict__ stats,
                                                      int* __restrict__ rounds_done) {
    __shared__ int s_n, s_w, s_next;
    int round = round0;
    unsigned flips = 0, loc[4] = {0, 0, 0, 0};
    for (int it = 0; it < max_rounds; it++, round++) {
        int* cand = (round & 1) ? cand1 : cand0;
        int* cand_next = (round & 1) ? cand0 : cand1;
        if (threadIdx.x == 0) {
            s_n = cnt[kCnCand + round % 3];
            s_w = 0;
            s_next = 0;
        }
        __syncthreads();
        const int n = s_n;
        if (n == 0) break;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const int e = cand[i];
            const int2 tt = reinterpret_cast<const int2*>(m.he_t)[e];
            if (tt.x < 0 || tt.y < 0) continue;
            const int2 vv = reinterpret_cast<const int2*>(m.he_v)[e];
            const int2 oo = reinterpret_cast<const int2*>(m.he_op)[e];
            bool want;
            int st;
            if (kKind == 0) {
                int s = incircle<true, true>(m.pos[vv.x], m.pos[oo.y], m.pos[vv.y], m.pos[oo.x], m.periodic, &st);
                bump(loc, st);
                want = s > 0;
            } else {
                float2 p0 = m.pos[vv.x], p1 = m.pos[vv.y], a0 = m.pos[oo.x], a1 = m.pos[oo.y];
                bool ccw0 = orient<true>(p0, p1, a0, m.periodic, &st) >= 0;
                bool ccw1 = orient<true>(p1, p0, a1, m.periodic, &st) >= 0;
                want = false;
                if (ccw0 != ccw1) {
                    float2 u0 = ccw0 ? p0 : p1, u1 = ccw0 ? p1 : p0, ua = ccw0 ? a0 : a1, op_inv = ccw0 ? a1 : a0;
                    float2 i0 = ccw0 ? p1 : p0, i1 = ccw0 ? p0 : p1, ia = ccw0 ? a1 : a0;
                    bool inB = orient<true>(u0, u1, op_inv, m.periodic, &st) > 0 &&
                               orient<true>(u1, ua, op_inv, m.periodic, &st) > 0 &&
                               orient<true>(ua, u0, op_inv, m.periodic, &st) > 0;
                    bool inC = !inB && orient<true>(i0, i1, ua, m.periodic, &st) < 0 &&
                               orient<true>(i1, ia, ua, m.periodic, &st) < 0 &&
                               orient<true>(ia, i0, ua, m.periodic, &st) < 0;
                    want = inB || inC;
                }
            }
            if (want) {
                unsigned long long key = claim_key(round, e);
                atomicMax(&claim[tt.x], key);
                atomicMax(&claim[tt.y], key);
                wants[atomicAdd(&s_w, 1)] = e;
            }
        }
        __syncthreads();
        const int nw = s_w;
        for (int i = threadIdx.x; i < nw; i += blockDim.x) {
            const int e = wants[i];
            const int2 tt = reinterpret_cast<const int2*>(m.he_t)[e];
            const unsigned long long key = claim_key(round, e);
            const int stamp = round + 1;
            if (claim[tt.x] == key && claim[tt.y] == key) {
                int h2, h3, h4, h5;
                flip_soa(m, e, h2, h3, h4, h5);
                flips++;
                int es[4] = {h2 >> 1, h3 >> 1, h4 >> 1, h5 >> 1};
                for (int k = 0; k < 4; k++)
                    if (atomicExch(&e_stamp[es[k]], stamp) != stamp) cand_next[atomicAdd(&s_next, 1)] = es[k];
            } else if (atomicExch(&e_stamp[e], stamp) != stamp) {
                cand_next[atomicAdd(&s_next, 1)] = e;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) cnt[kCnCand + (round + 1) % 3] = s_next;
        __syncthreads();
    }
    if (threadIdx.x == 0) *rounds_done = round - round0;
    unsigned f[1] = {flips};
    flush_stats(f, stats, kKind == 0 ? kStFlips : kStUntangleFlips);
    flush_stats(loc, stats, kKind == 0 ? kStIncircle0 : kStOrient0);
}

// exact non-Delaunay edge count (diagnostic)
__global__ void __launch_bounds__(kThreads) k_count_nd(DMesh m, unsigned long long* out) {
    unsigned loc = 0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m.ne; e += gridDim.x * blockDim.x) {
        const int2 tt = reinterpret_cast<const int2*>(m.he_t)[e];
        if (tt.x < 0 || tt.y < 0) continue;
        const int2 vv = reinterpret_cast<const int2*>(m.he_v)[e];
        const int2 oo = reinterpret_cast<const int2*>(m.he_op)[e];
        int st;
        if (incircle<true, true>(m.pos[vv.x], m.pos[oo.y], m.pos[vv.y], m.pos[oo.x], m.periodic, &st) > 0) loc++;
    }
    if (loc) atomicAdd(out, (unsigned long long)loc);
}

// the first non-Delaunay edges (diagnostic): up to 8 edge ids into out[1..],
// their count into out[0]
__global__ void __launch_bounds__(kThreads) k_list_nd(DMesh m, int* out) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m.ne; e += gridDim.x * blockDim.x) {
        const int2 tt = reinterpret_cast<const int2*>(m.he_t)[e];
        if (tt.x < 0 || tt.y < 0) continue;
        const int2 vv = reinterpret_cast<const int2*>(m.he_v)[e];
        const int2 oo = reinterpret_cast<const int2*>(m.he_op)[e];
        int st;
        if (incircle<true, true>(m.pos[vv.x], m.pos[oo.y], m.pos[vv.y], m.pos[oo.x], m.periodic, &st) > 0) {
            const int k = atomicAdd(out, 1);
            if (k < 8) out[1 + k] = e;
        }
    }
}

__global__ void k_set_ints(int* p, int n, int v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_set_u64(unsigned long long* p, int n, unsigned long long v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------- one ring
// out-degree by CW rotation he <- next(he^1) from v.he (kernels.cuh:67-76)
// (both walks split over the XCDs: each L2 serves one contiguous range of
// vertices and the half-edges around them)
__global__ void __launch_bounds__(kThreads) k_degree(DMesh m, int* __restrict__ deg) {
    int first, last, stride;
    xcd_split(m.nv, blockDim.x, first, last, stride);
    for (int base = first; base < last; base += stride) {
        const int v = base + threadIdx.x;
        if (v >= last) break;
        int h0 = m.v_he[v], h = h0, c = 0;
        do {
            c++;
            h = m.he_next[h ^ 1];
        } while (h != h0 && c < 65536);
        deg[v] = c;
    }
}
// (rows of a wave's vertices gathered in LDS and written with contiguous
// stores when they fit, as k_star_ring_csr)
__global__ void __launch_bounds__(kThreads) k_ring_csr(DMesh m, const int* __restrict__ off, int* __restrict__ nbr, int cap) {
    __shared__ int s_row[kThreads / 64][kRingWave];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int first, last, stride;
    xcd_split(m.nv, blockDim.x, first, last, stride);
    // each wave stages its own rows (wave-level syncs)
    for (int base = first; base < last; base += stride) {
        const int v = base + threadIdx.x;
        const int wb = base + (threadIdx.x & ~63);
        int w0 = 0, wn = 0;
        if (wb < last) {
            w0 = off[wb];
            wn = off[min(wb + 64, last)] - w0;
        }
        const bool staged = wn <= kRingWave;
        if (v < last) {
            int h0 = m.v_he[v], h = h0, k = off[v], end = off[v + 1];
            do {
                const int x = m.he_v[h ^ 1];
                if (k < end) {
                    if (staged) s_row[wv][k - w0] = x;
                    else if (k < cap) nbr[k] = x;
                }
                k++;
                h = m.he_next[h ^ 1];
            } while (h != h0 && k < end);
        }
        wave_sync();
        if (staged)
            for (int k = lane; k < wn; k += 64)
                if (w0 + k < cap) nbr[w0 + k] = s_row[wv][k];
        wave_sync();
    }
}
// oneRing<K> (layout 0) / oneRing2<K> (layout 1): kernels.cuh:54-121
// (a vertex with more than K - 1 neighbours keeps the first K - 1, as the
// reference does, and is counted in *overflow unless it is a box corner,
// whose fan of hull points is not a point's neighbourhood)
__global__ void __launch_bounds__(kThreads) k_ring_legacy(DMesh m, int K, int layout, int* __restrict__ ring,
                                                          int* __restrict__ overflow) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        int h0 = m.v_he[v], h = h0, c = 0;
        do {
            if (c + 2 > K) {
                if (v >= m.ncorner) atomicAdd(overflow, 1);
                break;
            }
            int w = m.he_v[h ^ 1];
            if (layout == 0) ring[(size_t)(c + 1) * m.nv + v] = w;
            else ring[(size_t)(c + 1) + (size_t)K * v] = w;
            c++;
            h = m.he_next[h ^ 1];
        } while (h != h0);
        if (layout == 0) ring[v] = c;
        else ring[(size_t)K * v] = c;
    }
}
// closestNeighbors (kernels.cuh:129-146) over the CSR ring
__global__ void __launch_bounds__(kThreads) k_closest(DMesh m, const int* __restrict__ off, const int* __restrict__ nbr,
                                                      int* __restrict__ out) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        float2 p = m.pos[v];
        int b = off[v], e = off[v + 1];
        int best = b < e ? nbr[b] : -1;
        float bd = 3.0e38f;
        for (int k = b; k < e; k++) {
            float2 q = m.pos[nbr[k]];
            float dx = p.x - q.x, dy = p.y - q.y;
            if (m.periodic) { dx = min_image_f(dx); dy = min_image_f(dy); }
            float d = __fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy));
            if (k == b || d < bd) { bd = d; best = nbr[k]; }
        }
        out[v] = best;
    }
}

// Structural check of the half-edge mesh with every index range-checked
// before it is followed (so a broken mesh is reported, never dereferenced):
// next^3 = id on triangles, one face per cycle, dest(h) = org(next(h)), op(h) =
// org(next(next(h))), t.he and v.he point back.  err[0] += violations.
__global__ void __launch_bounds__(kThreads) k_validate(DMesh m, int* __restrict__ err) {
    int bad = 0;
    const int stride = gridDim.x * blockDim.x;
    for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < m.nhe; h += stride) {
        const int n1 = m.he_next[h];
        if ((unsigned)n1 >= (unsigned)m.nhe) { bad++; continue; }
        const int n2 = m.he_next[n1];
        if ((unsigned)n2 >= (unsigned)m.nhe) { bad++; continue; }
        const int t = m.he_t[h];
        if (t >= 0 && m.he_next[n2] != h) bad++;  // triangles are 3-cycles (the outer face of the box is a 4-cycle)
        if (t < -1 || t >= m.nt || m.he_t[n1] != t) bad++;
        const int v = m.he_v[h];
        if ((unsigned)v >= (unsigned)m.nv) bad++;
        if (m.he_v[h ^ 1] != m.he_v[n1]) bad++;
        if (t >= 0 && m.he_op[h] != m.he_v[n2]) bad++;
    }
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < m.nt; t += stride) {
        const int e = m.t_he[t];
        if ((unsigned)e >= (unsigned)m.nhe || m.he_t[e] != t) bad++;
    }
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += stride) {
        const int e = m.v_he[v];
        if ((unsigned)e >= (unsigned)m.nhe || m.he_v[e] != v) bad++;
    }
    if (bad) atomicAdd(err, bad);
}

// triangles that are not strictly CCW (exact; diagnostic)
__global__ void __launch_bounds__(kThreads) k_count_inv(DMesh m, int* __restrict__ err) {
    int bad = 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < m.nt; t += gridDim.x * blockDim.x) {
        const int h0 = m.t_he[t];
        if ((unsigned)h0 >= (unsigned)m.nhe) continue;
        const int h1 = m.he_next[h0], h2 = m.he_next[h1];
        int st;
        if (orient<true>(m.pos[m.he_v[h0]], m.pos[m.he_v[h1]], m.pos[m.he_v[h2]], m.periodic, &st) <= 0) bad++;
    }
    if (bad) atomicAdd(err, bad);
}

// ring entry c (0-based) of vertex v in the legacy layouts of k_ring_legacy
__device__ __forceinline__ int ring_at(const int* ring, int layout, int stride, int nv, int v, int c) {
    return layout == 0 ? ring[(size_t)(c + 1) * nv + v] : ring[(size_t)(c + 1) + (size_t)stride * v];
}
__device__ __forceinline__ int ring_count(const int* ring, int layout, int stride, int v) {
    return layout == 0 ? ring[v] : ring[(size_t)stride * v];
}

// closestNeighbors<K> / closestNeighbors2<K> (kernels.cuh:129-163) over the
// legacy SoA / AoS ring (AoS indexed as oneRing2 writes it)
__global__ void __launch_bounds__(kThreads) k_closest_legacy(DMesh m, int K, int layout, const int* __restrict__ ring,
                                                             int* __restrict__ out) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        const float2 p = m.pos[v];
        const int cnt = ring_count(ring, layout, K, v);
        int best = cnt > 0 ? ring_at(ring, layout, K, m.nv, v, 0) : -1;
        float bd = 3.0e38f;
        for (int c = 0; c < cnt; c++) {
            const int w = ring_at(ring, layout, K, m.nv, v, c);
            const float2 q = m.pos[w];
            float dx = p.x - q.x, dy = p.y - q.y;
            if (m.periodic) { dx = min_image_f(dx); dy = min_image_f(dy); }
            const float d = __fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy));
            if (c == 0 || d < bd) { bd = d; best = w; }
        }
        out[v] = best;
    }
}

// getFRNN<K, M> / getFRNN2<K, M> (kernels.cuh:167-272): vertices reachable
// from v through the Delaunay graph along vertices within r of v, found by a
// depth-first walk over the ring; squared FP32 distance <= r*r.  Output in the
// ring's layout with M slots per vertex: count in slot 0 (the full count),
// at most M-1 ids kept.  The visited test scans the ids found so far, like
// the reference; walks past kMaxFound ids stop growing the set (count stays
// exact up to there) and are reported through *overflow.
constexpr int kFrnnStack = 64, kMaxFound = 256;
__global__ void __launch_bounds__(kThreads) k_frnn_graph(DMesh m, float rr, int K, int layout,
                                                         const int* __restrict__ ring, int M, int* __restrict__ out,
                                                         int* __restrict__ overflow) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        const float2 p = m.pos[v];
        int stack[kFrnnStack];
        int found[kMaxFound];
        int ns = 0, nf = 0;
        bool over = false;
        stack[ns++] = v;
        while (ns > 0) {
            const int s = stack[--ns];
            const int cnt = ring_count(ring, layout, K, s);
            for (int c = 0; c < cnt; c++) {
                const int w = ring_at(ring, layout, K, m.nv, s, c);
                if (w == v) continue;
                bool seen = false;
                for (int k = 0; k < nf && !seen; k++) seen = found[k] == w;
                if (seen) continue;
                const float2 q = m.pos[w];
                float dx = p.x - q.x, dy = p.y - q.y;
                if (m.periodic) { dx = min_image_f(dx); dy = min_image_f(dy); }
                if (__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)) > rr) continue;
                if (nf == kMaxFound || ns == kFrnnStack) { over = true; continue; }
                if (nf + 1 < M) {
                    if (layout == 0) out[(size_t)(nf + 1) * m.nv + v] = w;
                    else out[(size_t)(nf + 1) + (size_t)M * v] = w;
                }
                found[nf++] = w;
                stack[ns++] = w;
            }
        }
        if (layout == 0) out[v] = nf;
        else out[(size_t)M * v] = nf;
        if (over) atomicAdd(overflow, 1);
    }
}

}  // namespace mcl

#include "mcl_star.hpp"
// unsigned long longs after the statistics shards: the star counters (s_aux)
constexpr int kStatsTail = (kStarAux + 1) / 2;

// =========================================================================
// host side
// =========================================================================
constexpr int kPathProbe = 64;  // first interval between measurements of the slower step path
constexpr int kBigBlocks = 256;   // workgroups of the rounds that run the big exact stage


// the tail of a timestep (mcl_tri_timestep): CSR one-ring and grid FRNN query
struct TailReq {
    int* ring_off;
    int* ring_nbrs;
    int ring_cap;
    mcl_grid_t* g;
    float rad;
    int* frnn_off;
    int* frnn_nbrs;
    int frnn_cap;
    mcl_bgrid_t* bg = nullptr;  // an ensemble's per-instance grids (instead of g)
    bool has_frnn() const { return g || bg; }
};

struct mcl_tri {
    DMesh m;
    // flip rounds
    unsigned long long* claim = nullptr;  // nt
    int* e_stamp = nullptr;               // ne
    int* cand[2] = {nullptr, nullptr};    // ne each
    int* wants = nullptr;                 // ne
    int* cnt = nullptr;                   // kCnCount
    int* rounds_done = nullptr;
    unsigned long long* stats = nullptr;  // kStCount
    unsigned long long* scalar = nullptr;
    // kinetic step
    float2* acc = nullptr;
    float2* tgt = nullptr;
    int* v_mv = nullptr;
    int* v_rev = nullptr;
    int* v_fexp = nullptr;
    int* t_stamp = nullptr;
    int* bad[2] = {nullptr, nullptr};  // nt each
    int* rev = nullptr;                // nv: reverted in the current pass
    int* allrev = nullptr;             // nv: reverted during the iteration
    int* pend[2] = {nullptr, nullptr}; // nv each
    int* v_touch = nullptr;            // nv: stamp of the last repair that touched the vertex
    int* t_rep = nullptr;              // nt: stamp of the repair iteration that made the triangle
    int* t_rep_e = nullptr;            // nt: edge of that repair flip
    int* e_undo = nullptr;             // ne: undo dedupe stamps
    int* undo = nullptr;               // ne: repair flips to undo in the current pass
    int* stuck = nullptr;              // nt
    int* touched = nullptr;            // nv: vertices of repair-flipped quads
    int* v_soft = nullptr;             // nv: step stamp of a vertex's soft take-back
    int* r_tt = nullptr;               // nv: relocation target triangle
    int* r_tt2 = nullptr;              // nv: second triangle when the target is on an edge
    int* r_wait = nullptr;             // nv: relocation retries (waits, walks) in this step
    int* r_defer = nullptr;            // nv: relocations deferred past the Lawson rounds
    int* tail_out = nullptr;           // 8: results of the single-workgroup loops
    // star recomputation (torus kinetic step)
    int* star_deg[2] = {nullptr, nullptr};  // nv each
    int* star_nbr[2] = {nullptr, nullptr};  // star_rows_size(nv, ncorner) each
    int* star_tdeg = nullptr;               // nv: recomputed stars by list position
    int* star_tnbr = nullptr;               // star_rows_size(nv, ncorner)
    int star_cur = 0;                       // stars of the current mesh
    bool stars_valid = false;               // star_*[star_cur] describe the current mesh
    bool mesh_stale = false;                // half-edge arrays not yet rebuilt from the stars
    int star_min_vertices = kStarMinVertices;  // MCL_OPT_STAR_MIN_VERTICES
    int star_enable = 2;                       // MCL_OPT_STAR_ENABLE: 0 off, 1 every step, 2 the faster path
    int star_give_up = 40;                     // MCL_OPT_STAR_GIVE_UP: percent of vertices the first check may list
    double path_ms[2] = {-1.0, -1.0};          // last step time of the flip-based (0) and star (1) paths
    int path_since = 0;                        // steps since the slower path was last measured
    int path_probe = kPathProbe;               // steps between such measurements
    int star_skip = 0;                      // steps left before the star step is tried again
    int star_backoff = 0;                   // length of the last such pause
    int* s_ecnt = nullptr;                  // nv + 1
    int* s_tcnt = nullptr;                  // nv + 1
    int* s_mark = nullptr;                  // nv
    int* s_list = nullptr;                  // nv
    int* s_rlist = nullptr;                 // nv: vertices to re-check
    int* s_aux = nullptr;                   // kStarAux counters (see star_step), inside the stats allocation
    int* s_heavy = nullptr;                 // nv + 1: count, then rows for the full star kernel
    int* s_heavy2 = nullptr;                // nv + 1: the first pass's second list (left by its wave-per-vertex kernel)
    int* s_hint = nullptr;                  // nv: a further candidate per vertex, left by the checks
    int* c_bucket = nullptr;                // ncorner * kCornerCap * 3: corner fan entries
    int* c_count = nullptr;                 // ncorner
    int* d_stamp = nullptr;                 // 1: base of the star checks' mark stamps (advanced on the device)
    mcl_step_params_t* d_params = nullptr;  // the step's parameters, read by the captured move
    mcl_step_params_t* h_params = nullptr;  // pinned staging copy of them
    int graph = 0;                          // MCL_OPT_GRAPH: star steps (and timesteps) replayed from hipGraphs
    hipGraphExec_t g_head[2] = {nullptr, nullptr};  // star-step head per star_cur
    hipGraphExec_t g_post[3] = {nullptr, nullptr, nullptr};  // timestep tail: ring CSR per star_cur, grid FRNN
    const void* post_key[8] = {};           // the buffers g_post was captured with
    bool capturing = false;                 // kernel timers are not recorded under capture
    bool captured = false;                  // a graph was captured during the current step
    hipStream_t cap_stream = nullptr;       // stream the graphs are captured on (the caller's may be the null stream)
    hipEvent_t ev_step = nullptr;           // start of a timed kinetic step (path choice)
    hipEvent_t ev_aux = nullptr;            // after a star step's last counter copy: the end of its own work
    hipEvent_t ev_end = nullptr;            // end of a timed flip-based step
    bool star_end_event = false;            // the current step settled by stars: ev_aux ends it
    const TailReq* tail_req = nullptr;      // the calling timestep's tail (its FRNN part enqueued at the star read)
    bool frnn_done = false;                 // ... and that FRNN query holds the step's final positions
    bool ring_done = false;                 // ... and its one-ring part ran on the step's final stars
    hipStream_t side_stream = nullptr;      // a star step's FRNN part, beside its retry passes
    hipEvent_t ev_fork = nullptr;           // caller's stream -> side stream (after the first star check)
    hipEvent_t ev_join = nullptr;           // side stream -> caller's stream (end of the FRNN part)
    bool side_pending = false;              // the caller's stream has not yet waited for ev_join
    bool frnn_on_targets = false;           // a flip-based step's FRNN part ran on the targets (side stream)
    int* d_offt = nullptr;                  // 8: vertices off their targets (k_off_target)
    // one ring
    int* deg = nullptr;
    int* scan_tmp = nullptr;
    int* h_cnt = nullptr;  // pinned
    unsigned long long* h_stats = nullptr;  // pinned copy of the statistics shards
    bool h_stats_fresh = false;             // h_stats holds the counters of the current step
    KTimer timers[6];      // Lawson full round, check all, move, relocation apply, take-back apply, star recomputation
    int tail_threshold = 8192;  // queue length below which one workgroup finishes the rounds
    int revert_tail = 1024;     // bad-triangle count below which one workgroup finishes the take-backs
    int reloc_tail = 1024;      // relocation list length below which one workgroup finishes the rounds
    int big_grid_min = 2048;    // big-exact queue length from which those rounds run on kBigBlocks workgroups
    int star_head_passes = 4;   // star passes enqueued (and captured) before the first host read
    int round = 1;         // global round counter (claims, edge stamps)
    int stamp = 1;         // global stamp for vertex / triangle marks
    int grid_blocks = 1024;
};

namespace mcl {
int hip_fail(hipError_t e, const char* what, const char* file, int line) {
    fprintf(stderr, "[mcleap_hip] %s failed: %s (%s:%d)\n", what, hipGetErrorString(e), file, line);
    return MCL_ERR_HIP;
}
}  // namespace mcl

static int blocks_for(long long n, int cap = 8192) {
    long long b = (n + kThreads - 1) / kThreads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

template <class T>
static int dalloc(T** p, size_t n) {
    MCL_HIP(hipMalloc((void**)p, sizeof(T) * (n ? n : 1)));
    return MCL_OK;
}

static int ensure_mesh(mcl_tri* t, hipStream_t s);

static int tri_free(mcl_tri* t) {
    void* ptrs[] = {t->m.pos, t->m.he_next, t->m.he_v, t->m.he_t, t->m.he_op, t->m.v_he, t->m.v_pidx, t->m.t_he,
                    t->claim, t->e_stamp, t->cand[0], t->cand[1], t->wants, t->cnt, t->rounds_done, t->stats,
                    t->scalar, t->acc, t->tgt, t->v_mv, t->v_rev, t->v_fexp, t->t_stamp, t->bad[0], t->bad[1],
                    t->rev, t->allrev, t->pend[0], t->pend[1], t->deg, t->scan_tmp, t->v_touch, t->t_rep, t->t_rep_e, t->e_undo, t->undo, t->stuck, t->touched, t->r_tt, t->r_tt2, t->r_wait, t->r_defer, t->v_soft, t->tail_out,
                    t->star_deg[0], t->star_deg[1], t->star_nbr[0], t->star_nbr[1], t->s_ecnt, t->s_tcnt,
                    t->s_mark, t->s_list, t->star_tdeg, t->star_tnbr, t->s_rlist, t->s_heavy, t->s_heavy2,
                    t->s_hint, t->c_bucket, t->c_count, t->d_stamp, t->d_params, t->d_offt};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (int k = 0; k < 2; k++)
        if (t->g_head[k]) (void)hipGraphExecDestroy(t->g_head[k]);
    for (int k = 0; k < 3; k++)
        if (t->g_post[k]) (void)hipGraphExecDestroy(t->g_post[k]);
    if (t->h_params) (void)hipHostFree(t->h_params);
    for (hipEvent_t e : {t->ev_step, t->ev_aux, t->ev_end, t->ev_fork, t->ev_join})
        if (e) (void)hipEventDestroy(e);
    if (t->cap_stream) (void)hipStreamDestroy(t->cap_stream);
    if (t->side_stream) (void)hipStreamDestroy(t->side_stream);
    if (t->h_cnt) (void)hipHostFree(t->h_cnt);
    if (t->h_stats) (void)hipHostFree(t->h_stats);
    return MCL_OK;
}

// Host-side structure check of AoS arrays before they reach the device
// (kernels follow these indices unchecked): every index in range, finite
// positions, next^3 = id around each triangle with one face per cycle,
// dest(h) = org(next(h)), op(h) = org(next(next(h))), the outer face's
// half-edges (t = -1) chained among themselves, v.he leaving its vertex and
// t.he lying on its triangle.  MCL_ERR_INPUT on the first violation.
static int check_aos(int nv, int nhe, int nt, const float* pos, const int* v2, const int* he4, const int* t1) {
    auto bad = [](const char* what, long long i) {
        fprintf(stderr, "mcleap: mesh arrays rejected: %s (element %lld)\n", what, i);
        return MCL_ERR_INPUT;
    };
    if (pos)
        for (long long i = 0; i < 2LL * nv; i++)
            if (!std::isfinite(pos[i])) return bad("position not finite", i / 2);
    if (!he4 && !v2 && !t1) return MCL_OK;
    if (!he4 || !v2 || !t1) return bad("half-edges, vertices and triangles must come together", 0);
    for (int h = 0; h < nhe; h++) {
        const int nx = he4[4 * h], v = he4[4 * h + 1], t = he4[4 * h + 2], op = he4[4 * h + 3];
        if (nx < 0 || nx >= nhe) return bad("next out of range", h);
        if (v < 0 || v >= nv) return bad("origin out of range", h);
        if (t < -1 || t >= nt) return bad("triangle out of range", h);
        if (op < -1 || op >= nv) return bad("opposite vertex out of range", h);
    }
    for (int h = 0; h < nhe; h++) {
        const int nx = he4[4 * h], t = he4[4 * h + 2];
        if (he4[4 * nx + 2] != t) return bad("next leaves the face", h);
        if (he4[4 * nx + 1] != he4[4 * (h ^ 1) + 1]) return bad("org(next(h)) != dest(h)", h);
        if (t >= 0) {
            const int n2 = he4[4 * nx];
            if (he4[4 * n2] != h) return bad("next^3 != id on a triangle", h);
            if (he4[4 * h + 3] != he4[4 * n2 + 1]) return bad("op(h) != org(next(next(h)))", h);
        }
    }
    for (int v = 0; v < nv; v++) {
        const int hv = v2[2 * v + 1];
        if (v2[2 * v] < 0 || v2[2 * v] >= nv) return bad("vertex position index out of range", v);
        if (hv < 0 || hv >= nhe || he4[4 * hv + 1] != v) return bad("v.he does not leave its vertex", v);
    }
    for (int t = 0; t < nt; t++) {
        const int ht = t1[t];
        if (ht < 0 || ht >= nhe || he4[4 * ht + 2] != t) return bad("t.he not on its triangle", t);
    }
    return MCL_OK;
}

static int upload_aos(mcl_tri* t, const float* pos, const int* v2, const int* he4, const int* t1) {
    DMesh& m = t->m;
    std::vector<int> a(m.nhe), b(m.nhe), c(m.nhe), d(m.nhe);
    if (he4) {
        for (int h = 0; h < m.nhe; h++) {
            a[h] = he4[4 * h]; b[h] = he4[4 * h + 1]; c[h] = he4[4 * h + 2]; d[h] = he4[4 * h + 3];
        }
        MCL_HIP(hipMemcpy(m.he_next, a.data(), sizeof(int) * m.nhe, hipMemcpyHostToDevice));
        MCL_HIP(hipMemcpy(m.he_v, b.data(), sizeof(int) * m.nhe, hipMemcpyHostToDevice));
        MCL_HIP(hipMemcpy(m.he_t, c.data(), sizeof(int) * m.nhe, hipMemcpyHostToDevice));
        MCL_HIP(hipMemcpy(m.he_op, d.data(), sizeof(int) * m.nhe, hipMemcpyHostToDevice));
    }
    if (v2) {
        std::vector<int> pi(m.nv), vh(m.nv);
        for (int v = 0; v < m.nv; v++) { pi[v] = v2[2 * v]; vh[v] = v2[2 * v + 1]; }
        MCL_HIP(hipMemcpy(m.v_pidx, pi.data(), sizeof(int) * m.nv, hipMemcpyHostToDevice));
        MCL_HIP(hipMemcpy(m.v_he, vh.data(), sizeof(int) * m.nv, hipMemcpyHostToDevice));
    }
    if (t1) MCL_HIP(hipMemcpy(m.t_he, t1, sizeof(int) * m.nt, hipMemcpyHostToDevice));
    if (pos) MCL_HIP(hipMemcpy(m.pos, pos, sizeof(float2) * m.nv, hipMemcpyHostToDevice));
    MCL_HIP(hipDeviceSynchronize());
    return MCL_OK;
}

// Allocates a triangulation of nv vertices, nhe half-edges and nt triangles
// (the mesh arrays are left for the caller to fill) and clears its work
// arrays.  ncorner: leading corner vertices of a box mesh (0 on the torus).
int tri_make(int nv, int nhe, int nt, int periodic, int ncorner, mcl_tri** out) {
    if (!out || nv <= 0 || nhe <= 0 || (nhe & 1) || nt <= 0) return MCL_ERR_INPUT;
    if (nhe / 2 >= (1 << 28)) return MCL_ERR_CAPACITY;  // claim keys hold 28-bit edge ids
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MCL_ERR_NO_DEVICE;
    mcl_tri* t = new mcl_tri();
    DMesh& m = t->m;
    m.nv = nv; m.nhe = nhe; m.nt = nt; m.ne = nhe / 2; m.periodic = periodic ? 1 : 0;
    m.ncorner = periodic ? 0 : ncorner;
    int rc = MCL_OK;
#define A(p, n) if (rc == MCL_OK) rc = dalloc(&(p), (size_t)(n))
    A(m.pos, nv); A(m.he_next, nhe); A(m.he_v, nhe); A(m.he_t, nhe); A(m.he_op, nhe);
    A(m.v_he, nv); A(m.v_pidx, nv); A(m.t_he, nt);
    A(t->claim, nt); A(t->e_stamp, m.ne); A(t->cand[0], m.ne); A(t->cand[1], m.ne); A(t->wants, m.ne);
    A(t->cnt, kCnCount); A(t->rounds_done, 1); A(t->stats, (size_t)kStCount * kStatShards + kStatsTail); A(t->scalar, 4);
    A(t->acc, nv); A(t->tgt, nv); A(t->v_mv, nv); A(t->v_rev, nv); A(t->v_fexp, nv); A(t->t_stamp, nt);
    A(t->bad[0], nt); A(t->bad[1], nt); A(t->rev, nv); A(t->allrev, nv); A(t->pend[0], nv); A(t->pend[1], nv);
    A(t->deg, nv + 1); A(t->scan_tmp, scan_scratch_ints(nv + 1) + 16);
    A(t->v_touch, nv); A(t->t_rep, nt); A(t->t_rep_e, nt); A(t->e_undo, m.ne); A(t->undo, m.ne);
    A(t->stuck, nt); A(t->touched, nv); A(t->r_tt, nv); A(t->r_tt2, nv); A(t->r_wait, nv); A(t->r_defer, nv); A(t->v_soft, nv); A(t->tail_out, 8); A(t->d_offt, 8);
    if (periodic || m.ncorner > 0) {
        A(t->star_deg[0], nv); A(t->star_deg[1], nv);
        // vertex rows, then the corners' fans (star_row); star_tnbr also holds
        // the rebuild's half-edge ids in the same layout
        A(t->star_nbr[0], star_rows_size(nv, m.ncorner)); A(t->star_nbr[1], star_rows_size(nv, m.ncorner));
        A(t->star_tdeg, nv); A(t->star_tnbr, star_rows_size(nv, m.ncorner));
        A(t->s_ecnt, nv + 1); A(t->s_tcnt, nv + 1); A(t->s_mark, nv); A(t->s_list, nv);
        A(t->s_rlist, nv); A(t->s_heavy, nv + 1); A(t->s_heavy2, nv + 1); A(t->s_hint, nv); A(t->d_stamp, 1); A(t->d_params, 1);
        if (m.ncorner > 0) { A(t->c_bucket, (size_t)m.ncorner * kCornerCap * 3); A(t->c_count, m.ncorner); }
    }
#undef A
    if (rc == MCL_OK && hipHostMalloc((void**)&t->h_cnt, sizeof(int) * kCnCount) != hipSuccess) rc = MCL_ERR_HIP;
    if (rc == MCL_OK && hipHostMalloc((void**)&t->h_params, sizeof(mcl_step_params_t)) != hipSuccess) rc = MCL_ERR_HIP;
    // the star counters live right after the statistics shards, so that one
    // copy brings both to the host
    if (rc == MCL_OK && t->star_deg[0]) t->s_aux = reinterpret_cast<int*>(t->stats + (size_t)kStCount * kStatShards);
    for (hipEvent_t* e : {&t->ev_step, &t->ev_aux, &t->ev_end})
        if (rc == MCL_OK && hipEventCreate(e) != hipSuccess) rc = MCL_ERR_HIP;
    for (hipEvent_t* e : {&t->ev_fork, &t->ev_join})
        if (rc == MCL_OK && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) rc = MCL_ERR_HIP;
    if (rc == MCL_OK &&
        hipHostMalloc((void**)&t->h_stats, sizeof(unsigned long long) * ((size_t)kStCount * kStatShards + kStatsTail)) !=
            hipSuccess)
        rc = MCL_ERR_HIP;
    if (rc == MCL_OK) {
        hipLaunchKernelGGL(k_set_u64, dim3(blocks_for(nt)), dim3(kThreads), 0, 0, t->claim, nt, 0ull);
        hipLaunchKernelGGL(k_set_ints, dim3(blocks_for(m.ne)), dim3(kThreads), 0, 0, t->e_stamp, m.ne, 0);
        hipLaunchKernelGGL(k_set_ints, dim3(blocks_for(nv)), dim3(kThreads), 0, 0, t->v_mv, nv, 0);
        hipLaunchKernelGGL(k_set_ints, dim3(blocks_for(nv)), dim3(kThreads), 0, 0, t->v_rev, nv, 0);
        hipLaunchKernelGGL(k_set_ints, dim3(blocks_for(nv)), dim3(kThreads), 0, 0, t->v_fexp, nv, 0);
        hipLaunchKernelGGL(k_set_ints, dim3(blocks_for(nv)), dim3(kThreads), 0, 0, t->v_touch, nv, 0);
        hipLaunchKernelGGL(k_set_ints, dim3(blocks_for(nv)), dim3(kThreads), 0, 0, t->v_soft, nv, 0);
        hipLaunchKernelGGL(k_set_ints, dim3(blocks_for(nt)), dim3(kThreads), 0, 0, t->t_rep, nt, 0);
        hipLaunchKernelGGL(k_set_ints, dim3(blocks_for(m.ne)), dim3(kThreads), 0, 0, t->e_undo, m.ne, 0);
        hipLaunchKernelGGL(k_set_ints, dim3(blocks_for(nt)), dim3(kThreads), 0, 0, t->t_stamp, nt, 0);
        hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, 0, t->cnt, (int)kCnCount, 0);
        hipLaunchKernelGGL(k_set_u64, dim3(4), dim3(256), 0, 0, t->stats, (int)(kStCount * kStatShards), 0ull);
        if (t->s_mark) hipLaunchKernelGGL(k_set_ints, dim3(blocks_for(nv)), dim3(kThreads), 0, 0, t->s_mark, nv, 0);
        if (t->d_stamp) hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, 0, t->d_stamp, 1, 0);
        if (hipDeviceSynchronize() != hipSuccess) rc = MCL_ERR_HIP;
    }
    if (rc != MCL_OK) { tri_free(t); delete t; return rc; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess) t->grid_blocks = prop.multiProcessorCount * 8;
    *out = t;
    return MCL_OK;
}


extern "C" int mcl_tri_create(const float* pos, int npos, const int* v2, int nv, const int* he4, int nhe,
                              const int* t1, int nt, int periodic, mcl_tri_t** out) {
    if (!out || !pos || !v2 || !he4 || !t1 || npos != nv || nhe <= 0 || (nhe & 1) || nt <= 0) return MCL_ERR_INPUT;
    // box meshes keep their corners first: vertices 0..3, or 4i..4i+3 for
    // instance i of an ensemble (the star path needs it)
    int ncorner = 0;
    if (!periodic && nv > 4) {
        auto corner_at = [&](int v) {
            const int pi = v2[2 * v];
            return pi >= 0 && pi < npos && (fabsf(pos[2 * pi]) >= 1.0e6f || fabsf(pos[2 * pi + 1]) >= 1.0e6f);
        };
        int k = 0;
        while (k < nv && corner_at(k)) k++;
        bool ok = k > 0 && k % 4 == 0;
        for (int v = k; ok && v < nv; v++) ok = !corner_at(v);
        if (ok) ncorner = k;
    }
    MCL_TRY(check_aos(nv, nhe, nt, pos, v2, he4, t1));
    mcl_tri* t = nullptr;
    int rc = tri_make(nv, nhe, nt, periodic, ncorner, &t);
    if (rc != MCL_OK) return rc;
    rc = upload_aos(t, pos, v2, he4, t1);
    if (rc != MCL_OK) { tri_free(t); delete t; return rc; }
    tune_tri(t, periodic, ncorner, nv);
    *out = t;
    return MCL_OK;
}

extern "C" int mcl_tri_destroy(mcl_tri_t* t) {
    if (!t) return MCL_OK;
    (void)hipDeviceSynchronize();
    tri_free(t);
    delete t;
    return MCL_OK;
}

extern "C" int mcl_tri_sizes(const mcl_tri_t* t, int* npos, int* nv, int* nhe, int* nt) {
    if (!t) return MCL_ERR_INPUT;
    if (npos) *npos = t->m.nv;
    if (nv) *nv = t->m.nv;
    if (nhe) *nhe = t->m.nhe;
    if (nt) *nt = t->m.nt;
    return MCL_OK;
}

extern "C" void* mcl_tri_pos_ptr(mcl_tri_t* t) { return t ? (void*)t->m.pos : nullptr; }

extern "C" int mcl_tri_upload(mcl_tri_t* t, const float* pos, const int* v2, const int* he4, const int* t1) {
    if (!t) return MCL_ERR_INPUT;
    MCL_TRY(check_aos(t->m.nv, t->m.nhe, t->m.nt, pos, v2, he4, t1));
    t->stars_valid = false;
    t->mesh_stale = false;
    return upload_aos(t, pos, v2, he4, t1);
}

int mcl::tri_counters(mcl_tri_t* t, int* round, int* stamp, bool set) {
    if (!t || !round || !stamp) return MCL_ERR_INPUT;
    if (set) {
        t->round = *round;
        t->stamp = *stamp;
    } else {
        *round = t->round;
        *stamp = t->stamp;
    }
    return MCL_OK;
}

extern "C" int mcl_tri_is_periodic(const mcl_tri_t* t, int* periodic) {
    if (!t || !periodic) return MCL_ERR_INPUT;
    *periodic = t->m.periodic;
    return MCL_OK;
}

extern "C" int mcl_tri_download(mcl_tri_t* t, float* pos, int* v2, int* he4, int* t1) {
    if (!t) return MCL_ERR_INPUT;
    DMesh& m = t->m;
    MCL_HIP(hipDeviceSynchronize());
    MCL_TRY(ensure_mesh(t, 0));
    if (he4) {
        std::vector<int> a(m.nhe), b(m.nhe), c(m.nhe), d(m.nhe);
        MCL_HIP(hipMemcpy(a.data(), m.he_next, sizeof(int) * m.nhe, hipMemcpyDeviceToHost));
        MCL_HIP(hipMemcpy(b.data(), m.he_v, sizeof(int) * m.nhe, hipMemcpyDeviceToHost));
        MCL_HIP(hipMemcpy(c.data(), m.he_t, sizeof(int) * m.nhe, hipMemcpyDeviceToHost));
        MCL_HIP(hipMemcpy(d.data(), m.he_op, sizeof(int) * m.nhe, hipMemcpyDeviceToHost));
        for (int h = 0; h < m.nhe; h++) {
            he4[4 * h] = a[h]; he4[4 * h + 1] = b[h]; he4[4 * h + 2] = c[h]; he4[4 * h + 3] = d[h];
        }
    }
    if (v2) {
        std::vector<int> pi(m.nv), vh(m.nv);
        MCL_HIP(hipMemcpy(pi.data(), m.v_pidx, sizeof(int) * m.nv, hipMemcpyDeviceToHost));
        MCL_HIP(hipMemcpy(vh.data(), m.v_he, sizeof(int) * m.nv, hipMemcpyDeviceToHost));
        for (int v = 0; v < m.nv; v++) { v2[2 * v] = pi[v]; v2[2 * v + 1] = vh[v]; }
    }
    if (t1) MCL_HIP(hipMemcpy(t1, m.t_he, sizeof(int) * m.nt, hipMemcpyDeviceToHost));
    if (pos) MCL_HIP(hipMemcpy(pos, m.pos, sizeof(float2) * m.nv, hipMemcpyDeviceToHost));
    return MCL_OK;
}

extern "C" int mcl_tri_move(mcl_tri_t* t, const void* d_disp, void* stream) {
    if (!t || !d_disp) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_move_disp, dim3(blocks_for(t->m.nv)), dim3(kThreads), 0, s, t->m, (const float2*)d_disp);
    MCL_HIP(hipGetLastError());
    MCL_HIP(hipStreamSynchronize(s));
    return MCL_OK;
}

extern "C" int mcl_tri_integrate(mcl_tri_t* t, const void* d_dt, const void* d_vel, void* stream) {
    if (!t || !d_dt || !d_vel) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_integrate, dim3(blocks_for(t->m.nv)), dim3(kThreads), 0, s, t->m, (const float*)d_dt,
                       (const float2*)d_vel);
    MCL_HIP(hipGetLastError());
    MCL_HIP(hipStreamSynchronize(s));
    return MCL_OK;
}

// moveFlipflop on the host arrays: vertex v removed and re-inserted at
// pos + d by the relocation's removal / insertion (reloc_core), then, for
// MCL_MOVE_FLIPFLOP_DELAUNAY, Lawson flips from the edges of the triangles
// it rewrote; MCL_MOVE_UNTANGLING(_DELAUNAY): pos += d, B/C untangle passes
// until none flips, then (if no triangle stays inverted) Lawson flips.
extern "C" int mcl_hmesh_move_vertex(float* pos, int* he4, int nhe, int* t1, int nt, int* v2, int nv, int periodic,
                                     int v, const float* d, int mode, long long* flips) {
    if (!pos || !he4 || !t1 || !v2 || !d || v < 0 || v >= nv || mode < 0 || mode > 3) return MCL_ERR_INPUT;
    if (!std::isfinite(d[0]) || !std::isfinite(d[1])) return MCL_ERR_INPUT;
    MCL_TRY(check_aos(nv, nhe, nt, nullptr, v2, he4, t1));
    long long nf = 0;
    if (flips) *flips = 0;
    const int pv = v2[2 * v];
    float2 pt = make_float2(pos[2 * pv] + d[0], pos[2 * pv + 1] + d[1]);
    if (periodic) {
        pt.x -= floorf(pt.x);
        pt.y -= floorf(pt.y);
        if (pt.x >= 1.0f) pt.x = 0.0f;
        if (pt.y >= 1.0f) pt.y = 0.0f;
    }
    if (mode >= 2) {
        pos[2 * pv] = pt.x;
        pos[2 * pv + 1] = pt.y;
        for (int pass = 0; pass < 1024; pass++) {
            long long f = 0;
            MCL_TRY(mcl_hmesh_untangle(pos, he4, nhe, t1, nt, v2, nv, periodic, &f));
            nf += f;
            if (f == 0) break;
        }
        if (flips) *flips = nf;
        for (int t = 0; t < nt; t++) {  // any triangle still not strictly CCW?
            const int h0 = t1[t], h1 = he4[4 * h0], h2 = he4[4 * h1];
            const int a = v2[2 * he4[4 * h0 + 1]], b = v2[2 * he4[4 * h1 + 1]], c = v2[2 * he4[4 * h2 + 1]];
            if (mcl_orient2d(pos + 2 * a, pos + 2 * b, pos + 2 * c, periodic) <= 0) return MCL_ERR_UNRESOLVED;
        }
        if (mode == 3) {
            long long f = 0;
            const int rc = mcl_hmesh_delonize(pos, he4, nhe, t1, nt, v2, nv, periodic, nullptr, 0, &f);
            if (flips) *flips = nf + f;
            return rc;
        }
        return MCL_OK;
    }
    // the mesh as SoA host arrays (positions per vertex)
    std::vector<float2> P(nv);
    std::vector<int> nx(nhe), hv(nhe), ht(nhe), ho(nhe), vh(nv), th(t1, t1 + nt);
    for (int i = 0; i < nv; i++) {
        P[i] = make_float2(pos[2 * v2[2 * i]], pos[2 * v2[2 * i] + 1]);
        vh[i] = v2[2 * i + 1];
    }
    for (int h = 0; h < nhe; h++) {
        nx[h] = he4[4 * h]; hv[h] = he4[4 * h + 1]; ht[h] = he4[4 * h + 2]; ho[h] = he4[4 * h + 3];
    }
    DMesh m;
    m.nv = nv; m.nhe = nhe; m.nt = nt; m.ne = nhe / 2; m.periodic = periodic != 0;
    m.pos = P.data(); m.he_next = nx.data(); m.he_v = hv.data(); m.he_t = ht.data(); m.he_op = ho.data();
    m.v_he = vh.data(); m.t_he = th.data();
    if (pt.x == P[v].x && pt.y == P[v].y) return MCL_OK;
    // v's star: every triangle around it (a vertex on the outer face cannot be moved this way)
    int star[kMaxStar], deg = 0;
    {
        const int h0 = vh[v];
        int h = h0;
        do {
            if (deg >= kMaxStar || ht[h] < 0) return MCL_ERR_INPUT;
            star[deg++] = ht[h];
            h = nx[h ^ 1];
        } while (h != h0);
    }
    int on = -1, tt = -5;
    for (unsigned seed = 1; tt == -5 && seed <= 16; seed++) tt = locate_point<true>(m, pt, ht[vh[v]], &on, seed * 7919u + v);
    if (tt == -2) return MCL_ERR_DUPLICATE;  // the target is another vertex's position
    if (tt < 0) return MCL_ERR_INPUT;        // outside the triangulated domain
    const int tt2 = on >= 0 ? ht[on ^ 1] : -1;
    const int rc = reloc_core<true>(m, v, star, deg, tt, tt2, pt);
    // back to the caller's AoS arrays
    for (int i = 0; i < nv; i++) {
        v2[2 * i + 1] = vh[i];
        pos[2 * v2[2 * i]] = P[i].x;
        pos[2 * v2[2 * i] + 1] = P[i].y;
    }
    for (int h = 0; h < nhe; h++) {
        he4[4 * h] = nx[h]; he4[4 * h + 1] = hv[h]; he4[4 * h + 2] = ht[h]; he4[4 * h + 3] = ho[h];
    }
    for (int t = 0; t < nt; t++) t1[t] = th[t];
    if (rc != 0) return MCL_ERR_UNRESOLVED;
    if (mode == 1) {
        // Lawson flips from the half-edges of every triangle the relocation rewrote
        std::vector<int> seeds;
        auto add = [&](int T) {
            if (T < 0) return;
            int h = th[T];
            for (int k = 0; k < 3; k++, h = nx[h]) seeds.push_back(h);
        };
        for (int k = 0; k < deg; k++) add(star[k]);
        add(tt);
        add(tt2);
        return mcl_hmesh_delonize(pos, he4, nhe, t1, nt, v2, nv, periodic, seeds.data(), (int)seeds.size(), flips);
    }
    return MCL_OK;
}

// targets of an explicit move of listed vertices: every vertex's own
// position, then pos + d (wrapped on the torus) for the listed ones, whose
// retry counts are cleared
__global__ void __launch_bounds__(kThreads) k_targets_here(DMesh m, float2* __restrict__ tgt) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) tgt[v] = m.pos[v];
}
__global__ void __launch_bounds__(kThreads) k_targets_listed(DMesh m, const int* __restrict__ ids,
                                                             const float2* __restrict__ d, int n, int apply,
                                                             float2* __restrict__ tgt, int* __restrict__ r_wait) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int v = ids[i];
        const float2 q = m.pos[v];
        float2 x = make_float2(q.x + d[i].x, q.y + d[i].y);
        if (m.periodic) x = make_float2(wrap01(x.x), wrap01(x.y));
        tgt[v] = x;
        r_wait[v] = 0;
        if (apply) m.pos[v] = x;
    }
}

static int read_counts(mcl_tri* t, hipStream_t s);
static int relocate_rounds(mcl_tri* t, hipStream_t s, bool f32, int n, int* slot, int* rounds);
static int reloc_deferred(mcl_tri* t, hipStream_t s, bool f32, int variant, const int* list, int n, bool lawson,
                          long long* rounds, int* reloc_rounds);
template <int kKind>
static int run_rounds(mcl_tri* t, bool all, int variant, int max_rounds, hipStream_t s, long long* rounds_out);
static void stats_begin(mcl_tri* t, hipStream_t s);
static int stats_end(mcl_tri* t, hipStream_t s, mcl_stats_t* st, std::chrono::steady_clock::time_point t0);

extern "C" int mcl_tri_move_vertices(mcl_tri_t* t, const int* ids, const float* disp, int n, int mode,
                                     mcl_stats_t* st, void* stream) {
    if (!t || n < 0 || (n > 0 && (!ids || !disp)) || mode < 0 || mode > 3 || n > t->m.nv) return MCL_ERR_INPUT;
    {
        std::vector<int> chk(ids, ids + n);
        std::sort(chk.begin(), chk.end());
        for (int i = 0; i < n; i++)
            if (chk[i] < t->m.ncorner || chk[i] >= t->m.nv || (i && chk[i] == chk[i - 1])) return MCL_ERR_INPUT;
        for (int i = 0; i < 2 * n; i++)
            if (!std::isfinite(disp[i])) return MCL_ERR_INPUT;
    }
    hipStream_t s = (hipStream_t)stream;
    MCL_TRY(ensure_mesh(t, s));
    t->stars_valid = false;
    auto t0 = std::chrono::steady_clock::now();
    stats_begin(t, s);
    if (n == 0) return stats_end(t, s, st, t0);
    DMesh& m = t->m;
    // the ids into the relocation list, the displacements into acc (free between steps)
    MCL_HIP(hipMemcpyAsync(t->pend[0], ids, sizeof(int) * n, hipMemcpyHostToDevice, s));
    MCL_HIP(hipMemcpyAsync(t->acc, disp, sizeof(float2) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_targets_here, dim3(blocks_for(m.nv)), dim3(kThreads), 0, s, m, t->tgt);
    hipLaunchKernelGGL(k_targets_listed, dim3(blocks_for(n)), dim3(kThreads), 0, s, m, (const int*)t->pend[0],
                       (const float2*)t->acc, n, mode >= 2 ? 1 : 0, t->tgt, t->r_wait);
    MCL_HIP(hipGetLastError());
    int rc = MCL_OK;
    long long rounds = 0;
    if (mode < 2) {
        // the list again in allrev: relocate_rounds consumes pend[0] and pend[1]
        MCL_HIP(hipMemcpyAsync(t->allrev, t->pend[0], sizeof(int) * n, hipMemcpyDeviceToDevice, s));
        int slot = 0, rr = 0;
        rc = relocate_rounds(t, s, true, n, &slot, &rr);
        if (rc == MCL_OK && mode == 1) rc = run_rounds<0>(t, true, 0, 0, s, &rounds);
        if (rc == MCL_OK) rc = reloc_deferred(t, s, true, 0, t->allrev, n, mode == 1, &rounds, &rr);
    } else {
        rc = run_rounds<1>(t, true, 1, 0, s, &rounds);
        if (rc == MCL_ERR_ROUND_CAP) rc = MCL_OK;
        long long inv = 0;
        if (rc == MCL_OK) rc = mcl_tri_count_inverted(t, &inv, stream);
        if (rc == MCL_OK && inv) rc = MCL_ERR_UNRESOLVED;
        if (rc == MCL_OK && mode == 3) rc = run_rounds<0>(t, true, 0, 0, s, &rounds);
    }
    mcl_stats_t local{};
    mcl_stats_t* sp = st ? st : &local;
    sp->rounds = rounds;
    const int rc2 = stats_end(t, s, sp, t0);
    if (rc == MCL_OK && (sp->stalled > 0 || sp->duplicates > 0)) rc = MCL_ERR_UNRESOLVED;
    return rc != MCL_OK ? rc : rc2;
}

// ------------------------------------------------------------ round driver
static int read_counts(mcl_tri* t, hipStream_t s) {
    MCL_HIP(hipMemcpyAsync(t->h_cnt, t->cnt, sizeof(int) * kCnCount, hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    return MCL_OK;
}

// zero the counter slots used by the first round (round % 3 = r3):
// wants[r3] and cand[r3+1]; cand[r3] is set by the caller
__global__ void k_prep_round(int* cnt, int r3, int ncand) {
    if (threadIdx.x == 0) {
        cnt[kCnWants + r3] = 0;
        cnt[kCnCand + (r3 + 1) % 3] = 0;
        if (ncand >= 0) cnt[kCnCand + r3] = ncand;
    }
}

// Runs flip rounds (kind 0 = Lawson, 1 = B/C) starting from either all
// edges (all = true) or the queue already in cand[round % 2] with count in
// cnt[kCnCand + round % 3].  variant 0: full-grid rounds while the queue is
// large, then one workgroup finishes the rest (no host round trips until
// the end).  variant 1: one test + one flip launch per round with a host
// read of the queue length in between (FP64-first predicates).
static bool trace_on();
static bool validate_on();

// prints the first non-Delaunay edges with their quads (diagnostic)
static void trace_nd_edges(mcl_tri* t, hipStream_t s) {
    const DMesh& m = t->m;
    int* d = nullptr;
    if (hipMalloc(&d, sizeof(int) * 9) != hipSuccess) return;
    int h[9] = {0};
    hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, d, 9, 0);
    hipLaunchKernelGGL(k_list_nd, dim3(256), dim3(kThreads), 0, s, m, d);
    if (hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s) == hipSuccess && hipStreamSynchronize(s) == hipSuccess) {
        fprintf(stderr, "  [lawson cap] %d non-Delaunay edges\n", h[0]);
        for (int i = 0; i < std::min(h[0], 8); i++) {
            const int e = h[1 + i];
            int v[2], o[2];
            float2 p[4];
            (void)hipMemcpy(v, m.he_v + 2 * e, sizeof(v), hipMemcpyDeviceToHost);
            (void)hipMemcpy(o, m.he_op + 2 * e, sizeof(o), hipMemcpyDeviceToHost);
            const int ids[4] = {v[0], o[1], v[1], o[0]};
            for (int k = 0; k < 4; k++) (void)hipMemcpy(&p[k], m.pos + ids[k], sizeof(float2), hipMemcpyDeviceToHost);
            fprintf(stderr, "    edge %d: %d (%.9g, %.9g) %d (%.9g, %.9g) %d (%.9g, %.9g) %d (%.9g, %.9g)\n", e, ids[0],
                    p[0].x, p[0].y, ids[1], p[1].x, p[1].y, ids[2], p[2].x, p[2].y, ids[3], p[3].x, p[3].y);
        }
    }
    (void)hipFree(d);
}

template <int kKind>
static int run_rounds(mcl_tri* t, bool all, int variant, int max_rounds, hipStream_t s, long long* rounds_out) {
    DMesh& m = t->m;
    const int cap = max_rounds > 0 ? max_rounds : 4096;
    int r = 0;
    if (all) {
        hipLaunchKernelGGL(k_prep_round, dim3(1), dim3(64), 0, s, t->cnt, t->round % 3, 0);
    }
    bool first = all;
    long long done = 0;
    if (variant == 1) {
        bool need_tail = false;
        for (; r < cap; r++) {
            const int R = t->round;
            int* cand = t->cand[R & 1];
            int* cand_next = t->cand[(R + 1) & 1];
            const int stat = kKind == 0 ? kStFlips : kStUntangleFlips;
            if (first) {
                hipLaunchKernelGGL((k_round_test<kKind, true, false, false>), dim3(blocks_for(m.ne)), dim3(kThreads), 0, s,
                                   m, cand, t->cnt, R, m.ne, t->wants, t->claim, t->e_stamp, cand_next, t->stats);
            } else {
                int n = t->h_cnt[kCnCand + R % 3];
                hipLaunchKernelGGL((k_round_test<kKind, false, false, false>), dim3(blocks_for(n)), dim3(kThreads), 0, s,
                                   m, cand, t->cnt, R, 0, t->wants, t->claim, t->e_stamp, cand_next, t->stats);
            }
            MCL_TRY(read_counts(t, s));
            int nw = t->h_cnt[kCnWants + R % 3];
            hipLaunchKernelGGL(k_round_flip, dim3(blocks_for(nw)), dim3(kThreads), 0, s, m, t->cnt, R, t->wants, t->claim,
                               t->e_stamp, cand_next, t->stats, stat);
            MCL_HIP(hipGetLastError());
            MCL_TRY(read_counts(t, s));
            t->round++;
            done++;
            first = false;
            if (t->h_cnt[kCnCand + (R + 1) % 3] == 0) { r++; break; }
            if (nw == 0) { need_tail = true; r++; break; }  // only big-exact cases left
        }
        if (!need_tail) {
            if (rounds_out) *rounds_out = done;
            if (r >= cap && t->h_cnt[kCnCand + t->round % 3] != 0) return MCL_ERR_ROUND_CAP;
            return MCL_OK;
        }
        const int R = t->round;
        hipLaunchKernelGGL(k_round_tail<kKind>, dim3(1), dim3(1024), 0, s, m, t->cnt, R, cap - r + 1, t->cand[0],
                           t->cand[1], t->wants, t->claim, t->e_stamp, t->stats, t->rounds_done);
        MCL_HIP(hipGetLastError());
        int tail_rounds = 0;
        MCL_HIP(hipMemcpyAsync(&tail_rounds, t->rounds_done, sizeof(int), hipMemcpyDeviceToHost, s));
        MCL_TRY(read_counts(t, s));
        t->round = R + tail_rounds + 1;
        done += tail_rounds;
        if (rounds_out) *rounds_out = done;
        if (t->h_cnt[kCnCand + (R + tail_rounds) % 3] != 0) return MCL_ERR_ROUND_CAP;
        return MCL_OK;
    }
    // variant 0: full-grid rounds while the queue is large, then the
    // single-workgroup tail finishes the small remainder
    const int gb = t->grid_blocks;
    // MCL_TRACE and MCL_VALIDATE set: every round on the full grid, each
    // followed by an exact count of the triangles that are not strictly CCW
    const bool watch = trace_on() && validate_on();
    int* d_inv = nullptr;
    if (watch && hipMalloc(&d_inv, sizeof(int)) != hipSuccess) d_inv = nullptr;
    for (int k = 0; r < cap; k++, r++) {
        if (k > 0) {
            MCL_TRY(read_counts(t, s));
            const int left = t->h_cnt[kCnCand + t->round % 3];
            // no flip wanted last round: what is queued needs the big exact
            // stage (below)
            if (t->h_cnt[kCnWants + (t->round - 1) % 3] == 0 && left > 0) break;
            if (d_inv) {
                int inv = 0;
                hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, d_inv, 1, 0);
                hipLaunchKernelGGL(k_count_inv, dim3(gb), dim3(kThreads), 0, s, m, d_inv);
                MCL_HIP(hipMemcpyAsync(&inv, d_inv, sizeof(int), hipMemcpyDeviceToHost, s));
                MCL_HIP(hipStreamSynchronize(s));
                if (inv || k < 4 || k % 256 == 0)
                    fprintf(stderr, "  [lawson round %d] queue %d, triangles not strictly CCW %d\n", k, left, inv);
                if (inv) trace_nd_edges(t, s);
                if (inv || left == 0) break;
            } else if (left < t->tail_threshold) {
                break;
            }
        }
        const int R = t->round;
        int* cand = t->cand[R & 1];
        int* cand_next = t->cand[(R + 1) & 1];
        const int stat = kKind == 0 ? kStFlips : kStUntangleFlips;
        if (first) {
            t->timers[0].mark(s);
            if (kKind == 0 && !m.periodic)
                hipLaunchKernelGGL((k_round_test<kKind, true, true, false, true>), dim3(gb), dim3(kThreads), 0, s, m, cand,
                                   t->cnt, R, m.ne, t->wants, t->claim, t->e_stamp, cand_next, t->stats);
            else
                hipLaunchKernelGGL((k_round_test<kKind, true, true, false>), dim3(gb), dim3(kThreads), 0, s, m, cand,
                                   t->cnt, R, m.ne, t->wants, t->claim, t->e_stamp, cand_next, t->stats);
            t->timers[0].mark(s);
        }
        else
            hipLaunchKernelGGL((k_round_test<kKind, false, true, false>), dim3(gb), dim3(kThreads), 0, s, m, cand, t->cnt, R, 0,
                               t->wants, t->claim, t->e_stamp, cand_next, t->stats);
        hipLaunchKernelGGL(k_round_flip, dim3(gb), dim3(kThreads), 0, s, m, t->cnt, R, t->wants, t->claim, t->e_stamp,
                           cand_next, t->stats, stat);
        t->round++;
        done++;
        first = false;
    }
    if (d_inv) (void)hipFree(d_inv);
    // a long queue of cases for the big exact stage (box meshes with many
    // points on the sides, next to the corners): rounds on kBigBlocks
    // workgroups, whose per-lane private arrays bound how many can run; the
    // single-workgroup tail takes the rest
    for (; r < cap; r++) {
        MCL_TRY(read_counts(t, s));
        if (t->h_cnt[kCnCand + t->round % 3] < t->big_grid_min) break;
        const int R = t->round;
        const int stat = kKind == 0 ? kStFlips : kStUntangleFlips;
        hipLaunchKernelGGL((k_round_test<kKind, false, true, true>), dim3(kBigBlocks), dim3(kThreads), 0, s, m,
                           t->cand[R & 1], t->cnt, R, 0, t->wants, t->claim, t->e_stamp, t->cand[(R + 1) & 1],
                           t->stats);
        hipLaunchKernelGGL(k_round_flip, dim3(kBigBlocks), dim3(kThreads), 0, s, m, t->cnt, R, t->wants, t->claim,
                           t->e_stamp, t->cand[(R + 1) & 1], t->stats, stat);
        MCL_HIP(hipGetLastError());
        t->round++;
        done++;
    }
    const int R = t->round;
    const int tail_cap = cap - r;
    if (tail_cap > 0) {
        hipLaunchKernelGGL(k_round_tail<kKind>, dim3(1), dim3(1024), 0, s, m, t->cnt, R, tail_cap, t->cand[0], t->cand[1],
                           t->wants, t->claim, t->e_stamp, t->stats, t->rounds_done);
    }
    MCL_HIP(hipGetLastError());
    int tail_rounds = 0;
    if (tail_cap > 0) MCL_HIP(hipMemcpyAsync(&tail_rounds, t->rounds_done, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_TRY(read_counts(t, s));
    t->round = R + tail_rounds + 1;  // +1 keeps claim keys of the next call above the tail's
    done += tail_rounds;
    if (rounds_out) *rounds_out = done;
    if (t->h_cnt[kCnCand + (R + tail_rounds) % 3] != 0) {
        if (trace_on()) trace_nd_edges(t, s);
        return MCL_ERR_ROUND_CAP;
    }
    return MCL_OK;
}

static void stats_begin(mcl_tri* t, hipStream_t s) {
    t->h_stats_fresh = false;
    hipLaunchKernelGGL(k_set_u64, dim3(4), dim3(256), 0, s, t->stats, (int)(kStCount * kStatShards), 0ull);
}
// the statistics counters summed over their shards
static int stats_read(mcl_tri* t, hipStream_t s, unsigned long long (&h)[kStCount]) {
    const size_t nraw = (size_t)kStCount * kStatShards;
    const unsigned long long* raw = t->h_stats;
    if (!t->h_stats_fresh) {
        MCL_HIP(hipMemcpyAsync(t->h_stats, t->stats, sizeof(unsigned long long) * nraw, hipMemcpyDeviceToHost, s));
        MCL_HIP(hipStreamSynchronize(s));
    }
    t->h_stats_fresh = false;
    for (int k = 0; k < kStCount; k++) {
        unsigned long long a = 0;
        for (int j = 0; j < kStatShards; j++) a += raw[(size_t)k * kStatShards + j];
        h[k] = a;
    }
    return MCL_OK;
}
static int stats_end(mcl_tri* t, hipStream_t s, mcl_stats_t* st, std::chrono::steady_clock::time_point t0) {
    if (!st) return MCL_OK;
    unsigned long long h[kStCount];
    MCL_TRY(stats_read(t, s, h));
    long long rounds = st->rounds;
    memset(st, 0, sizeof(*st));
    st->rounds = rounds;
    st->flips = (long long)h[kStFlips];
    st->untangle_flips = (long long)h[kStUntangleFlips];
    st->reverts = (long long)h[kStReverts];
    st->stalled = (long long)h[kStStalled] + (long long)h[kStRelocFailed];
    st->duplicates = (long long)h[kStRelDup];
    st->relocated = (long long)h[kStRelocated];
    st->relocated_grid = (long long)h[kStRelocGrid];
    st->inverted = (long long)h[kStInverted];
    for (int k = 0; k < 4; k++) {
        st->pred_orient[k] = (long long)h[kStOrient0 + k];
        st->pred_incircle[k] = (long long)h[kStIncircle0 + k];
    }
    st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return MCL_OK;
}

extern "C" int mcl_tri_delonize(mcl_tri_t* t, int variant, int max_rounds, mcl_stats_t* st, void* stream) {
    if (!t) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    MCL_TRY(ensure_mesh(t, s));
    t->stars_valid = false;
    auto t0 = std::chrono::steady_clock::now();
    stats_begin(t, s);
    long long rounds = 0;
    int rc = run_rounds<0>(t, true, variant, max_rounds, s, &rounds);
    if (st) st->rounds = rounds;
    int rc2 = stats_end(t, s, st, t0);
    return rc != MCL_OK ? rc : rc2;
}

extern "C" int mcl_tri_count_inverted(mcl_tri_t* t, long long* out, void* stream) {
    if (!t || !out) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    MCL_TRY(ensure_mesh(t, s));
    hipLaunchKernelGGL(k_prep_round, dim3(1), dim3(64), 0, s, t->cnt, 0, -1);
    hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, t->cnt + kCnMisc, 1, 0);
    hipLaunchKernelGGL((k_check_all<true>), dim3(t->grid_blocks), dim3(kThreads), 0, s, t->m, (int*)nullptr, t->cnt,
                       (int)kCnMisc, t->stats);
    MCL_HIP(hipGetLastError());
    MCL_TRY(read_counts(t, s));
    *out = t->h_cnt[kCnMisc];
    return MCL_OK;
}

extern "C" int mcl_tri_count_non_delaunay(mcl_tri_t* t, long long* out, void* stream) {
    if (!t || !out) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    MCL_TRY(ensure_mesh(t, s));
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, t->scalar, 1, 0ull);
    hipLaunchKernelGGL(k_count_nd, dim3(256), dim3(kThreads), 0, s, t->m, t->scalar);
    MCL_HIP(hipGetLastError());
    unsigned long long h = 0;
    MCL_HIP(hipMemcpyAsync(&h, t->scalar, sizeof(h), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    *out = (long long)h;
    return MCL_OK;
}

extern "C" int mcl_tri_untangle(mcl_tri_t* t, int max_rounds, mcl_stats_t* st, void* stream) {
    if (!t) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    MCL_TRY(ensure_mesh(t, s));
    t->stars_valid = false;
    auto t0 = std::chrono::steady_clock::now();
    stats_begin(t, s);
    long long rounds = 0;
    int rc = run_rounds<1>(t, true, 1, max_rounds, s, &rounds);
    if (st) st->rounds = rounds;
    MCL_TRY(stats_end(t, s, st, t0));
    if (rc != MCL_OK && rc != MCL_ERR_ROUND_CAP) return rc;
    long long inv = 0;
    MCL_TRY(mcl_tri_count_inverted(t, &inv, stream));
    if (st) st->inverted = inv;
    return inv ? MCL_ERR_UNRESOLVED : rc;
}

// ------------------------------------------------------------- kinetic step
// Step state machine (one iteration per pass of the outer loop):
//   move      -- iteration 0 moves every non-corner vertex to its target;
//                later iterations move only pending vertices, by 2^-fexp of
//                their remaining displacement
//   check     -- triangles (all, or around movers) that are not strictly CCW
//   repair    -- rounds of flips of an edge between a bad and a good triangle
//                whose two new triangles are both strictly CCW (the B-step
//                of Host_Triangulation.h:445-450 plus its on-edge case).  A
//                triangle made by a repair is not repaired again in the same
//                iteration, so every repair can later be undone on its own.
//   revert    -- each remaining bad triangle either undoes the repair that
//                made it or sends its largest-displacement mover back to the
//                accepted position; passes repeat until no triangle is bad.
//                This terminates: with all repairs undone and all movers
//                back, the mesh is the iteration's valid starting state.
//   Lawson    -- flip rounds to the exact Delaunay triangulation
//   pending   -- movers short of their target retry next iteration
namespace {
struct StepCtx {
    mcl_tri* t;
    hipStream_t s;
    bool f32;
};

template <bool kF32>
int check_list(mcl_tri* t, hipStream_t s, const int* vlist, int slot_list, int n, int* bad, int slot_bad) {
    int ts = ++t->stamp;
    hipLaunchKernelGGL((k_check_ring<kF32>), dim3(blocks_for(n)), dim3(kThreads), 0, s, t->m, vlist, t->cnt, slot_list,
                       t->t_stamp, ts, bad, t->cnt, slot_bad, t->stats);
    MCL_HIP(hipGetLastError());
    return MCL_OK;
}
}  // namespace

static int zero_slots(mcl_tri* t, hipStream_t s, int first, int n) {
    hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, t->cnt + first, n, 0);
    MCL_HIP(hipGetLastError());
    return MCL_OK;
}

// repair rounds on the bad list in bad[*cur] (count slot kCnBad + *cur);
// returns with the remaining bad triangles in bad[*cur]
static bool trace_on() {
    static const bool on = getenv("MCL_TRACE") != nullptr;
    return on;
}

// MCL_VALIDATE set: check the mesh structure after every phase of a step and
// stop with MCL_ERR_TOPOLOGY (naming the phase) at the first violation
static bool validate_on() { return getenv("MCL_VALIDATE") != nullptr; }
static int validate(mcl_tri* t, hipStream_t s, const char* phase) {
    if (!validate_on()) return MCL_OK;
    hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, t->tail_out + 7, 1, 0);
    hipLaunchKernelGGL(k_validate, dim3(t->grid_blocks), dim3(kThreads), 0, s, t->m, t->tail_out + 7);
    MCL_HIP(hipGetLastError());
    int e = 0;
    MCL_HIP(hipMemcpyAsync(&e, t->tail_out + 7, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    if (e) {
        fprintf(stderr, "mcleap: mesh structure broken after %s (%d violations)\n", phase, e);
        return MCL_ERR_TOPOLOGY;
    }
    if (trace_on()) {
        hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, t->tail_out + 7, 1, 0);
        hipLaunchKernelGGL(k_count_inv, dim3(t->grid_blocks), dim3(kThreads), 0, s, t->m, t->tail_out + 7);
        MCL_HIP(hipGetLastError());
        MCL_HIP(hipMemcpyAsync(&e, t->tail_out + 7, sizeof(int), hipMemcpyDeviceToHost, s));
        MCL_HIP(hipStreamSynchronize(s));
        fprintf(stderr, "  [validate] after %s: %d triangles not strictly CCW\n", phase, e);
    }
    return MCL_OK;
}

static int repair_rounds(mcl_tri* t, hipStream_t s, bool f32, int* cur, int rep_stamp, int* nrep) {
    DMesh& m = t->m;
    *nrep = 0;
    for (int r = 0; r < 64; r++) {
        int nb = t->h_cnt[kCnBad + *cur];
        if (nb == 0) break;
        const int R = t->round, nxt = 1 - *cur;
        hipLaunchKernelGGL(k_prep_round, dim3(1), dim3(64), 0, s, t->cnt, R % 3, 0);
        MCL_TRY(zero_slots(t, s, kCnBad + nxt, 1));
        if (f32)
            hipLaunchKernelGGL((k_filter_bad<true>), dim3(blocks_for(nb)), dim3(kThreads), 0, s, m, t->bad[*cur], t->cnt,
                               (int)kCnBad + *cur, t->bad[nxt], (int)kCnBad + nxt, t->e_stamp, R, t->cand[R & 1], t->stats);
        else
            hipLaunchKernelGGL((k_filter_bad<false>), dim3(blocks_for(nb)), dim3(kThreads), 0, s, m, t->bad[*cur], t->cnt,
                               (int)kCnBad + *cur, t->bad[nxt], (int)kCnBad + nxt, t->e_stamp, R, t->cand[R & 1], t->stats);
        MCL_TRY(read_counts(t, s));
        *cur = nxt;
        int nc = t->h_cnt[kCnCand + R % 3];
        if (t->h_cnt[kCnBad + nxt] == 0 || nc == 0) break;
        if (f32)
            hipLaunchKernelGGL((k_round_test<2, false, true, false>), dim3(blocks_for(nc)), dim3(kThreads), 0, s, m,
                               t->cand[R & 1], t->cnt, R, 0, t->wants, t->claim, t->e_stamp, t->cand[(R + 1) & 1], t->stats,
                               (const int*)t->t_rep, rep_stamp);
        else
            hipLaunchKernelGGL((k_round_test<2, false, false, false>), dim3(blocks_for(nc)), dim3(kThreads), 0, s, m,
                               t->cand[R & 1], t->cnt, R, 0, t->wants, t->claim, t->e_stamp, t->cand[(R + 1) & 1], t->stats,
                               (const int*)t->t_rep, rep_stamp);
        MCL_TRY(read_counts(t, s));
        int nw = t->h_cnt[kCnWants + R % 3];
        t->round++;
        if (nw == 0) break;
        MCL_TRY(zero_slots(t, s, kCnErr, 1));
        hipLaunchKernelGGL(k_repair_flip, dim3(blocks_for(nw)), dim3(kThreads), 0, s, m, t->cnt, R, t->wants, t->claim,
                           t->t_rep, t->t_rep_e, rep_stamp, t->v_touch, t->touched, (int)kCnTouched, t->stats);
        MCL_HIP(hipGetLastError());
        MCL_TRY(read_counts(t, s));
        *nrep += t->h_cnt[kCnErr];
    }
    return MCL_OK;
}

// take back moves / undo repairs until no triangle is bad; starts from the
// bad list in bad[*cur]
static int revert_passes(mcl_tri* t, hipStream_t s, bool f32, int* cur, int mv_stamp, int rep_stamp, int* undos,
                         bool* stuck) {
    DMesh& m = t->m;
    *stuck = false;
    const RevertArgs rargs{t->acc, t->v_mv, mv_stamp, t->t_rep, t->t_rep_e, rep_stamp, t->e_undo, t->undo,
                           t->v_rev, t->rev, t->stuck};
    for (int pass = 0; pass < 4096; pass++) {
        int nb = t->h_cnt[kCnBad + *cur];
        if (nb == 0) break;
        if (f32 && nb < t->revert_tail) {
            // finish in one workgroup: no host round trip per pass
            const int max_passes = 4096 - pass;
            const int stamp0 = t->stamp + 1;
            t->stamp += 2 * max_passes + 2;
            hipLaunchKernelGGL((k_revert_tail<true>), dim3(1), dim3(1024), 0, s, m, t->bad[0], t->bad[1], *cur, t->cnt,
                               rargs, t->t_rep, stamp0, max_passes, t->v_fexp, t->v_soft, mv_stamp, t->allrev,
                               t->t_stamp, t->stats, t->tail_out);
            MCL_HIP(hipGetLastError());
            int out[4];
            MCL_HIP(hipMemcpyAsync(out, t->tail_out, sizeof(out), hipMemcpyDeviceToHost, s));
            MCL_TRY(read_counts(t, s));
            if (trace_on()) fprintf(stderr, "  [revert tail] passes %d undos %d stuck %d\n", out[0], out[3], out[2]);
            *undos += out[3];
            *cur = out[1];
            if (out[2]) *stuck = true;
            return MCL_OK;
        }
        const int rev_stamp = ++t->stamp, undo_stamp = rev_stamp;
        MCL_TRY(zero_slots(t, s, kCnRev, 2));   // rev list, undo list
        MCL_TRY(zero_slots(t, s, kCnMisc, 1));  // stuck
        hipLaunchKernelGGL(k_revert_mark, dim3(blocks_for(nb)), dim3(kThreads), 0, s, m, t->bad[*cur], t->cnt,
                           (int)kCnBad + *cur, rargs, undo_stamp, rev_stamp);
        MCL_TRY(read_counts(t, s));
        if (t->h_cnt[kCnMisc] > 0) { *stuck = true; return MCL_OK; }
        const int nr = t->h_cnt[kCnRev], nu = t->h_cnt[kCnRev + 1];
        if (trace_on()) fprintf(stderr, "  [revert %d] bad %d rev %d undo %d\n", pass, nb, nr, nu);
        const int nxt = 1 - *cur;
        MCL_TRY(zero_slots(t, s, kCnBad + nxt, 1));
        if (nr > 0) {
            t->timers[4].mark(s);
            if (f32)
                hipLaunchKernelGGL((k_revert_apply<true>), dim3(blocks_for(nr)), dim3(kThreads), 0, s, m, t->rev, t->cnt,
                                   (int)kCnRev, t->acc, t->v_fexp, t->v_soft, mv_stamp, t->allrev, t->cnt,
                                   (int)kCnRev + 2, t->stats);
            else
                hipLaunchKernelGGL((k_revert_apply<false>), dim3(blocks_for(nr)), dim3(kThreads), 0, s, m, t->rev, t->cnt,
                                   (int)kCnRev, t->acc, t->v_fexp, t->v_soft, mv_stamp, t->allrev, t->cnt,
                                   (int)kCnRev + 2, t->stats);
            t->timers[4].mark(s);
        }
        if (nu > 0) {
            *undos += nu;
            if (f32)
                hipLaunchKernelGGL((k_undo_apply<true>), dim3(blocks_for(nu)), dim3(kThreads), 0, s, m, t->undo, t->cnt,
                                   (int)kCnRev + 1, t->t_rep, t->bad[nxt], (int)kCnBad + nxt, t->stats);
            else
                hipLaunchKernelGGL((k_undo_apply<false>), dim3(blocks_for(nu)), dim3(kThreads), 0, s, m, t->undo, t->cnt,
                                   (int)kCnRev + 1, t->t_rep, t->bad[nxt], (int)kCnBad + nxt, t->stats);
        }
        if (nr > 0) {
            if (f32) MCL_TRY(check_list<true>(t, s, t->rev, kCnRev, nr, t->bad[nxt], kCnBad + nxt));
            else MCL_TRY(check_list<false>(t, s, t->rev, kCnRev, nr, t->bad[nxt], kCnBad + nxt));
        }
        MCL_HIP(hipGetLastError());
        MCL_TRY(read_counts(t, s));
        *cur = nxt;
        if (nr == 0 && nu == 0) { *stuck = true; return MCL_OK; }
    }
    return MCL_OK;
}

// relocation rounds for the vertices listed in pend[*slot] (count in
// kCnPend); each round re-locates, claims and applies; losers retry
static int relocate_rounds(mcl_tri* t, hipStream_t s, bool f32, int n, int* slot, int* rounds) {
    DMesh& m = t->m;
    *rounds = 0;
    int cur = *slot;
    for (int r = 0; r < 256 && n > 0; r++) {
        if (f32 && n < t->reloc_tail) {
            const int max_rounds = 256 - r;
            const int R0 = t->round;
            hipLaunchKernelGGL((k_reloc_tail<true>), dim3(1), dim3(1024), 0, s, m, t->pend[0], t->pend[1], cur, n, R0,
                               max_rounds, t->r_tt, t->r_tt2, t->r_wait, t->claim, t->tgt, t->stats, t->tail_out);
            MCL_HIP(hipGetLastError());
            int out[3];
            MCL_HIP(hipMemcpyAsync(out, t->tail_out, sizeof(out), hipMemcpyDeviceToHost, s));
            MCL_HIP(hipStreamSynchronize(s));
            if (trace_on()) fprintf(stderr, "  [reloc tail] rounds %d left %d\n", out[0], out[2]);
            t->round = R0 + out[0] + 1;
            *rounds += out[0];
            *slot = out[1];
            return out[2] == 0 ? MCL_OK : MCL_ERR_ROUND_CAP;
        }
        const int R = t->round++;
        const int nxt = 1 - cur;
        hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, t->cnt + kCnPend, 1, n);
        hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, t->cnt + kCnMove, 1, 0);
        if (f32)
            hipLaunchKernelGGL((k_reloc_locate<true>), dim3(blocks_for(n)), dim3(kThreads), 0, s, m, t->pend[cur], t->cnt,
                               (int)kCnPend, t->tgt, t->r_tt, t->r_tt2, R);
        else
            hipLaunchKernelGGL((k_reloc_locate<false>), dim3(blocks_for(n)), dim3(kThreads), 0, s, m, t->pend[cur], t->cnt,
                               (int)kCnPend, t->tgt, t->r_tt, t->r_tt2, R);
        hipLaunchKernelGGL(k_reloc_claim, dim3(blocks_for(n)), dim3(kThreads), 0, s, m, t->pend[cur], t->cnt, (int)kCnPend,
                           R, t->r_tt, t->r_tt2, t->claim);
        t->timers[3].mark(s);
        if (f32)
            hipLaunchKernelGGL((k_reloc_apply<true>), dim3(blocks_for(n)), dim3(kThreads), 0, s, m, t->pend[cur], t->cnt,
                               (int)kCnPend, R, t->r_tt, t->r_tt2, t->r_wait, t->claim, t->tgt, t->pend[nxt], t->cnt,
                               (int)kCnMove, t->stats);
        else
            hipLaunchKernelGGL((k_reloc_apply<false>), dim3(blocks_for(n)), dim3(kThreads), 0, s, m, t->pend[cur], t->cnt,
                               (int)kCnPend, R, t->r_tt, t->r_tt2, t->r_wait, t->claim, t->tgt, t->pend[nxt], t->cnt,
                               (int)kCnMove, t->stats);
        t->timers[3].mark(s);
        MCL_HIP(hipGetLastError());
        MCL_TRY(read_counts(t, s));
        MCL_TRY(validate(t, s, "a relocation round"));
        if (trace_on()) fprintf(stderr, "  [reloc %d] list %d retry %d\n", r, n, t->h_cnt[kCnMove]);
        n = t->h_cnt[kCnMove];
        cur = nxt;
        (*rounds)++;
    }
    *slot = cur;
    return n == 0 ? MCL_OK : MCL_ERR_ROUND_CAP;
}

// The vertices of the relocation list `list` (n, in allrev or r_defer)
// whose relocation met a star above kMaxStar.  Relocation re-inserts
// vertices without Delaunay flips, so a vertex opposite an edge on which many
// points are inserted (points clamped onto a side of the box, say) gains an
// edge per point; after Lawson rounds (lawson) its degree is an ordinary
// Delaunay one again, and it is relocated again (then Lawson rounds again),
// at most kDeferPasses times.  What is still deferred after that, or
// without Lawson rounds, counts as stalled.
constexpr int kDeferPasses = 4;
static int reloc_deferred(mcl_tri* t, hipStream_t s, bool f32, int variant, const int* list, int n, bool lawson,
                          long long* rounds, int* reloc_rounds) {
    const int* in = list;
    for (int pass = 0; n > 0; pass++) {
        int* out = in == t->r_defer ? t->allrev : t->r_defer;
        hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, t->cnt + kCnDefer, 1, 0);
        hipLaunchKernelGGL(k_reloc_deferred, dim3(blocks_for(n)), dim3(kThreads), 0, s, in, n, (const int*)t->r_tt,
                           out, t->cnt + kCnDefer);
        MCL_HIP(hipGetLastError());
        MCL_TRY(read_counts(t, s));
        const int nd = t->h_cnt[kCnDefer];
        if (nd == 0) return MCL_OK;
        if (trace_on()) fprintf(stderr, "  [reloc deferred %d] %d vertices\n", pass, nd);
        if (!lawson || pass == kDeferPasses) {
            hipLaunchKernelGGL(k_stat_add, dim3(1), dim3(64), 0, s, t->stats, (int)kStRelocFailed,
                               (unsigned long long)nd);
            hipLaunchKernelGGL(k_stat_add, dim3(1), dim3(64), 0, s, t->stats, (int)kStRelStar,
                               (unsigned long long)nd);
            MCL_HIP(hipGetLastError());
            return MCL_OK;
        }
        hipLaunchKernelGGL(k_reloc_init, dim3(blocks_for(nd)), dim3(kThreads), 0, s, (const int*)out, nd, t->pend[0],
                           t->r_wait);
        MCL_HIP(hipGetLastError());
        int slot = 0, rr = 0;
        MCL_TRY(relocate_rounds(t, s, f32, nd, &slot, &rr));
        MCL_TRY(validate(t, s, "deferred relocation rounds"));
        *reloc_rounds += rr;
        long long lr = 0;
        MCL_TRY(run_rounds<0>(t, true, variant, 0, s, &lr));
        MCL_TRY(validate(t, s, "Lawson rounds after deferred relocations"));
        *rounds += lr;
        in = out;
        n = nd;
    }
    return MCL_OK;
}

// One kinetic step:
//   move      -- every non-corner vertex goes to its target
//   check     -- triangles that are not strictly CCW
//   repair    -- rounds of flips of an edge between a bad and a good triangle
//                whose two new triangles are both strictly CCW (the B-step
//                of Host_Triangulation.h:445-450 plus its on-edge case).  A
//                triangle made by a repair is not repaired again in the same
//                step, so every repair can be undone on its own.
//   revert    -- each remaining bad triangle either undoes the repair that
//                made it or sends its largest-displacement vertex back to its
//                previous position, until no triangle is bad (this
//                terminates: with every repair undone and every vertex back,
//                the mesh is the previous, valid, state)
//   relocate  -- vertices sent back are removed from their stars and
//                re-inserted at their targets
//   Lawson    -- flip rounds to the exact Delaunay triangulation
static int fork_side(mcl_tri* t, hipStream_t s, int nxt);

static int step_impl(mcl_tri* t, const mcl_step_params_t* p, const float2* disp, mcl_stats_t* st, void* stream) {
    t->stars_valid = false;  // flips and relocations change the topology
    hipStream_t s = (hipStream_t)stream;
    DMesh& m = t->m;
    auto t0 = std::chrono::steady_clock::now();
    stats_begin(t, s);
    const int variant = p->variant;
    const int gb = t->grid_blocks;
    const bool f32 = variant == 0;
    static const bool trace = getenv("MCL_TRACE") != nullptr;

    const int mv_stamp = ++t->stamp;
    t->timers[2].mark(s);
    hipLaunchKernelGGL(k_move_motion, dim3(blocks_for(m.nv)), dim3(kThreads), 0, s, m, t->acc, t->tgt, t->v_mv,
                       t->v_fexp, mv_stamp, *p, (const mcl_step_params_t*)nullptr, disp, StepPrologue{});
    t->timers[2].mark(s);
    // the timestep's FRNN part beside the whole step, on the targets (where
    // every vertex ends unless one stalls: checked after the step)
    MCL_TRY(fork_side(t, s, t->star_cur));
    if (t->side_pending) t->frnn_on_targets = true;
    MCL_TRY(zero_slots(t, s, kCnBad, 11));  // bad, rev, move, pend, err, misc, touched
    t->timers[1].mark(s);
    if (f32)
        hipLaunchKernelGGL((k_check_all<true>), dim3(gb), dim3(kThreads), 0, s, m, t->bad[0], t->cnt, (int)kCnBad, t->stats);
    else
        hipLaunchKernelGGL((k_check_all<false>), dim3(gb), dim3(kThreads), 0, s, m, t->bad[0], t->cnt, (int)kCnBad, t->stats);
    t->timers[1].mark(s);
    MCL_HIP(hipGetLastError());
    MCL_TRY(read_counts(t, s));
    const long long inverted0 = t->h_cnt[kCnBad];
    int cur = 0, nrep = 0, nundo = 0, reloc_rounds = 0;
    bool stuck = false;
    MCL_TRY(validate(t, s, "move"));
    MCL_TRY(repair_rounds(t, s, f32, &cur, mv_stamp, &nrep));
    MCL_TRY(validate(t, s, "repair rounds"));
    const int bad_after = t->h_cnt[kCnBad + cur];
    MCL_TRY(revert_passes(t, s, f32, &cur, mv_stamp, mv_stamp, &nundo, &stuck));
    MCL_TRY(validate(t, s, "take-back passes"));
    if (stuck) return MCL_ERR_TOPOLOGY;  // cannot happen: see k_revert_mark
    const int nrev = t->h_cnt[kCnRev + 2];
    int rc = MCL_OK;
    if (nrev > 0) {
        // relocate the vertices that were sent back (list allrev -> pend[0])
        hipLaunchKernelGGL(k_reloc_init, dim3(blocks_for(nrev)), dim3(kThreads), 0, s, (const int*)t->allrev, nrev,
                           t->pend[0], t->r_wait);
        MCL_HIP(hipGetLastError());
        int slot = 0;
        rc = relocate_rounds(t, s, f32, nrev, &slot, &reloc_rounds);
        if (rc == MCL_OK) rc = validate(t, s, "relocation rounds");
    }
    long long rounds = 0;
    if (rc == MCL_OK) rc = run_rounds<0>(t, true, variant, 0, s, &rounds);
    if (rc == MCL_OK) rc = validate(t, s, "Lawson rounds");
    if (rc == MCL_OK && nrev > 0) rc = reloc_deferred(t, s, f32, variant, t->allrev, nrev, true, &rounds, &reloc_rounds);
    if (trace) {
        unsigned long long h[kStCount];
        if (stats_read(t, s, h) == MCL_OK)
            fprintf(stderr, "[step] relocation failures: walk %llu star %llu removal %llu put-back %llu; duplicates %llu\n",
                    h[kStRelWalk], h[kStRelStar], h[kStRelStuck], h[kStRelPutBack], h[kStRelDup]);
    }
    if (trace)
        fprintf(stderr, "[step] inverted %lld repairs %d bad_after_repair %d undone %d reverted %d reloc_rounds %d lawson %lld\n",
                inverted0, nrep, bad_after, nundo, nrev, reloc_rounds, rounds);
    if (st) st->rounds = rounds;
    int rc2 = stats_end(t, s, st, t0);
    if (st) {
        st->iterations = 1 + reloc_rounds;
        st->inverted = inverted0;
    }
    return rc != MCL_OK ? rc : rc2;
}

// stars of the current mesh into star_*[star_cur] (after any topology change
// made by the flip-based paths)
static int ensure_stars(mcl_tri* t, hipStream_t s) {
    if (t->stars_valid) return MCL_OK;
    hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, t->s_aux, kStarAux, 0);
    hipLaunchKernelGGL(k_stars_from_mesh, dim3(blocks_for(t->m.nv)), dim3(256), 0, s, t->m, t->star_deg[t->star_cur],
                       t->star_nbr[t->star_cur], t->s_aux);
    MCL_HIP(hipGetLastError());
    int fail = 0;
    MCL_HIP(hipMemcpyAsync(&fail, t->s_aux, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    if (fail) return MCL_ERR_CAPACITY;  // a degree above kStarCap (kCornerCap at a corner)
    t->stars_valid = true;
    return MCL_OK;
}

// Rebuilds the half-edge arrays from the stars after star steps.  Consistent
// stars with the degree sum 2E give E edges and, each triangle being listed
// by its three vertices, T = 2E/3 triangles; the counts are checked again
// here before any write.
static int ensure_mesh(mcl_tri* t, hipStream_t s) {
    if (!t->mesh_stale) return MCL_OK;
    DMesh& m = t->m;
    const int c = t->star_cur;
    hipLaunchKernelGGL(k_star_counts, dim3(blocks_for(m.nv)), dim3(256), 0, s, m.nv, m.ncorner, (const int*)t->star_deg[c],
                       (const int*)t->star_nbr[c], t->s_ecnt, t->s_tcnt);
    MCL_TRY(scan_exclusive(t->s_ecnt, t->s_ecnt, m.nv, t->scan_tmp, s));
    MCL_TRY(scan_exclusive(t->s_tcnt, t->s_tcnt, m.nv, t->scan_tmp, s));
    int tot[2];
    MCL_HIP(hipMemcpyAsync(&tot[0], t->s_ecnt + m.nv, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipMemcpyAsync(&tot[1], t->s_tcnt + m.nv, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    if (tot[0] != m.ne || tot[1] != m.nt) {
        fprintf(stderr, "mcleap: stars give %d edges and %d triangles, the mesh has %d and %d\n", tot[0], tot[1], m.ne,
                m.nt);
        return MCL_ERR_TOPOLOGY;
    }
    // hid rows in star_tnbr (scratch between star steps)
    hipLaunchKernelGGL(k_star_hid, dim3(blocks_for(m.nv)), dim3(256), 0, s, m, (const int*)t->star_deg[c],
                       (const int*)t->star_nbr[c], (const int*)t->s_ecnt, t->star_tnbr);
    hipLaunchKernelGGL(k_star_link, dim3(blocks_for(m.nv)), dim3(256), 0, s, m, (const int*)t->star_deg[c],
                       (const int*)t->star_nbr[c], (const int*)t->star_tnbr, (const int*)t->s_tcnt);
    hipLaunchKernelGGL(k_star_face3, dim3(blocks_for(m.nt)), dim3(256), 0, s, m);
    MCL_HIP(hipGetLastError());
    t->mesh_stale = false;
    return validate(t, s, "star rebuild");
}

// Kinetic step on the torus by star recomputation (see mcl_star.hpp).
// Returns MCL_OK with *done = false when the stars could not be settled; the
// positions are then back at the start of the step and the mesh untouched.
//
// The step's first star passes (MCL_OPT_STAR_HEAD_PASSES of them, default 4:
// the move, the first pass over
// every vertex with its consistency check and give-up test, and retries of the
// listed vertices) are enqueued without any host read: every list length
// stays on the device, kernels over an empty list return at once, and the
// check stamps come from a device counter.  With MCL_OPT_GRAPH this head is
// captured once per star buffer parity and replayed as one hipGraph launch
// (its parameters copied to the device first).  The host then reads the
// counters once; the rare step that has not settled continues eagerly, one
// read per pass.

static int launch_tail(mcl_tri* t, const TailReq& q, hipStream_t s, int part, int c, bool from_stars,
                       const float2* pos = nullptr);

namespace {
// launch geometry of the star kernels
struct StarGrid {
    mcl_tri* t;
    // blocks for n vertices, kv vertices per block iteration (a multiple of 8
    // when large enough, for the XCD-aware split)
    dim3 full(long long nverts, int kv) const {
        long long b = std::max<long long>(1, std::min<long long>((nverts + kv - 1) / kv, 4LL * t->grid_blocks));
        if (b >= 64) b &= ~7LL;
        return dim3((unsigned)b);
    }
    // list lengths stay on the device: grids sized for a long list
    dim3 list() const { return dim3((unsigned)(std::min<long long>(4LL * t->grid_blocks, 2048) & ~7LL)); }
    // the same numbers of threads in blocks of kStarBlock (k_star)
    dim3 star_full(long long nverts, int kv) const {
        const long long scale = 256 / kStarBlock;
        long long b = std::max<long long>(1, std::min<long long>((nverts + kv - 1) / kv, 4LL * scale * t->grid_blocks));
        if (b >= 64) b &= ~7LL;
        return dim3((unsigned)b);
    }
    // (twice the threads of list(): the retry's wave-per-vertex kernel, 162 ->
    // 137 us over ~81k vertices at 1M; half the threads measured 172 us)
    dim3 star_list() const { return dim3(list().x * (256 / kStarBlock) * 2); }
};

// stars of the listed vertices (all without a list; count *nlist_dev) into
// (odeg, onbr): the filter-only kernel, two vertices per wave, then the full
// kernel on the vertices it could not finish (more than 64 candidates, or a
// test the filter could not decide)
// Counters of the star passes (s_aux, and the heavy-list count s_heavy[0])
// are zeroed by a kernel that runs before their producer and after their
// last reader, not by launches of their own: the step's prologue (in the
// move) zeroes all of them, each check zeroes s_heavy for the next pass's
// stars, each retry's filter kernel zeroes the re-check count [8] for its
// scatter, and each scatter zeroes the list count its check fills.  The list
// count alternates between [1] (the first check, and the checks of even
// retries) and [20] (odd retries), so that a pass's list is not zeroed while
// that pass still reads its count.
__host__ __device__ constexpr int star_list_slot(int r) { return (r & 1) ? 20 : 1; }

// stars of the listed vertices (all without a list; count *nlist_dev) into
// (odeg, onbr).  First pass (no list): the 2-ring filter kernel, two vertices
// per wave and at most 32 candidates each; the vertices it leaves (more
// candidates, or a hull it could not certify) go through the same kernel with
// a whole wave per vertex and up to 64 candidates (its own list, whose count
// the first launch zeroes), and what that leaves to the full kernel.  Retries:
// the filter kernel over old and new first rings with their new stars, a wave
// per vertex, then the full kernel on what it leaves (more than 64
// candidates, a failed own star, a hull it could not certify).  reset,
// give_up: as k_star's.
void star_pass_stars(mcl_tri* t, hipStream_t s, int cur, const int* vlist, const int* nlist_dev, int depth,
                     const int* bdeg, const int* bnbr, int* odeg, int* onbr, int* reset, int give_up) {
    DMesh& m = t->m;
    const StarGrid G{t};
    const int* heavy_rows = t->s_heavy;
    if (!vlist) {
        hipLaunchKernelGGL((k_star<true, true, 32, kCandRing2>), G.star_full(m.nv, kStarBlock / 32), dim3(kStarBlock), 0, s, m, vlist, 0,
                           nlist_dev, (const int*)nullptr, (const int*)nullptr, depth, (const int*)t->star_deg[cur],
                           (const int*)t->star_nbr[cur], bdeg, bnbr, odeg, onbr, t->s_aux, t->s_heavy,
                           (const float2*)t->acc, (const float2*)t->tgt, t->stats, t->s_heavy2, give_up,
                           t->s_aux + 19, (const int*)nullptr);
        // (a wave for each of its ~2.4k vertices at 1M: 2048 blocks, 24 -> 16 us against 512)
        hipLaunchKernelGGL((k_star<true, true, 64, kCandRing2>), dim3(2048), dim3(kStarBlock), 0, s, m, vlist, 0, nlist_dev,
                           (const int*)(t->s_heavy + 1), (const int*)t->s_heavy, depth,
                           (const int*)t->star_deg[cur], (const int*)t->star_nbr[cur], bdeg, bnbr, odeg, onbr,
                           t->s_aux, t->s_heavy2, (const float2*)t->acc, (const float2*)t->tgt, t->stats,
                           (int*)nullptr, -1, (int*)nullptr, (const int*)nullptr);
        heavy_rows = t->s_heavy2;
    } else {
        hipLaunchKernelGGL((k_star<true, true, 64, kCandRaw>), G.star_list(), dim3(kStarBlock), 0, s, m, vlist, 0, nlist_dev,
                           (const int*)nullptr, (const int*)nullptr, depth, (const int*)t->star_deg[cur],
                           (const int*)t->star_nbr[cur], bdeg, bnbr, odeg, onbr, t->s_aux, t->s_heavy,
                           (const float2*)t->acc, (const float2*)t->tgt, t->stats, reset, give_up, t->s_aux + 19,
                           (const int*)t->s_hint);
    }
    hipLaunchKernelGGL((k_star<true, false, 64, kCandHash>), dim3(512), dim3(kStarBlock), 0, s, m, vlist, 0, nlist_dev,
                       heavy_rows + 1, heavy_rows, depth, (const int*)t->star_deg[cur],
                       (const int*)t->star_nbr[cur], bdeg, bnbr, odeg, onbr, t->s_aux, (int*)nullptr,
                       (const float2*)t->acc, (const float2*)t->tgt, t->stats, (int*)nullptr, -1, (int*)nullptr,
                       vlist ? (const int*)t->s_hint : (const int*)nullptr);
}

// s_aux: [0] failed stars, [1] / [20] listed vertices (star_list_slot),
// [2..3] degree sum of the valid stars, [4..7] failures by reason, [8]
// vertices to re-check, [10] vertices handed to the full kernel, [12..13]
// vertices retried, [14] retry passes with a non-empty list, [16..18] corner
// fans, [19] gave up, [21] filter-kernel vertices whose hull elimination
// failed its certification (they took the successor scan), [22] / [23]
// vertices the wave-per-vertex filter kernels handed on (more candidates
// than lanes / an uncertified hull)
//
// The first check (r = 0) covers every vertex and lists the inconsistent
// ones; later ones (retry r) cover the re-check list, and also list a
// neighbour whose star lacks the vertex, with the vertex as its candidate
// (only there: at the first check most such pairs settle without it, and
// listing them all costs a third more recomputations).  mark_nbrs: list the
// star neighbours of the inconsistent vertices too (otherwise each conflict
// lists only the side that loses its incircle test).  stamp_off: this
// check's offset from the step's stamp base.
void star_check(mcl_tri* t, hipStream_t s, int nxt, int r, bool mark_nbrs, int stamp_off) {
    DMesh& m = t->m;
    const StarGrid G{t};
    const bool first = r == 0;
    // the first check (every vertex): one thread per vertex, 8 tiles of 256
    // vertices per block (1, 4 and 16 measured slower: 259, 126 and 163 us
    // against 115 at 1M); the later ones (lists): 8 lanes per vertex
    if (first)
        hipLaunchKernelGGL(k_star_check_all, G.full(m.nv, 2048), dim3(256), 0, s, m.nv, (const int*)nullptr,
                           (const int*)nullptr, (const int*)t->star_deg[nxt], (const int*)t->star_nbr[nxt],
                           (const float2*)m.pos, m.periodic, m.ncorner, t->s_mark,
                           (const int*)t->d_stamp, stamp_off, t->s_list, t->s_aux + star_list_slot(r),
                           (long long*)(t->s_aux + 2), t->s_heavy, (int*)nullptr);
    else
        hipLaunchKernelGGL(k_star_check, G.list(), dim3(256), 0, s, m.nv, (const int*)t->s_rlist,
                           (const int*)(t->s_aux + 8), (const int*)t->star_deg[nxt], (const int*)t->star_nbr[nxt],
                           (const float2*)m.pos, m.periodic, m.ncorner, mark_nbrs ? 1 : 0, t->s_mark,
                           (const int*)t->d_stamp, stamp_off, t->s_list, t->s_aux + star_list_slot(r),
                           (long long*)(t->s_aux + 2), t->s_heavy, t->s_hint);
}

// the most vertices the first check may list before the step gives up on the
// stars (MCL_OPT_STAR_GIVE_UP percent of them)
static int star_give_up_limit(const mcl_tri* t) {
    return (int)std::min<long long>((long long)t->star_give_up * t->m.nv / 100, t->m.nv);
}

// retry pass r (1 <= r < kStarPasses): recompute the listed vertices from the
// 2-rings in the previous mesh and in the stars found so far (a vertex whose
// star failed, having left its 2-ring, searches one ring deeper each pass),
// copy them in and re-check.  The first retry applies the give-up limit to
// the first check's list.
void star_retry(mcl_tri* t, hipStream_t s, int cur, int nxt, int r) {
    const StarGrid G{t};
    int* in_count = t->s_aux + star_list_slot(r - 1);
    star_pass_stars(t, s, cur, t->s_list, in_count, 2 + r, t->star_deg[nxt], t->star_nbr[nxt], t->star_tdeg,
                    t->star_tnbr, t->s_aux + 8, r == 1 ? star_give_up_limit(t) : -1);
    hipLaunchKernelGGL(k_star_scatter, G.list(), dim3(256), 0, s, (const int*)t->s_list, (const int*)in_count,
                       (const int*)t->star_tdeg, (const int*)t->star_tnbr, t->star_deg[nxt], t->star_nbr[nxt],
                       t->s_mark, (const int*)t->d_stamp, 2 * r, t->s_rlist, t->s_aux + 8,
                       (long long*)(t->s_aux + 2), (long long*)(t->s_aux + 12), t->s_aux + star_list_slot(r));
    star_check(t, s, nxt, r, r >= kStarDecidePasses, 2 * r + 1);
}

// corner fans from this step's stars (box meshes), into aux[16..18]
void star_corner_fans(mcl_tri* t, hipStream_t s, int cur, int nxt) {
    DMesh& m = t->m;
    if (m.ncorner == 0) return;
    hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, t->s_aux + 16, 3, 0);
    hipLaunchKernelGGL(k_set_ints, dim3(blocks_for(m.ncorner)), dim3(256), 0, s, t->c_count, m.ncorner, 0);
    hipLaunchKernelGGL(k_corner_collect, dim3(blocks_for(m.nv)), dim3(256), 0, s, m.nv, m.ncorner,
                       (const int*)t->star_deg[nxt], (const int*)t->star_nbr[nxt], t->c_bucket, t->c_count);
    hipLaunchKernelGGL(k_corner_fans, dim3((m.ncorner + 63) / 64), dim3(64), 0, s, m.nv, m.ncorner,
                       (const int*)t->c_bucket, (const int*)t->c_count, (const int*)t->star_deg[cur],
                       (const int*)t->star_nbr[cur], t->star_deg[nxt], t->star_nbr[nxt], t->s_aux + 18,
                       (long long*)(t->s_aux + 16));
}

// the head of a star step (see above); pdev: the step parameters in device
// memory (graph), else p by value.  Ends with the corner fans, so that the
// counters read after it describe a complete set of stars.
int star_head(mcl_tri* t, hipStream_t s, int cur, int nxt, const mcl_step_params_t* p,
              const mcl_step_params_t* pdev, const float2* disp, int mv_stamp) {
    DMesh& m = t->m;
    // the statistics, the star counters and the heavy-list count start at
    // zero, and the check stamps move on (the move's block 0)
    const StepPrologue pr{t->stats, (int)(kStCount * kStatShards), t->s_aux, kStarAux, t->s_heavy, t->d_stamp,
                          kStampSpan, t->s_hint};
    t->timers[2].mark(s);
    hipLaunchKernelGGL(k_move_motion, dim3(blocks_for(m.nv)), dim3(kThreads), 0, s, m, t->acc, t->tgt, t->v_mv,
                       t->v_fexp, mv_stamp, *p, pdev, disp, pr);
    t->timers[2].mark(s);
    t->timers[5].mark(s);
    star_pass_stars(t, s, cur, nullptr, nullptr, 2, nullptr, nullptr, t->star_deg[nxt], t->star_nbr[nxt], nullptr, -1);
    t->timers[5].mark(s);
    star_check(t, s, nxt, 0, false, 1);
    MCL_TRY(fork_side(t, s, nxt));
    for (int r = 1; r < t->star_head_passes; r++) star_retry(t, s, cur, nxt, r);
    star_corner_fans(t, s, cur, nxt);
    return MCL_OK;
}
}  // namespace

// The FRNN part of the calling timestep's tail (grid sort and CSR query) on a
// second stream, on the targets the move kernel wrote (where every vertex
// ends the step unless it stalls or is left on a duplicate target): from the
// first check of a star step on, or beside the whole of a flip-based step.
// It reads the targets only, and runs beside launches over short lists and
// host round trips that leave most of the GPU idle, instead of after them.
// The caller's stream waits for it before the timestep returns (join_side);
// it runs again after the step when a vertex is not at its target.  Eager
// steps only (a captured head keeps the tail on the caller's stream), and
// meshes of MCL_OPT_STAR_MIN_VERTICES or more (below, where launches and
// host round trips dominate, the query takes less time than the fork, the
// join and the check of the targets).
static int fork_side(mcl_tri* t, hipStream_t s, int nxt) {
    if (t->capturing || t->graph || !t->tail_req || !t->tail_req->has_frnn() || t->m.nv < t->star_min_vertices)
        return MCL_OK;
    if (!t->side_stream) MCL_HIP(hipStreamCreateWithFlags(&t->side_stream, hipStreamNonBlocking));
    MCL_HIP(hipEventRecord(t->ev_fork, s));
    MCL_HIP(hipStreamWaitEvent(t->side_stream, t->ev_fork, 0));
    t->side_pending = true;
    MCL_TRY(launch_tail(t, *t->tail_req, t->side_stream, 0, nxt, true, t->tgt));
    MCL_HIP(hipEventRecord(t->ev_join, t->side_stream));
    t->frnn_done = true;
    return MCL_OK;
}

// the caller's stream waits for the side stream's work (if any is pending)
static int join_side(mcl_tri* t, hipStream_t s) {
    if (!t->side_pending) return MCL_OK;
    t->side_pending = false;
    MCL_HIP(hipStreamWaitEvent(s, t->ev_join, 0));
    return MCL_OK;
}

static void drop_graphs(mcl_tri* t) {
    for (int k = 0; k < 2; k++) {
        if (t->g_head[k]) (void)hipGraphExecDestroy(t->g_head[k]);
        t->g_head[k] = nullptr;
    }
    for (int k = 0; k < 3; k++) {
        if (t->g_post[k]) (void)hipGraphExecDestroy(t->g_post[k]);
        t->g_post[k] = nullptr;
    }
}

// captures the work `enqueue(cs)` puts on the capture stream cs as a graph
// and instantiates it (the graph is then launched on the caller's stream)
template <class F>
static int capture_graph(mcl_tri* t, hipGraphExec_t* out, F&& enqueue) {
    if (!t->cap_stream) MCL_HIP(hipStreamCreateWithFlags(&t->cap_stream, hipStreamNonBlocking));
    hipStream_t s = t->cap_stream;
    hipGraph_t g = nullptr;
    MCL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    t->capturing = true;
    t->captured = true;
    enqueue(s);
    t->capturing = false;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    }
    const hipError_t e2 = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) return hip_fail(e2, "hipGraphInstantiate", __FILE__, __LINE__);
    return MCL_OK;
}

static int enqueue_ring_csr(mcl_tri* t, hipStream_t s, int c, bool from_stars, int* off, int* nbrs, int cap);

// Enqueues a part of a timestep's tail on s: the grid's counting sort of the
// positions with its CSR FRNN query (part 0), or the CSR one-ring (part 1),
// from the stars in buffer c (from_stars) or from the half-edges.  With
// MCL_OPT_GRAPH (the one-ring only from stars): one launch of a graph
// captured per part (and star buffer), captured again when the caller's
// buffers change.  Only device work, no host read: running a part again over
// the same positions and stars gives the same arrays.
static int launch_tail(mcl_tri* t, const TailReq& q, hipStream_t s, int part, int c, bool from_stars,
                       const float2* pos) {
    if (!pos) pos = t->m.pos;
    auto enqueue = [&](hipStream_t cs) -> int {
        if (part == 0 && q.g) MCL_TRY(grid_enqueue_frnn(q.g, pos, q.rad, q.frnn_off, q.frnn_nbrs, q.frnn_cap, cs));
        if (part == 0 && q.bg) {
            MCL_TRY(mcl_bgrid_build(q.bg, pos, cs));
            MCL_TRY(mcl_bgrid_frnn_csr(q.bg, q.frnn_off, q.frnn_nbrs, q.frnn_cap, nullptr, cs));
        }
        if (part == 1 && q.ring_off) MCL_TRY(enqueue_ring_csr(t, cs, c, from_stars, q.ring_off, q.ring_nbrs, q.ring_cap));
        return MCL_OK;
    };
    if ((part == 0 && !q.has_frnn()) || (part == 1 && !q.ring_off)) return MCL_OK;
    if (!t->graph || (part == 1 && !from_stars)) return enqueue(s);
    float radv = q.rad;
    const void* key[8] = {q.ring_off, q.ring_nbrs, (const void*)(intptr_t)q.ring_cap,
                          q.g ? (const void*)q.g : (const void*)q.bg,
                          (const void*)(intptr_t)*(const int*)&radv, q.frnn_off, q.frnn_nbrs,
                          (const void*)(intptr_t)q.frnn_cap};
    if (memcmp(key, t->post_key, sizeof(key)) != 0) {
        for (int k = 0; k < 3; k++) {
            if (t->g_post[k]) (void)hipGraphExecDestroy(t->g_post[k]);
            t->g_post[k] = nullptr;
        }
        memcpy(t->post_key, key, sizeof(key));
    }
    hipGraphExec_t* ge = &t->g_post[part == 0 ? 2 : c];
    if (!*ge) {
        int trc = MCL_OK;
        MCL_TRY(capture_graph(t, ge, [&](hipStream_t cs) { trc = enqueue(cs); }));
        MCL_TRY(trc);
    }
    MCL_HIP(hipGraphLaunch(*ge, s));
    return MCL_OK;
}

static int star_step(mcl_tri* t, const mcl_step_params_t* p, const float2* disp, mcl_stats_t* st, hipStream_t s,
                     bool* done) {
    DMesh& m = t->m;
    *done = false;
    if (ensure_stars(t, s) != MCL_OK) return MCL_OK;
    const int cur = t->star_cur, nxt = 1 - cur;
    const bool trace = trace_on();
    if (t->graph && !disp && !trace) {
        if (!t->g_head[cur]) {
            // v_mv marks matter only to the flip-based step, which writes its own
            const int mv_stamp = ++t->stamp;
            MCL_TRY(capture_graph(t, &t->g_head[cur], [&](hipStream_t cs) {
                (void)star_head(t, cs, cur, nxt, p, t->d_params, nullptr, mv_stamp);
            }));
        }
        *t->h_params = *p;
        MCL_HIP(hipMemcpyAsync(t->d_params, t->h_params, sizeof(mcl_step_params_t), hipMemcpyHostToDevice, s));
        MCL_HIP(hipGraphLaunch(t->g_head[cur], s));
    } else {
        MCL_TRY(star_head(t, s, cur, nxt, p, nullptr, disp, ++t->stamp));
    }
    MCL_HIP(hipGetLastError());
    int aux[kStarAux];
    // the counters, and the statistics with them (one wait).  At the first
    // read the FRNN part of the calling timestep's tail is enqueued behind
    // the copies before the host waits for them, so that the GPU runs it
    // while the host takes its decision.  It reads the positions only, which
    // later passes change only by sending a vertex found on top of another
    // back to its previous position (counted in the stalled and duplicate
    // statistics); if one does, or the step falls back, it runs again after
    // the step (frnn_done false).
    unsigned long long moved_back = 0;
    auto read_aux = [&](int pass, bool first) -> int {
        // the statistics shards and, right after them, the star counters: one copy
        MCL_HIP(hipMemcpyAsync(t->h_stats, t->stats, sizeof(unsigned long long) * ((size_t)kStCount * kStatShards + kStatsTail),
                               hipMemcpyDeviceToHost, s));
        MCL_HIP(hipEventRecord(t->ev_aux, s));
        if (first && t->tail_req && t->tail_req->has_frnn() && !t->frnn_done) {
            MCL_TRY(launch_tail(t, *t->tail_req, s, 0, nxt, true));
            t->frnn_done = true;
        }
        // with the FRNN part on the side stream, the one-ring part on the
        // stars as they stand (final when this read finds the step settled,
        // else run again after the step), so that the GPU has work while the
        // host waits
        if (first && t->tail_req && t->tail_req->ring_off && t->side_pending) {
            MCL_TRY(launch_tail(t, *t->tail_req, s, 1, nxt, true));
            t->ring_done = true;
        }
        MCL_HIP(hipEventSynchronize(t->ev_aux));
        memcpy(aux, t->h_stats + (size_t)kStCount * kStatShards, sizeof(aux));
        t->h_stats_fresh = true;
        unsigned long long mb = 0;
        for (int j = 0; j < kStatShards; j++)
            mb += t->h_stats[(size_t)kStStalled * kStatShards + j] + t->h_stats[(size_t)kStRelDup * kStatShards + j];
        // the FRNN part holds the final positions unless a vertex was sent
        // back to its previous position after it started: after the first
        // read (caller's stream), or at any point of the step (side stream,
        // which starts during the head)
        if (first) moved_back = t->side_pending ? 0 : mb;
        if (mb != moved_back) t->frnn_done = false;
        if (trace)
            fprintf(stderr,
                    "  [stars pass %d] inconsistent %d (failed stars so far %d: candidates %d undecided %d open %d "
                    "outside %d; full-kernel vertices %d; filter-kernel vertices that took the successor scan %d; "
                    "wave-per-vertex filter kernels: more candidates than lanes %d, uncertified hulls %d)\n",
                    pass, aux[star_list_slot(pass)], aux[0], aux[4], aux[5], aux[6], aux[7], aux[10], aux[21], aux[22],
                    aux[23]);
        return MCL_OK;
    };
    bool settled = false;
    int listed = 0;
    for (int r = t->star_head_passes - 1; r < kStarPasses; r++) {
        if (r >= t->star_head_passes) {
            t->ring_done = false;  // the stars change
            star_retry(t, s, cur, nxt, r);
            star_corner_fans(t, s, cur, nxt);
            MCL_HIP(hipGetLastError());
        }
        MCL_TRY(read_aux(r, r == t->star_head_passes - 1));
        listed = aux[star_list_slot(r)];
        // the give-up limit (applied on the device by the first retry)
        if (r == 0 && listed > star_give_up_limit(t)) break;
        if (listed == 0) {
            long long sum_deg, sum_corner = 0;
            memcpy(&sum_deg, aux + 2, sizeof(sum_deg));
            if (m.ncorner) memcpy(&sum_corner, aux + 16, sizeof(sum_corner));
            // each half-edge once: the degree sum is 2E; every corner fan chained
            settled = sum_deg + sum_corner == (long long)m.nhe && aux[18] == 0 && aux[19] == 0;
            if (trace && !settled)
                fprintf(stderr, "  [stars] degree sum %lld + corners %lld vs %d half-edges, corner failures %d\n",
                        sum_deg, sum_corner, m.nhe, aux[18]);
            break;
        }
    }
    if (!settled) t->frnn_done = t->ring_done = false;
    if (!settled && trace && listed > 0) {
        // the first vertices still listed: id, position, star degree (-1: failed)
        const int k = std::min(listed, 12);
        int ids[12];
        MCL_HIP(hipMemcpyAsync(ids, t->s_list, sizeof(int) * k, hipMemcpyDeviceToHost, s));
        MCL_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < k; i++) {
            float2 q;
            int d = 0;
            if (ids[i] < 0 || ids[i] >= m.nv) continue;
            MCL_HIP(hipMemcpyAsync(&q, m.pos + ids[i], sizeof(q), hipMemcpyDeviceToHost, s));
            MCL_HIP(hipMemcpyAsync(&d, t->star_deg[nxt] + ids[i], sizeof(d), hipMemcpyDeviceToHost, s));
            MCL_HIP(hipStreamSynchronize(s));
            fprintf(stderr, "  [stars] listed %d at (%.9g, %.9g) degree %d\n", ids[i], q.x, q.y, d);
        }
    }
    if (!settled) {
        // back to the start of the step; the mesh and its stars are untouched
        MCL_HIP(hipMemcpyAsync(m.pos, t->acc, sizeof(float2) * m.nv, hipMemcpyDeviceToDevice, s));
        MCL_HIP(hipStreamSynchronize(s));
        if (trace_on()) fprintf(stderr, "  [stars] not settled: flip-based step instead\n");
        return MCL_OK;
    }
    // the stars are the triangulation; the half-edge arrays are rebuilt from
    // them when an entry point reads them (ensure_mesh)
    t->star_cur = nxt;
    t->stars_valid = true;
    t->mesh_stale = true;
    t->star_end_event = true;  // nothing of the step's own work follows its last counter read
    if (validate_on()) MCL_TRY(ensure_mesh(t, s));
    if (st) {
        long long retried;
        memcpy(&retried, aux + 12, sizeof(retried));
        st->iterations = 1 + aux[14];
        st->star_retries = retried;
    }
    *done = true;
    return MCL_OK;
}

// The star step; false in *done when it did not settle (the flip-based step
// then ran instead, st->fallbacks = 1, its host time in *flip_ms).
static int star_or_fallback(mcl_tri* t, const mcl_step_params_t* p, const float2* disp, mcl_stats_t* st, void* stream,
                            bool* done, double* flip_ms = nullptr) {
    hipStream_t s = (hipStream_t)stream;
    auto t0 = std::chrono::steady_clock::now();
    t->h_stats_fresh = false;  // (the star head's prologue zeroes the statistics)
    mcl_stats_t tmp{};
    *done = false;
    int rc = star_step(t, p, disp, &tmp, s, done);
    if (rc != MCL_OK) return rc;
    if (*done) {
        MCL_TRY(stats_end(t, s, st, t0));
        if (st) {
            st->iterations = tmp.iterations;
            st->star_retries = tmp.star_retries;
        }
        return MCL_OK;
    }
    auto t1 = std::chrono::steady_clock::now();
    MCL_TRY(ensure_mesh(t, s));
    rc = step_impl(t, p, disp, st, stream);
    if (st) st->fallbacks = 1;
    if (flip_ms && rc == MCL_OK) {
        MCL_HIP(hipStreamSynchronize(s));
        *flip_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    }
    return rc;
}

// Path choice.  MCL_OPT_STAR_ENABLE = 2 (default): each step takes the path
// that measured faster (time on the stream over the step's own work, a star
// step that fell back counting its attempt and the flip-based step after it;
// the star path's time averaged over its steps, the flip path's last); the
// other path is measured again after kPathProbe x (slow / fast) steps, so
// that probing costs at most 1 / kPathProbe of the run, and twice as late
// each time the measurement keeps the choice (up to 4096 steps); a change in
// the motion still moves the choice.  = 1: the star step every step, except that after
// a fallback it is not tried for 1, 2, 4, ... 64 steps.  Meshes below
// MCL_OPT_STAR_MIN_VERTICES (where the number of launches dominates) take the
// flip-based step.
static int kinetic_step(mcl_tri* t, const mcl_step_params_t* p, const float2* disp, mcl_stats_t* st, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const bool star_ok = (t->m.periodic || t->m.ncorner > 0) && p->variant == 0 && t->star_deg[0] &&
                         t->star_enable != 0 && t->m.nv >= t->star_min_vertices;
    if (!star_ok) {
        MCL_TRY(ensure_mesh(t, s));
        return step_impl(t, p, disp, st, stream);
    }
    if (t->star_enable == 2) {
        int path;
        bool probe = false;
        if (t->path_ms[1] < 0.0) path = 1;
        else if (t->path_ms[0] < 0.0) path = 0;
        else {
            path = t->path_ms[1] <= t->path_ms[0] ? 1 : 0;
            if (++t->path_since >= t->path_probe) {
                path = 1 - path;
                probe = true;
                t->path_since = 0;
            }
        }
        // The path's time on the stream, between events: from the point the
        // GPU reaches the step (after the previous step's work) to the end of
        // the path's own work (a settled star step: its last counter read,
        // before any timestep tail it enqueued; the flip-based step: its
        // end), host round trips inside the step included.  No host wait for
        // the previous work.
        t->captured = false;
        t->star_end_event = false;
        MCL_HIP(hipEventRecord(t->ev_step, s));
        int rc;
        double flip_ms = -1.0;
        if (path == 1) {
            bool done = false;
            rc = star_or_fallback(t, p, disp, st, stream, &done, &flip_ms);
        } else {
            MCL_TRY(ensure_mesh(t, s));
            rc = step_impl(t, p, disp, st, stream);
        }
        if (rc != MCL_OK) return rc;
        hipEvent_t end = t->ev_aux;
        if (!t->star_end_event) {
            end = t->ev_end;
            MCL_HIP(hipEventRecord(end, s));
        }
        MCL_HIP(hipEventSynchronize(end));
        float ev_ms = 0.0f;
        MCL_HIP(hipEventElapsedTime(&ev_ms, t->ev_step, end));
        const double ms = ev_ms;
        if (t->captured) return MCL_OK;  // a step that captured a graph is not a measurement of its path
        if (probe) {  // the probed path stays slower: probe half as often
            const bool keeps = ms > t->path_ms[1 - path];
            t->path_probe = keeps ? std::min(2 * t->path_probe, 4096) : kPathProbe;
        }
        // the star path's cost is averaged over its steps (weight 0.2 per
        // step): a rare step that falls back to flips (attempt + flip step)
        // does not turn the choice for the next ~64 x (slow / fast) steps,
        // while a star step that keeps falling back soon costs its attempt
        // plus the flip step
        if (path == 1 && t->path_ms[1] > 0.0) t->path_ms[1] = 0.8 * t->path_ms[1] + 0.2 * ms;
        else t->path_ms[path] = ms;
        if (flip_ms >= 0.0) t->path_ms[0] = flip_ms;  // a fallback is also a flip-step measurement
        if (t->path_ms[0] > 0.0 && t->path_ms[1] > 0.0) {
            // a probe costs (slow - fast); after kPathProbe * slow / fast steps
            // of the fast path that is at most 1 / kPathProbe of the time
            const double r = std::max(t->path_ms[0], t->path_ms[1]) / std::min(t->path_ms[0], t->path_ms[1]);
            t->path_probe = std::max(t->path_probe, (int)std::min(4096.0, kPathProbe * r));
        }
        return MCL_OK;
    }
    if (t->star_skip > 0) {
        t->star_skip--;
        MCL_TRY(ensure_mesh(t, s));
        return step_impl(t, p, disp, st, stream);
    }
    bool done = false;
    MCL_TRY(star_or_fallback(t, p, disp, st, stream, &done));
    if (done) {
        t->star_backoff = 0;
    } else {
        t->star_backoff = t->star_backoff ? std::min(2 * t->star_backoff, 64) : 1;
        t->star_skip = t->star_backoff;
    }
    return MCL_OK;
}

// vertices not at their target (exact compare); the first 4 into n[2..5]
// (n[1]: listed so far)
__global__ void __launch_bounds__(kThreads) k_off_target(DMesh m, const float2* __restrict__ target, int* __restrict__ n) {
    int c = 0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nv; v += gridDim.x * blockDim.x) {
        const float2 p = m.pos[v], q = target[v];
        if (!(p.x == q.x && p.y == q.y)) {
            c++;
            const int k = atomicAdd(&n[1], 1);
            if (k < 4) n[2 + k] = v;
        }
    }
    if (c) atomicAdd(n, c);
}

// The exact Delaunay triangulation of device points, built on the GPU: a
// valid triangulation of the points at row positions (mcl_gbuild.hip), then
// flip-based kinetic steps to the points' own positions (repeated while a
// vertex stalled).  One box or torus triangulation (ninst = 1), or an
// ensemble of ninst box triangulations of counts[i] points each; vertex
// numbering as mcl_hmesh_build's and mcl_hmesh_build_batch's.
static int build_impl(const void* d_xy, const int* counts, int ninst, int periodic, void* d_perm, mcl_tri_t** out,
                      void* stream) {
    if (!out || !d_xy || !counts || ninst <= 0 || (periodic && ninst != 1)) return MCL_ERR_INPUT;
    long long n = 0;
    for (int i = 0; i < ninst; i++) {
        if (counts[i] < 0) return MCL_ERR_INPUT;
        n += counts[i];
    }
    if (n <= 0) return MCL_ERR_INPUT;
    const long long base = periodic ? 0 : 4LL * ninst;
    const long long nhe = 6 * n + (periodic ? 0 : 10LL * ninst);
    if (nhe >= (1LL << 31)) return MCL_ERR_CAPACITY;
    hipStream_t s = (hipStream_t)stream;
    mcl_tri* t = nullptr;
    MCL_TRY(tri_make((int)(n + base), (int)nhe, (int)(2 * n + (periodic ? 0 : 2LL * ninst)), periodic, (int)base, &t));
    float2* target = nullptr;
    int* offc = nullptr;
    int rc = dalloc(&target, (size_t)(n + base));
    if (rc == MCL_OK) rc = dalloc(&offc, 8);
    if (rc == MCL_OK) rc = gbuild(t->m, (const float2*)d_xy, counts, ninst, target, (int*)d_perm, s);
    if (rc == MCL_OK) rc = validate(t, s, "device build");
    mcl_step_params_t p{};
    p.variant = 0;
    for (int it = 0; rc == MCL_OK; it++) {
        // straight to the targets; after 8 passes, every other pass goes half
        // way (a vertex that keeps stalling on the straight move meets other
        // links on the way)
        p.motion = (it >= 8 && (it & 1) == 0) ? kMotionTargetHalf : kMotionTarget;
        rc = step_impl(t, &p, target, nullptr, stream);
        if (rc != MCL_OK) break;
        hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, offc, 6, 0);
        hipLaunchKernelGGL(k_off_target, dim3(blocks_for(t->m.nv)), dim3(kThreads), 0, s, t->m, (const float2*)target,
                           offc);
        int oc[6] = {0, 0, 0, 0, 0, 0};
        if (hipMemcpyAsync(oc, offc, sizeof(oc), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = MCL_ERR_HIP;
            break;
        }
        const int off = oc[0];
        if (trace_on()) {
            fprintf(stderr, "  [device build] pass %d: %d vertices off target\n", it, off);
            for (int k = 0; k < std::min(off, 4); k++) {
                float2 a, b;
                int deg = 0;
                (void)hipMemcpy(&a, t->m.pos + oc[2 + k], sizeof(a), hipMemcpyDeviceToHost);
                (void)hipMemcpy(&b, target + oc[2 + k], sizeof(b), hipMemcpyDeviceToHost);
                int h0 = 0, h = 0;
                (void)hipMemcpy(&h0, t->m.v_he + oc[2 + k], sizeof(int), hipMemcpyDeviceToHost);
                h = h0;
                do {  // degree, by the half-edge walk
                    int tw = 0;
                    (void)hipMemcpy(&tw, t->m.he_next + (h ^ 1), sizeof(int), hipMemcpyDeviceToHost);
                    h = tw;
                    deg++;
                } while (h != h0 && deg < 4096);
                fprintf(stderr, "    vertex %d at (%.9g, %.9g) target (%.9g, %.9g) degree %d\n", oc[2 + k], a.x, a.y,
                        b.x, b.y, deg);
            }
        }
        if (off == 0) break;
        if (it == 31) rc = MCL_ERR_UNRESOLVED;
    }
    if (target) (void)hipFree(target);
    if (offc) (void)hipFree(offc);
    if (rc != MCL_OK) { tri_free(t); delete t; return rc; }
    tune_tri(t, periodic, (int)base, t->m.nv);
    *out = t;
    return MCL_OK;
}

extern "C" int mcl_tri_build(const void* d_xy, int n, int periodic, void* d_perm, mcl_tri_t** out, void* stream) {
    return build_impl(d_xy, &n, 1, periodic, d_perm, out, stream);
}

extern "C" int mcl_batch_build(const void* d_xy, const int* counts, int ninst, void* d_perm, mcl_tri_t** out,
                               void* stream) {
    return build_impl(d_xy, counts, ninst, 0, d_perm, out, stream);
}

// A step that leaves a vertex short of its target (other than a duplicate
// target) reports MCL_ERR_UNRESOLVED; the mesh is still valid and exactly
// Delaunay for the positions it holds.
extern "C" int mcl_tri_step(mcl_tri_t* t, const mcl_step_params_t* p, mcl_stats_t* st, void* stream) {
    if (!t || !p) return MCL_ERR_INPUT;
    mcl_stats_t local{};
    mcl_stats_t* sp = st ? st : &local;
    int rc = kinetic_step(t, p, nullptr, sp, stream);
    if (rc == MCL_OK && sp->stalled > 0) rc = MCL_ERR_UNRESOLVED;
    return rc;
}

extern "C" int mcl_tri_move_legalize(mcl_tri_t* t, const void* d_disp, int variant, mcl_stats_t* st, void* stream) {
    if (!t || !d_disp || (variant != 0 && variant != 1)) return MCL_ERR_INPUT;
    mcl_step_params_t p{};
    p.variant = variant;
    mcl_stats_t local{};
    mcl_stats_t* sp = st ? st : &local;
    int rc = kinetic_step(t, &p, (const float2*)d_disp, sp, stream);
    if (rc == MCL_OK && sp->stalled > 0) rc = MCL_ERR_UNRESOLVED;
    return rc;
}

extern "C" void* mcl_tri_target_ptr(mcl_tri_t* t) { return t ? (void*)t->tgt : nullptr; }

// ---------------------------------------------------------------- one ring
extern "C" int mcl_tri_one_ring(mcl_tri_t* t, int K, int layout, void* d_ring, int* overflow, void* stream) {
    if (!t || !d_ring || K < 2 || (layout != 0 && layout != 1)) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    MCL_TRY(ensure_mesh(t, s));
    hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, t->cnt + kCnMisc, 1, 0);
    hipLaunchKernelGGL(k_ring_legacy, dim3(blocks_for(t->m.nv)), dim3(kThreads), 0, s, t->m, K, layout, (int*)d_ring,
                       t->cnt + kCnMisc);
    MCL_HIP(hipGetLastError());
    int ov = 0;
    MCL_HIP(hipMemcpyAsync(&ov, t->cnt + kCnMisc, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    if (overflow) *overflow = ov;
    return ov ? MCL_ERR_CAPACITY : MCL_OK;
}

// the CSR one-ring, asynchronous (cap >= nhe: every half-edge leaves one
// vertex), from the stars in buffer c when from_stars, else from the half-edges
static int enqueue_ring_csr(mcl_tri* t, hipStream_t s, int c, bool from_stars, int* off, int* nbrs, int cap) {
    // offsets: the scan of the star degrees (from the stars), or of the degrees
    // counted on the half-edges
    if (!from_stars)
        hipLaunchKernelGGL(k_degree, dim3(blocks_for(t->m.nv)), dim3(kThreads), 0, s, t->m, off);
    MCL_TRY(scan_exclusive(from_stars ? (const int*)t->star_deg[c] : off, off, t->m.nv, t->scan_tmp, s));
    if (from_stars)
        hipLaunchKernelGGL(k_star_ring_csr, dim3(blocks_for(t->m.nv)), dim3(256), 0, s, t->m.nv, t->m.ncorner,
                           (const int*)t->star_deg[c], (const int*)t->star_nbr[c], (const int*)off, nbrs, cap);
    else
        hipLaunchKernelGGL(k_ring_csr, dim3(blocks_for(t->m.nv)), dim3(kThreads), 0, s, t->m, off, nbrs, cap);
    MCL_HIP(hipGetLastError());
    return MCL_OK;
}

// the whole timestep with the tail q (see mcl_tri_timestep)
static int timestep_impl(mcl_tri* t, const mcl_step_params_t* p, const TailReq& q, mcl_stats_t* st, hipStream_t s) {
    const bool want = q.ring_off || q.has_frnn();
    // a star step enqueues the FRNN part on the side stream after its first
    // check (a flip-based step right after its move); frnn_done: it need not
    // run again
    t->tail_req = want ? &q : nullptr;
    t->frnn_done = t->ring_done = false;
    t->frnn_on_targets = false;
    int rc = mcl_tri_step(t, p, st, s);
    t->tail_req = nullptr;
    // the FRNN part run on the side stream (if any) is part of this call's work
    const int jrc = join_side(t, s);
    if (rc != MCL_OK && rc != MCL_ERR_UNRESOLVED) return rc;
    MCL_TRY(jrc);
    if (!want) return rc;
    if (t->frnn_on_targets && t->frnn_done) {
        // a flip-based step's FRNN part ran on the targets: it holds the final
        // positions when every vertex is at its target
        t->frnn_on_targets = false;
        hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, t->d_offt, 6, 0);
        hipLaunchKernelGGL(k_off_target, dim3(blocks_for(t->m.nv)), dim3(kThreads), 0, s, t->m,
                           (const float2*)t->tgt, t->d_offt);
        int off = 0;
        MCL_HIP(hipMemcpyAsync(&off, t->d_offt, sizeof(int), hipMemcpyDeviceToHost, s));
        MCL_HIP(hipStreamSynchronize(s));
        if (off != 0 || rc != MCL_OK) t->frnn_done = false;
    }
    if (!t->frnn_done) MCL_TRY(launch_tail(t, q, s, 0, t->star_cur, t->stars_valid));
    if (!t->stars_valid) MCL_TRY(ensure_mesh(t, s));
    if (!t->ring_done || !t->stars_valid) MCL_TRY(launch_tail(t, q, s, 1, t->star_cur, t->stars_valid));
    return rc;
}

extern "C" int mcl_tri_timestep(mcl_tri_t* t, const mcl_step_params_t* p, void* d_ring_off, void* d_ring_nbrs,
                                int ring_cap, mcl_grid_t* g, float rad, void* d_frnn_off, void* d_frnn_nbrs,
                                int frnn_cap, mcl_stats_t* st, void* stream) {
    if (!t || !p || (d_ring_off && (!d_ring_nbrs || ring_cap < t->m.nhe)) || (g && !d_frnn_off)) return MCL_ERR_INPUT;
    const TailReq q{(int*)d_ring_off, (int*)d_ring_nbrs, ring_cap, g, rad, (int*)d_frnn_off, (int*)d_frnn_nbrs,
                    frnn_cap};
    return timestep_impl(t, p, q, st, (hipStream_t)stream);
}

extern "C" int mcl_tri_timestep_batch(mcl_tri_t* t, const mcl_step_params_t* p, void* d_ring_off, void* d_ring_nbrs,
                                      int ring_cap, mcl_bgrid_t* bg, void* d_frnn_off, void* d_frnn_nbrs,
                                      int frnn_cap, mcl_stats_t* st, void* stream) {
    if (!t || !p || (d_ring_off && (!d_ring_nbrs || ring_cap < t->m.nhe)) || (bg && !d_frnn_off)) return MCL_ERR_INPUT;
    if (bg) {
        int npoints = 0, first = 0, ncells = 0;
        MCL_TRY(mcl_bgrid_sizes(bg, &npoints, &first, &ncells));
        if (first + npoints > t->m.nv) return MCL_ERR_INPUT;  // its points must be vertices of this mesh
    }
    TailReq q{(int*)d_ring_off, (int*)d_ring_nbrs, ring_cap, nullptr, 0.0f, (int*)d_frnn_off, (int*)d_frnn_nbrs,
              frnn_cap};
    q.bg = bg;
    return timestep_impl(t, p, q, st, (hipStream_t)stream);
}

extern "C" int mcl_tri_one_ring_csr(mcl_tri_t* t, void* d_offsets, void* d_nbrs, int cap, int* total, void* stream) {
    if (!t || !d_offsets) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    int* off = (int*)d_offsets;
    // with current stars the ring comes from them (same order as the walk)
    const bool from_stars = t->stars_valid;
    const int c = t->star_cur;
    // offsets: the scan of the star degrees (from the stars), or of the degrees
    // counted on the half-edges
    if (!from_stars)
        hipLaunchKernelGGL(k_degree, dim3(blocks_for(t->m.nv)), dim3(kThreads), 0, s, t->m, off);
    MCL_TRY(scan_exclusive(from_stars ? (const int*)t->star_deg[c] : off, off, t->m.nv, t->scan_tmp, s));
    auto fill = [&]() {
        if (from_stars)
            hipLaunchKernelGGL(k_star_ring_csr, dim3(blocks_for(t->m.nv)), dim3(256), 0, s, t->m.nv, t->m.ncorner,
                               (const int*)t->star_deg[c], (const int*)t->star_nbr[c], (const int*)off, (int*)d_nbrs, cap);
        else
            hipLaunchKernelGGL(k_ring_csr, dim3(blocks_for(t->m.nv)), dim3(kThreads), 0, s, t->m, off, (int*)d_nbrs, cap);
    };
    if (!total && d_nbrs && cap >= t->m.nhe) {
        // every half-edge leaves exactly one vertex: the total is nhe, no read-back needed
        fill();
        MCL_HIP(hipGetLastError());
        return MCL_OK;
    }
    int tot = 0;
    MCL_HIP(hipMemcpyAsync(&tot, off + t->m.nv, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    if (total) *total = tot;
    if (d_nbrs && cap > 0) {
        fill();
        MCL_HIP(hipGetLastError());
        MCL_HIP(hipStreamSynchronize(s));
    }
    return tot > cap ? MCL_ERR_CAPACITY : MCL_OK;
}

extern "C" int mcl_tri_closest(mcl_tri_t* t, const void* d_offsets, const void* d_nbrs, void* d_out, void* stream) {
    if (!t || !d_offsets || !d_nbrs || !d_out) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_closest, dim3(blocks_for(t->m.nv)), dim3(kThreads), 0, s, t->m, (const int*)d_offsets,
                       (const int*)d_nbrs, (int*)d_out);
    MCL_HIP(hipGetLastError());
    MCL_HIP(hipStreamSynchronize(s));
    return MCL_OK;
}

extern "C" int mcl_tri_closest_ring(mcl_tri_t* t, int K, int layout, const void* d_ring, void* d_out, void* stream) {
    if (!t || !d_ring || !d_out || K < 2 || (layout != 0 && layout != 1)) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_closest_legacy, dim3(blocks_for(t->m.nv)), dim3(kThreads), 0, s, t->m, K, layout,
                       (const int*)d_ring, (int*)d_out);
    MCL_HIP(hipGetLastError());
    MCL_HIP(hipStreamSynchronize(s));
    return MCL_OK;
}

extern "C" int mcl_tri_frnn_graph(mcl_tri_t* t, float r, int K, int layout, const void* d_ring, int M, void* d_out,
                                  int* overflow, void* stream) {
    if (!t || !d_ring || !d_out || K < 2 || M < 1 || (layout != 0 && layout != 1) || !(r >= 0.0f)) return MCL_ERR_INPUT;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, s, t->cnt + kCnMisc, 1, 0);
    hipLaunchKernelGGL(k_frnn_graph, dim3(blocks_for(t->m.nv)), dim3(kThreads), 0, s, t->m, r * r, K, layout,
                       (const int*)d_ring, M, (int*)d_out, t->cnt + kCnMisc);
    MCL_HIP(hipGetLastError());
    int ov = 0;
    MCL_HIP(hipMemcpyAsync(&ov, t->cnt + kCnMisc, sizeof(int), hipMemcpyDeviceToHost, s));
    MCL_HIP(hipStreamSynchronize(s));
    if (overflow) *overflow = ov;
    return ov ? MCL_ERR_CAPACITY : MCL_OK;
}

extern "C" int mcl_tri_set_option(mcl_tri_t* t, int option, long long value) {
    if (!t) return MCL_ERR_INPUT;
    switch (option) {
        case MCL_OPT_STAR_MIN_VERTICES:
            if (value < 0) return MCL_ERR_INPUT;
            t->star_min_vertices = (int)std::min<long long>(value, INT32_MAX);
            return MCL_OK;
        case MCL_OPT_STAR_ENABLE:
            if (value < 0 || value > 2) return MCL_ERR_INPUT;
            t->star_enable = (int)value;
            t->path_ms[0] = t->path_ms[1] = -1.0;
            t->path_since = 0;
            t->path_probe = kPathProbe;
            t->star_skip = t->star_backoff = 0;
            return MCL_OK;
        case MCL_OPT_STAR_GIVE_UP:
            if (value < 1 || value > 100) return MCL_ERR_INPUT;
            t->star_give_up = (int)value;
            drop_graphs(t);
            return MCL_OK;
        case MCL_OPT_GRAPH:
            if (value < 0 || value > 1) return MCL_ERR_INPUT;
            t->graph = (int)value;
            return MCL_OK;
        case MCL_OPT_TAIL_THRESHOLD:
            if (value < 0 || value > (1 << 24)) return MCL_ERR_INPUT;
            t->tail_threshold = (int)value;
            return MCL_OK;
        case MCL_OPT_REVERT_TAIL:
            if (value < 0 || value > (1 << 24)) return MCL_ERR_INPUT;
            t->revert_tail = (int)value;
            return MCL_OK;
        case MCL_OPT_RELOC_TAIL:
            if (value < 0 || value > (1 << 24)) return MCL_ERR_INPUT;
            t->reloc_tail = (int)value;
            return MCL_OK;
        case MCL_OPT_BIG_GRID_MIN:
            if (value < 1 || value > (1 << 30)) return MCL_ERR_INPUT;
            t->big_grid_min = (int)value;
            return MCL_OK;
        case MCL_OPT_STAR_HEAD_PASSES:
            if (value < 1 || value > kStarPasses) return MCL_ERR_INPUT;
            t->star_head_passes = (int)value;
            drop_graphs(t);
            return MCL_OK;
        default:
            return MCL_ERR_INPUT;
    }
}

extern "C" int mcl_tri_get_option(const mcl_tri_t* t, int option, long long* value) {
    if (!t || !value) return MCL_ERR_INPUT;
    switch (option) {
        case MCL_OPT_STAR_MIN_VERTICES: *value = t->star_min_vertices; return MCL_OK;
        case MCL_OPT_STAR_ENABLE: *value = t->star_enable; return MCL_OK;
        case MCL_OPT_STAR_GIVE_UP: *value = t->star_give_up; return MCL_OK;
        case MCL_OPT_GRAPH: *value = t->graph; return MCL_OK;
        case MCL_OPT_TAIL_THRESHOLD: *value = t->tail_threshold; return MCL_OK;
        case MCL_OPT_REVERT_TAIL: *value = t->revert_tail; return MCL_OK;
        case MCL_OPT_RELOC_TAIL: *value = t->reloc_tail; return MCL_OK;
        case MCL_OPT_BIG_GRID_MIN: *value = t->big_grid_min; return MCL_OK;
        case MCL_OPT_STAR_HEAD_PASSES: *value = t->star_head_passes; return MCL_OK;
        default: return MCL_ERR_INPUT;
    }
}

extern "C" int mcl_tri_kernel_timer(mcl_tri_t* t, int enable) {
    if (!t) return MCL_ERR_INPUT;
    for (auto& k : t->timers) k.reset(enable != 0);
    return MCL_OK;
}
extern "C" int mcl_tri_kernel_timer_read(mcl_tri_t* t, int which, double* total_ms, long long* launches) {
    if (!t || which < 0 || which > 5) return MCL_ERR_INPUT;
    double ms = t->timers[which].total_ms(launches);
    if (total_ms) *total_ms = ms;
    return MCL_OK;
}
