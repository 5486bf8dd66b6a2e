/*
 * mcleap_hip.h -- C ABI of libmcleap_hip.so, the MI355X (gfx950) engine for
 * kinetic 2-D Delaunay maintenance of moving points.
 *
 * The reference (hporro/mcleap) is header-only C++/CUDA with no ABI; its
 * callers use three C++ types.  Each entry point below replaces the
 * reference member named in its comment (path:line in the reference tree).
 * The C++ drop-in headers include/mcleap/{Host_Triangulation,
 * Device_Triangulation,gridCount2d}.h wrap these calls with the reference's
 * class and method names.
 *
 * Conventions
 *  - Every function returns an int status (MCL_OK = 0).  No exception or
 *    C++ type crosses the ABI.
 *  - Device buffers are passed as plain device pointers (void*), sizes as int.
 *    Positions are float2 pairs (x, y); "AoS" arrays use the reference record
 *    layouts: Vertex{pos_index, he} = 2 ints, HalfEdge{next, v, t, op} =
 *    4 ints, Triangle{he} = 1 int (Common_Triangulation.h:42-57).
 *  - `stream` is a hipStream_t (NULL = the default stream).  Calls are
 *    synchronous on return unless stated, like the reference (which calls
 *    cudaDeviceSynchronize after every launch, e.g. Device_Triangulation.h:343).
 *  - One handle per host thread.
 */
#ifndef MCLEAP_HIP_H
#define MCLEAP_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define MCL_OK 0
#define MCL_ERR_HIP 1          /* a HIP runtime call failed */
#define MCL_ERR_INPUT 2        /* invalid argument / point outside the domain */
#define MCL_ERR_DUPLICATE 3    /* duplicate point rejected at load */
#define MCL_ERR_TOPOLOGY 4     /* mesh invariant broken / cannot build */
#define MCL_ERR_CAPACITY 5     /* caller buffer too small (count still reported) */
#define MCL_ERR_UNRESOLVED 6   /* untangling left inverted triangles */
#define MCL_ERR_ROUND_CAP 7    /* flip rounds exceeded the cap */
#define MCL_ERR_NO_DEVICE 8    /* no usable gfx950 device */

/* --------------------------------------------------------------- statistics */
typedef struct {
    long long rounds;          /* Lawson flip rounds executed */
    long long flips;           /* Lawson flips */
    long long untangle_flips;  /* B/C-step flips (untangle) */
    long long reverts;         /* vertex moves taken back because a triangle inverted */
    long long iterations;      /* move/legalise iterations of the kinetic step */
    long long stalled;         /* vertices that could not reach their target */
    long long inverted;        /* inverted triangles seen after the first move */
    long long relocated;       /* vertices removed and re-inserted at their targets */
    long long pred_orient[4];  /* orientation decided at FP32 / FP64 / exact-small / exact-big */
    long long pred_incircle[4];/* incircle decided at FP32 / FP64 / exact-small / exact-big */
    double ms;                 /* wall time of the call, host clock */
    long long relocated_grid;  /* of `relocated`, those done by the full-grid apply launches (kernel timer 3) */
    long long star_retries;    /* torus step: vertices whose star was recomputed from the 3-ring */
    long long fallbacks;       /* torus step: 1 when the stars did not settle and the flip-based step ran */
    long long duplicates;      /* vertices left at their previous position because another vertex
                                  holds exactly the same target (not counted in `stalled`) */
} mcl_stats_t;

/* =========================================================================
 * Host triangulation build  (HostTriangulation, Host_Triangulation.h:24-98)
 * ========================================================================= */
typedef struct mcl_hmesh mcl_hmesh_t;

/* HostTriangulation() + addDelaunayPoints(points) (Host_Triangulation.h:107-137,
 * 296-300).  periodic = 0: the reference layout with 4 corners at +-1e7
 * (vertex ids 0..3) and input point perm[k] as vertex 4+k.  periodic = 1:
 * unit-torus triangulation of points in [0,1)^2, no corners, vertex k = perm[k].
 * Exact predicates; duplicates are rejected (MCL_ERR_DUPLICATE). */
int mcl_hmesh_build(const float* xy, int n, int periodic, mcl_hmesh_t** out);
/* sizes of m_pos, m_v, m_he, m_t (Host_Triangulation.h:27-30) */
int mcl_hmesh_sizes(const mcl_hmesh_t* m, int* npos, int* nv, int* nhe, int* nt);
/* copy out the AoS arrays (any pointer may be NULL); perm has n entries */
int mcl_hmesh_export(const mcl_hmesh_t* m, float* pos, int* v2, int* he4, int* t1, int* perm);
int mcl_hmesh_free(mcl_hmesh_t* m);
/* An ensemble of ninst independent box triangulations as ONE mesh, so that
 * every device call advances all of them together (no counterpart in the
 * reference, which builds one HostTriangulation per instance,
 * Host_Triangulation.h:107-137, 296-300).  xy holds the instances' points back
 * to back, counts[i] of them for instance i.  Vertex layout: the 4 corners of
 * every instance first (instance i: 4i..4i+3, at +-1e7 like a single box
 * mesh), then instance 0's points, instance 1's, ...; perm maps the points'
 * vertex order to indices into xy.  Instances are built in nthreads host
 * threads (0: all cores). */
int mcl_hmesh_build_batch(const float* xy, const int* counts, int ninst, int nthreads, mcl_hmesh_t** out);
/* ninst + 1 offsets each: instance i's points are vertices
 * vstart[i]..vstart[i+1]-1, its half-edges hestart[i]..hestart[i+1]-1, its
 * triangles tstart[i]..tstart[i+1]-1.  MCL_ERR_INPUT for a single mesh. */
int mcl_hmesh_batch_offsets(const mcl_hmesh_t* m, int* vstart, int* hestart, int* tstart);

/* Host-side operations on the caller's AoS arrays (pos: float pairs indexed
 * by Vertex.pos_index; he4/t1/v2 as above), exact predicates, in place. */
/* orient2d / inCircle (Helpers_Triangulation.h:86-151) with exact signs;
 * incircle ties broken by the symbolic perturbation (never 0 unless the four
 * points are collinear).  periodic != 0: minimum-image differences. */
int mcl_orient2d(const float* a, const float* b, const float* c, int periodic);
int mcl_incircle(const float* a, const float* b, const float* c, const float* d, int periodic);
/* f2to2 of the edge of he_index (Common_Triangulation.h:69-127, the flip of
 * delonizeEdge/untangleEdge); check != 0 refuses (MCL_ERR_TOPOLOGY) a flip
 * that would leave a triangle not strictly CCW */
int mcl_hmesh_flip(float* pos, int* he4, int nhe, int* t1, int nt, int* v2, int nv, int periodic,
                   int he_index, int check);
/* delonize / delonizeVertex (Host_Triangulation.h:305-319, 323-346,
 * 368-395): Lawson flips from a work stack of edges (all edges, or the edges
 * of the given half-edges `seeds`) until every interior edge passes the exact
 * incircle test.  A reflex quad (possible only on a tangled mesh) is left
 * unflipped and reported as MCL_ERR_UNRESOLVED. */
int mcl_hmesh_delonize(float* pos, int* he4, int nhe, int* t1, int nt, int* v2, int nv, int periodic,
                       const int* seeds, int nseeds, long long* flips);
/* moveFlipflop / moveFlipflopDelaunay / moveUntangling / moveUntanglingDelaunay
 * (Host_Triangulation.h:77-80, declared and left unimplemented there,
 * :408-413) on the host arrays: vertex v moves by d (wrapped on the torus).
 * MCL_MOVE_FLIPFLOP: v is removed from its star (flips down to degree 3 and a
 * 3 -> 1 merge) and re-inserted at the target (1 -> 3, or 2 -> 4 on an edge);
 * every triangle stays strictly CCW, the mesh need not be Delaunay.
 * MCL_MOVE_FLIPFLOP_DELAUNAY: then Lawson flips from the rewritten triangles.
 * MCL_MOVE_UNTANGLING: v simply moves, then B/C untangle passes run until
 * none flips; MCL_ERR_UNRESOLVED if a triangle stays inverted.
 * MCL_MOVE_UNTANGLING_DELAUNAY: then Lawson flips over every edge.
 * A target on another vertex gives MCL_ERR_DUPLICATE, one outside the mesh
 * (or a vertex on the outer face) MCL_ERR_INPUT; *flips = flips made. */
#define MCL_MOVE_FLIPFLOP 0
#define MCL_MOVE_FLIPFLOP_DELAUNAY 1
#define MCL_MOVE_UNTANGLING 2
#define MCL_MOVE_UNTANGLING_DELAUNAY 3
int mcl_hmesh_move_vertex(float* pos, int* he4, int nhe, int* t1, int nt, int* v2, int nv, int periodic,
                          int v, const float* d, int mode, long long* flips);
/* untangle (Host_Triangulation.h:417-471): one pass of the B/C steps over all
 * edges with exact orientation; *flips = flips made */
int mcl_hmesh_untangle(float* pos, int* he4, int nhe, int* t1, int nt, int* v2, int nv, int periodic,
                       long long* flips);

/* =========================================================================
 * Device triangulation  (DeviceTriangulation, Device_Triangulation.h:128-223)
 * ========================================================================= */
typedef struct mcl_tri mcl_tri_t;
typedef struct mcl_grid mcl_grid_t; /* the uniform grid of GridCount2d, declared below */

/* DeviceTriangulation(HostTriangulation*): allocateMemory + transferToDevice
 * (Device_Triangulation.h:232-236, 255-285).  The AoS host arrays are stored
 * on the device as SoA (he_next, he_v, he_t, he_op, v_he, t_he, float2 pos).
 * periodic != 0 marks a torus mesh (minimum-image geometry). */
int mcl_tri_create(const float* pos, int npos, const int* v2, int nv, const int* he4, int nhe,
                   const int* t1, int nt, int periodic, mcl_tri_t** out);
/* ~DeviceTriangulation / freeMemory (Device_Triangulation.h:238-241, 287-306) */
int mcl_tri_destroy(mcl_tri_t* t);
/* m_pos_size, m_v_size, m_he_size, m_t_size (Device_Triangulation.h:134-137) */
int mcl_tri_sizes(const mcl_tri_t* t, int* npos, int* nv, int* nhe, int* nt);
/* transferToHost (Device_Triangulation.h:245-253): SoA -> AoS; any may be NULL */
int mcl_tri_download(mcl_tri_t* t, float* pos, int* v2, int* he4, int* t1);
/* HostTriangulation() + addDelaunayPoints(points) + DeviceTriangulation(host)
 * (Host_Triangulation.h:107-137, 296-300; Device_Triangulation.h:232-306) in
 * one call, built on the GPU from n device points (float2): a valid
 * triangulation at row positions, then kinetic steps to the points.
 * periodic = 0: box layout, points inside the corners; periodic = 1: the
 * unit torus, points in [0,1)^2, n >= 16, rows of points less than half the
 * torus apart (else MCL_ERR_INPUT).  The result is the exact Delaunay
 * triangulation with the vertex numbering of mcl_hmesh_build (vertex base + k
 * = point d_perm[k], base 4 in a box, 0 on the torus; d_perm may be NULL).
 * MCL_ERR_DUPLICATE for repeated points; MCL_ERR_UNRESOLVED when vertices
 * cannot reach their points (all points on one line). */
int mcl_tri_build(const void* d_xy, int n, int periodic, void* d_perm, mcl_tri_t** out, void* stream);
/* mcl_tri_build for an ensemble of ninst independent box triangulations held
 * as one mesh (no counterpart in the reference, which builds one
 * HostTriangulation per instance, Host_Triangulation.h:107-137, 296-300).
 * d_xy holds the instances' points back to back on the device, counts (a
 * host array) gives each instance's number of points (0 allowed, total > 0).
 * Result: the exact Delaunay triangulation of every instance, with the
 * vertex layout and numbering of mcl_hmesh_build_batch (corners of instance
 * i at 4i..4i+3, then every instance's points in its own Hilbert order;
 * vertex 4 ninst + k = point d_perm[k] of d_xy).  Half-edges and triangles
 * of one instance are not contiguous. */
int mcl_batch_build(const void* d_xy, const int* counts, int ninst, void* d_perm, mcl_tri_t** out, void* stream);
/* 1 for a torus triangulation, 0 for a box (or ensemble of boxes) */
int mcl_tri_is_periodic(const mcl_tri_t* t, int* periodic);

/* ------------------------------------------------------------ state files
 * Save / resume (no counterpart in the reference, whose state lives only in
 * the host object's vectors, Host_Triangulation.h:27-30).  The file holds
 * exactly those AoS arrays (pos, v2, he4, t1) after a 40-byte header
 * ("MCLTRI01", version 1, periodic, npos, nv, nhe, nt, and the library's
 * flip-round and stamp counters, 0 from mcl_state_save) and
 * ends with an FNV-1a 64 checksum of the arrays; little-endian.  Loading a
 * file that is truncated, corrupted or not such a file gives MCL_ERR_INPUT.
 * Steps resumed from a saved state reproduce the uninterrupted run: the
 * motion is a function of (seed, step, vertex). */
int mcl_state_save(const char* path, int periodic, const float* pos, int npos, const int* v2, int nv,
                   const int* he4, int nhe, const int* t1, int nt);
int mcl_state_info(const char* path, int* periodic, int* npos, int* nv, int* nhe, int* nt);
/* any array may be NULL (skipped, still checksummed) */
int mcl_state_load(const char* path, float* pos, int* v2, int* he4, int* t1);
/* transferToHost + mcl_state_save; mcl_state_load + mcl_tri_create */
int mcl_tri_save(mcl_tri_t* t, const char* path);
int mcl_tri_load(const char* path, mcl_tri_t** out);
/* transferToDevice (Device_Triangulation.h:255-263) */
int mcl_tri_upload(mcl_tri_t* t, const float* pos, const int* v2, const int* he4, const int* t1);
/* public m_pos (device glm::vec2*, Device_Triangulation.h:139): float2 device pointer */
void* mcl_tri_pos_ptr(mcl_tri_t* t);

/* movePoints(glm::vec2* d): pos[i] += d[i] for non-corner vertices
 * (Device_Triangulation.h:395-402, kernels.cuh:38-43); d is a device float2 array */
int mcl_tri_move(mcl_tri_t* t, const void* d_disp, void* stream);
/* integrateMovePoints(float* dt, glm::vec2* v): pos[i] += v[i]*dt[i]
 * (Device_Triangulation.h:403-410, kernels.cuh:45-50) */
int mcl_tri_integrate(mcl_tri_t* t, const void* d_dt, const void* d_vel, void* stream);

/* moveFlipflop / moveFlipflopDelaunay / moveUntangling / moveUntanglingDelaunay
 * (Device_Triangulation.h:183-186, stubs returning false at :411-417) for n
 * vertices at once: ids[i] (host array, distinct, not box corners) moves by
 * disp[2i], disp[2i+1] (host array).  Modes as mcl_hmesh_move_vertex; the
 * flip-flop modes run the kinetic step's relocation rounds on the listed
 * vertices (claims, parallel removal and re-insertion), then for
 * MCL_MOVE_FLIPFLOP_DELAUNAY the Lawson rounds.  MCL_ERR_UNRESOLVED (mesh
 * still valid) when a vertex did not reach its target (st->stalled,
 * st->duplicates) or, in the untangling modes, triangles stay inverted.
 * The targets are readable through mcl_tri_target_ptr afterwards. */
int mcl_tri_move_vertices(mcl_tri_t* t, const int* ids, const float* disp, int n, int mode, mcl_stats_t* st,
                          void* stream);
/* delonize / delonize2 (Device_Triangulation.h:324-390): parallel Lawson flip
 * rounds until no edge is illegal.  Requires a mesh with no inverted
 * triangle.  Exact incircle; MCL_ERR_ROUND_CAP if max_rounds is hit
 * (max_rounds <= 0: unlimited).  variant 0 = device-driven rounds with
 * worklists and FP32-first predicates, variant 1 = one launch per round with
 * a host round trip and an FP64-first predicate (the reference's loop shape). */
int mcl_tri_delonize(mcl_tri_t* t, int variant, int max_rounds, mcl_stats_t* st, void* stream);
/* untangle / untangle2 (Device_Triangulation.h:422-495): rounds of B/C-step
 * flips of creased edges (kernels.cuh:302-362, Host_Triangulation.h:425-471)
 * with exact orientation.  Returns MCL_ERR_UNRESOLVED when inverted
 * triangles remain (the reference prints "couldn't fix everything" and
 * returns false). */
int mcl_tri_untangle(mcl_tri_t* t, int max_rounds, mcl_stats_t* st, void* stream);
/* number of triangles that are not strictly counter-clockwise (exact) */
int mcl_tri_count_inverted(mcl_tri_t* t, long long* out, void* stream);
/* number of interior edges that fail the exact Delaunay test */
int mcl_tri_count_non_delaunay(mcl_tri_t* t, long long* out, void* stream);

/* ---------------------------------------------------- kinetic step driver */
/* synthetic displacement fields of the benchmark workloads */
#define MCL_MOTION_GAUSS 0     /* Gaussian noise sigma, clamped to [0, 1-2^-24]^2 */
#define MCL_MOTION_SHEAR 1     /* torus: (a sin 2 pi y, a cos 2 pi x) + noise sigma, wrapped */

typedef struct {
    int motion;                /* MCL_MOTION_* */
    float sigma;               /* Gaussian noise standard deviation */
    float amp;                 /* shear-flow amplitude */
    unsigned long long seed;   /* counter-based RNG key */
    unsigned long long step;   /* step index (RNG counter) */
    int variant;               /* 0 = device-driven (fast), 1 = plain (host round trips, FP64-first) */
    int max_iterations;        /* move/legalise iterations before stalling vertices (<=0: 64) */
} mcl_step_params_t;

/* One kinetic step: displace every non-corner vertex by the synthetic motion,
 * and restore the exact Delaunay triangulation.  On the torus (variant 0)
 * every vertex recomputes its Delaunay star from its 2-ring in the previous
 * mesh (exact incircle on the inverted candidates), inconsistent stars are
 * recomputed from the 3-ring, and the half-edge arrays are rebuilt from the
 * stars (half-edge and triangle ids are renumbered; vertex ids are kept);
 * if the stars do not settle, the flip-based step below runs instead.
 * The flip-based step (box, and variant 1 everywhere): 
 * keep the mesh valid (repair flips; moves that would still invert a triangle
 * are taken back, and those vertices are removed and re-inserted at their
 * targets), and flip back to the exact Delaunay triangulation.  Replaces the
 * move + untangle2 + delonize2 sequence of moving_device_mesh.cu:119-155.
 * st->stalled counts vertices left short of their target for another reason
 * than a duplicate target (st->duplicates): a relocation whose target stays
 * occupied by a vertex that does not leave, or a star above 64 triangles. */
int mcl_tri_step(mcl_tri_t* t, const mcl_step_params_t* p, mcl_stats_t* st, void* stream);
/* movePoints(d) followed by untangle + delonize (moving_device_mesh.cu:119-155)
 * as one call that always ends on a valid mesh that is exactly Delaunay:
 * the kinetic step of mcl_tri_step with the caller's device displacement
 * array d (float2 per vertex, corners ignored; on the torus the targets
 * wrap into [0,1)).  Box targets must stay inside the corner square. */
int mcl_tri_move_legalize(mcl_tri_t* t, const void* d_disp, int variant, mcl_stats_t* st, void* stream);
/* Both step calls return MCL_ERR_UNRESOLVED (after finishing the step) when a
 * vertex could not reach its target for any reason other than a duplicate
 * target (st->stalled > 0); the mesh is then still valid and exactly
 * Delaunay for the positions it holds. */
/* the targets of the last step (device float2 per vertex; corners hold their
 * own positions): after a step every vertex is at its target except the
 * st->duplicates vertices that share a target with another vertex (no
 * counterpart in the reference, whose moves have no targets,
 * kernels.cuh:38-50) */
void* mcl_tri_target_ptr(mcl_tri_t* t);

/* One whole timestep of the reference's driver loop (moving_device_mesh.cu:
 * 109-243: move, untangle + legalize, one-ring, grid update + FRNN query):
 * mcl_tri_step, then, asynchronously, the CSR one-ring into d_ring_off /
 * d_ring_nbrs (as mcl_tri_one_ring_csr; ring_cap >= nhe) when d_ring_off is
 * given, and the grid g's counting sort of the new positions with its CSR
 * FRNN query into d_frnn_off / d_frnn_nbrs (as mcl_grid_build +
 * mcl_grid_frnn_csr with total = NULL: writes past frnn_cap dropped,
 * d_frnn_off[nump] the full count) when g is given (g sized for every vertex).
 * The grid part runs on a second stream of the library's, on the targets
 * (mcl_tri_target_ptr), beside the step (again after it, on `stream`, when a
 * vertex did not reach its target); `stream` waits for it before the call
 * returns, so every output is ordered on `stream` as if enqueued there.
 * With MCL_OPT_GRAPH the tail is one graph launch.  Status as mcl_tri_step. */
int mcl_tri_timestep(mcl_tri_t* t, const mcl_step_params_t* p, void* d_ring_off, void* d_ring_nbrs, int ring_cap,
                     mcl_grid_t* g, float rad, void* d_frnn_off, void* d_frnn_nbrs, int frnn_cap, mcl_stats_t* st,
                     void* stream);

/* ------------------------------------------------------------- queries */
/* oneRing<K>(int*) (layout 0, SoA: ring[(c+1)*nv + i], count in ring[i]) and
 * oneRing2<K>(int*) (layout 1, AoS: ring[(c+1) + K*i], count in ring[K*i]),
 * kernels.cuh:54-121: CW rotation from v.he, at most K-1 neighbours kept.
 * d_ring is caller-allocated device memory of nv*K ints.  Where the reference
 * truncates silently (kernels.cuh:68, "TODO: do something about this case"),
 * *overflow (may be NULL) receives the number of vertices other than box
 * corners with more than K-1 neighbours, and the call returns
 * MCL_ERR_CAPACITY when it is not 0 (the truncated rings are still written;
 * a corner's fan of hull points is truncated without being counted). */
int mcl_tri_one_ring(mcl_tri_t* t, int K, int layout, void* d_ring, int* overflow, void* stream);
/* one-ring as CSR (no cap): d_offsets nv+1 ints, d_nbrs cap ints; same CW
 * order as above.  *total = number of entries (MCL_ERR_CAPACITY if > cap).
 * With total == NULL and cap >= nhe the call is asynchronous (the entry
 * count always equals the half-edge count). */
int mcl_tri_one_ring_csr(mcl_tri_t* t, void* d_offsets, void* d_nbrs, int cap, int* total,
                         void* stream);
/* closestNeighbors<K>(ring, out) (kernels.cuh:129-146) over a CSR one-ring */
int mcl_tri_closest(mcl_tri_t* t, const void* d_offsets, const void* d_nbrs, void* d_out,
                    void* stream);
/* closestNeighbors<K> / closestNeighbors2<K> (Device_Triangulation.h:519-526,
 * 537-544; kernels.cuh:129-163) over the ring written by mcl_tri_one_ring
 * with the same K and layout (the AoS form reads ring[(c+1) + K*i], the
 * layout oneRing2 writes). */
int mcl_tri_closest_ring(mcl_tri_t* t, int K, int layout, const void* d_ring, void* d_out,
                         void* stream);
/* getFRNN<K,M>(r, ring, out) / getFRNN2<K,M> (Device_Triangulation.h:528-535,
 * 546-553; kernels.cuh:167-272): the vertices reachable from each vertex
 * through the Delaunay graph along vertices within r of it (FP32 squared
 * distance <= r*r), from the ring of mcl_tri_one_ring (K, layout).  Output
 * in the same layout with M slots per vertex: the count in slot 0, at most
 * M-1 ids.  MCL_ERR_CAPACITY (with *overflow = number of vertices) when a
 * walk needs more than 256 ids or 64 stacked vertices. */
int mcl_tri_frnn_graph(mcl_tri_t* t, float r, int K, int layout, const void* d_ring, int M, void* d_out,
                       int* overflow, void* stream);

/* =========================================================================
 * Uniform-grid FRNN  (GridCount2d, gridCount2d.h:23-48)
 * ========================================================================= */
#define MCL_GRID_ROWMAJOR 0    /* reference cell key (y*nx + x) % (nx*ny) */
#define MCL_GRID_MORTON 1      /* Morton-ordered cell key */
#define MCL_GRID_PERIODIC 2    /* flag: wrap cells, minimum-image distances */

/* GridCount2d(numP, min, cell_size, num_cells) (gridCount2d.h:50-71) */
int mcl_grid_create(int nump, float minx, float miny, float csx, float csy, int ncx, int ncy,
                    int flags, mcl_grid_t** out);
int mcl_grid_destroy(mcl_grid_t* g);
/* update(pos[, vel]) (gridCount2d.h:119-135, 174-260): counting sort by cell;
 * permutes the caller's device arrays in place (vel may be NULL).  Within a
 * cell, points keep their previous relative order (the reference's atomic
 * order is nondeterministic). */
int mcl_grid_update(mcl_grid_t* g, void* d_pos, void* d_vel, void* stream);
/* d_ind: new slot of old point i (gridCount2d.h:40, 146-153); perm: old id of slot s */
int mcl_grid_index_ptrs(mcl_grid_t* g, void** d_ind, void** d_perm, void** d_cell_start,
                        void** d_cell_count);
/* apply_f_frnn(f, pos, rad) (gridCount2d.h:262-302) as CSR over sorted slots:
 * for each slot i, every slot j != i in the 3x3 cells around i with
 * r2 = dx*dx + dy*dy <= rad*rad (FP32, no contraction).  d_offsets has
 * nump+1 ints, d_nbrs cap ints.  *total = pair count (MCL_ERR_CAPACITY if > cap).
 * With total == NULL the call is asynchronous; entries past cap are dropped
 * and d_offsets[nump] still holds the full count.  The count pass leaves each
 * slot's hits in the grid for the fill pass, so queries on one grid (this
 * call, mcl_tri_timestep's tail) must not overlap in time. */
int mcl_grid_frnn_csr(mcl_grid_t* g, const void* d_pos, float rad, void* d_offsets, void* d_nbrs,
                      int cap, long long* total, void* stream);
/* the counting sort of update() without touching the source: the sorted
 * positions stay in the grid (mcl_grid_sorted_ptr); asynchronous */
int mcl_grid_build(mcl_grid_t* g, const void* d_src_pos, void* stream);
void* mcl_grid_sorted_ptr(mcl_grid_t* g);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): mcl_grid_create + mcl_grid_build + mcl_grid_frnn_csr in one
 * call that allocates nothing.  The grid (geometry and flags as
 * mcl_grid_create) sorts the nump device points d_pos by cell into a
 * library-owned static device array; the result equals the handle path's:
 * d_offsets (nump + 1 ints) and d_nbrs (cap ints) as mcl_grid_frnn_csr over
 * the sorted points, d_perm (nump ints, may be NULL) the point index of each
 * sorted slot (mcl_grid_index_ptrs' perm).  At most 1,048,576 points and
 * 1,048,576 cell keys (cells, or the power-of-two Morton square with
 * MCL_GRID_MORTON): MCL_ERR_CAPACITY beyond.  The static array (about 30 MB)
 * is held by every GPU the library runs on until the process ends and is
 * shared by all calls on that GPU: calls that overlap on one GPU (other
 * streams, other host threads) race.  total == NULL: asynchronous; else as
 * mcl_grid_frnn_csr. */
int mcl_grid_frnn_csr_noalloc(const void* d_pos, int nump, float minx, float miny, float csx, float csy,
                              int ncx, int ncy, int flags, float rad, void* d_perm, void* d_offsets,
                              void* d_nbrs, int cap, long long* total, void* stream);
/* mean_num_particle_in_cell (gridCount2d.h:304-334) */
int mcl_grid_mean_occupancy(mcl_grid_t* g, float* out, void* stream);

/* --------------------------------------------------------------- options */
/* No counterpart in the reference (its step has no variants): choices of the
 * fast kinetic step (variant 0).
 * MCL_OPT_STAR_MIN_VERTICES: meshes with at least this many vertices (default
 *   65536) may restore the triangulation by per-vertex star recomputation
 *   (torus, and box meshes with their corners first); smaller ones, where
 *   the number of launches dominates, use the flip-based step.  0 = always.
 *   The same bound decides whether mcl_tri_timestep runs its grid part on a
 *   second stream beside the step (at least this many vertices).
 * MCL_OPT_STAR_ENABLE: 0 = never use the star recomputation; 1 = every step
 *   (after a step where it did not settle, skipped for 1, 2, 4 ... 64 steps);
 *   2 (default) = the path whose last step was faster, the slower one timed
 *   again after 64 x (slow / fast) steps, twice as late each time it stays
 *   slower (up to 4096 steps).
 * MCL_OPT_STAR_GIVE_UP: the star step falls back to the flip-based step at
 *   once when its first consistency check lists more than this percentage of
 *   the vertices (1..100, default 40; 100 = always run the retry passes).
 * MCL_OPT_GRAPH: 1 = the star step's head (move, first pass, checks and the
 *   retries before the first host read) and mcl_tri_timestep's tail are
 *   captured once as hipGraphs and replayed every step (the results are those
 *   of the eager launches); 0 (default) = eager launches.
 * Round-loop parameters of the flip-based step (the reference hard-codes its
 * launch parameters instead, e.g. blocksize = 128, Device_Triangulation.h:151):
 * MCL_OPT_TAIL_THRESHOLD: Lawson queue length below which one workgroup
 *   finishes the remaining rounds in one launch (default 8192).
 * MCL_OPT_REVERT_TAIL / MCL_OPT_RELOC_TAIL: the same for take-back passes and
 *   relocation rounds (default 1024 each).
 * MCL_OPT_BIG_GRID_MIN: queue length of big-exact-only Lawson cases from
 *   which they run on 256 workgroups before the tail (default 2048).
 * MCL_OPT_STAR_HEAD_PASSES: star passes enqueued (and, with MCL_OPT_GRAPH,
 *   captured) before the first host read (1..12, default 4).
 * Defaults are overridden at creation by the autotuner's table
 * (mcl_tuning_path). */
#define MCL_OPT_STAR_MIN_VERTICES 1
#define MCL_OPT_STAR_ENABLE 2
#define MCL_OPT_STAR_GIVE_UP 3
#define MCL_OPT_GRAPH 4
#define MCL_OPT_TAIL_THRESHOLD 5
#define MCL_OPT_REVERT_TAIL 6
#define MCL_OPT_RELOC_TAIL 7
#define MCL_OPT_BIG_GRID_MIN 8
#define MCL_OPT_STAR_HEAD_PASSES 9
int mcl_tri_set_option(mcl_tri_t* t, int option, long long value);
int mcl_tri_get_option(const mcl_tri_t* t, int option, long long* value);
/* GridCount2d's launch parameter (the reference's public blocksize = 64,
 * gridCount2d.h:35): MCL_GRID_OPT_BLOCK, threads per block of the grid's
 * sort and query kernels, 64, 128 or 256 (default 256, or the tuning table's). */
#define MCL_GRID_OPT_BLOCK 1
int mcl_grid_set_option(mcl_grid_t* g, int option, long long value);
int mcl_grid_get_option(const mcl_grid_t* g, int option, long long* value);
/* The autotuner's table in use: the file MCL_TUNING names (MCL_TUNING=off:
 * none), else tuning.json next to libmcleap_hip.so, written by
 * mcleap_amd/autotune.py; "" when no table was loaded.  Each entry (domain
 * torus / box / ensemble, log2 of the vertex count) sets the options above
 * and grid_block for meshes and grids of its class and size (within a factor
 * of 4). */
const char* mcl_tuning_path(void);

/* ------------------------------------------------------- kernel timers */
/* GPU time of the main kernels, from HIP event pairs recorded on the launch
 * stream around every launch while enabled (no synchronisation is added).
 * Triangulation kernels: 0 = first Lawson round over all edges,
 * 1 = orientation check of all triangles, 2 = move, 3 = relocation apply
 * (full-grid launches), 4 = take-back apply (full-grid launches),
 * 5 = star recomputation (first pass of the star step).
 * Grid kernels: 0 = FRNN count pass, 1 = FRNN fill pass, 2 = grid build. */
int mcl_tri_kernel_timer(mcl_tri_t* t, int enable);
int mcl_tri_kernel_timer_read(mcl_tri_t* t, int which, double* total_ms, long long* launches);
int mcl_grid_kernel_timer(mcl_grid_t* g, int enable);
int mcl_grid_kernel_timer_read(mcl_grid_t* g, int which, double* total_ms, long long* launches);

/* =========================================================================
 * Batched grids for an ensemble (no counterpart in the reference, where each
 * instance would hold its own GridCount2d, gridCount2d.h:23-48)
 * ========================================================================= */
typedef struct mcl_bgrid mcl_bgrid_t;
/* Instance i's points are entries vstart[i]..vstart[i+1]-1 of the position
 * array (mcl_hmesh_batch_offsets), all in [0,1)^2.  Its grid has
 * nc_i = max(3, floor(1/radius[i])) cells per side (clamped at the border)
 * and its query radius is radius[i]; MCL_ERR_INPUT when a cell is smaller
 * than the radius. */
int mcl_bgrid_create(int ninst, const int* vstart, const float* radius, mcl_bgrid_t** out);
int mcl_bgrid_destroy(mcl_bgrid_t* g);
/* points sorted (npoints = vstart[ninst] - vstart[0], first = vstart[0]),
 * total cells */
int mcl_bgrid_sizes(const mcl_bgrid_t* g, int* npoints, int* first, int* ncells);
/* counting sort of every instance's points by (instance, cell) in one pass
 * (GridCount2d::update, gridCount2d.h:194-258, for all grids at once); the
 * sorted positions stay in the grid.  Asynchronous on `stream`. */
int mcl_bgrid_build(mcl_bgrid_t* g, const void* d_pos, void* stream);
void* mcl_bgrid_sorted_ptr(mcl_bgrid_t* g);
/* d_ind[v - first] = sorted slot of point v, d_perm[s] = v - first of slot s,
 * d_cell_start = ncells + 1 slot offsets (instance i's cells after instance
 * i-1's, row-major) */
int mcl_bgrid_index_ptrs(mcl_bgrid_t* g, void** d_ind, void** d_perm, void** d_cell_start);
/* FRNN CSR over the sorted points (apply_f_frnn + SaveNeighborsFunctor,
 * gridCount2d.h:262-302, with each instance's own radius): d_offsets has
 * npoints + 1 entries, neighbours are sorted slots of the same instance.
 * total = NULL: asynchronous, writes past cap dropped; else as
 * mcl_grid_frnn_csr. */
int mcl_bgrid_frnn_csr(mcl_bgrid_t* g, void* d_offsets, void* d_nbrs, int cap, long long* total, void* stream);
/* HIP-event timers: 0 FRNN count, 1 FRNN fill, 2 build */
int mcl_bgrid_kernel_timer(mcl_bgrid_t* g, int enable);
int mcl_bgrid_kernel_timer_read(mcl_bgrid_t* g, int which, double* total_ms, long long* launches);
/* mcl_tri_timestep for an ensemble (no counterpart in the reference, whose
 * driver loop steps one triangulation, moving_device_mesh.cu:109-243): the
 * step, the CSR one-ring, and the batched grids' sort of the new positions
 * with their CSR FRNN query (as mcl_bgrid_build + mcl_bgrid_frnn_csr with
 * total = NULL) when bg is given; bg's points must be vertices of t. */
int mcl_tri_timestep_batch(mcl_tri_t* t, const mcl_step_params_t* p, void* d_ring_off, void* d_ring_nbrs,
                           int ring_cap, mcl_bgrid_t* bg, void* d_frnn_off, void* d_frnn_nbrs, int frnn_cap,
                           mcl_stats_t* st, void* stream);

/* ----------------------------------------------------------------- misc */
/* library build string and the device it runs on */
const char* mcl_version(void);
int mcl_device_info(char* name, int len, int* cu_count);

#ifdef __cplusplus
}
#endif
#endif /* MCLEAP_HIP_H */
