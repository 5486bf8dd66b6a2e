/*
 * oracle.h -- CPU restatement of mcleap's kinetic Delaunay path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library (mcleap_amd/libmcleap_hip.so) never links or calls it.
 *
 * Contents and the reference code each part restates:
 *   predicates  -- Helpers_Triangulation.h:86-151 (verbatim FP64 / FP32-sub
 *                  forms) and exact expansion arithmetic written from
 *                  Shewchuk, CMU-CS-96-140 (cited at experiments/predicates.c:26-32)
 *   topology    -- Common_Triangulation.h:42-209 (Vertex/HalfEdge/Triangle,
 *                  f2to2, f1to3) plus f2to4 (a stub at :211 in the reference)
 *   host path   -- Host_Triangulation.h:107-560 (construction, delonize,
 *                  untangle, oneRing, getFRNN)
 *   grid FRNN   -- gridCount2d.h:89-172, 262-295
 *   rng         -- std::mt19937 + libstdc++ uniform_real_distribution, used
 *                  by the reference tests (tests/grid2d_frnn.cu:211-218,
 *                  tests/test_frnn.cu:21-28)
 */
#pragma once
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- records */
typedef struct { int pos_index, he; } orc_vertex;     /* Common_Triangulation.h:42-45 */
typedef struct { int next, v, t, op; } orc_halfedge;  /* Common_Triangulation.h:47-53 */
typedef struct { int he; } orc_triangle;              /* Common_Triangulation.h:55-57 */

/* ------------------------------------------------------------- predicates */
/* verbatim reference forms (sign with the EPS dead zone) */
int orc_orient2d_ref(float ax, float ay, float bx, float by, float cx, float cy);
int orc_incircle_ref(float ax, float ay, float bx, float by, float cx, float cy,
                     float dx, float dy);
/* exact signs on double inputs (floats are passed widened). periodic != 0:
 * coordinates live on the unit torus and every difference is taken as the
 * minimum image. */
int orc_orient2d_exact(const double* a, const double* b, const double* c, int periodic);
int orc_incircle_exact(const double* a, const double* b, const double* c,
                       const double* d, int periodic);
/* incircle with the symbolic-perturbation tie rule: never returns 0.
 * Lifts are perturbed by infinitesimals ordered by the lexicographic (x, y)
 * order of the points' canonical coordinates (largest point = largest
 * perturbation), so ties resolve the same way for any vertex numbering. */
int orc_incircle_sos(const double* a, const double* b, const double* c,
                     const double* d, int periodic);
/* counters of how predicate calls were decided (filter / exact) */
void orc_pred_stats(long long* out4);
void orc_pred_stats_reset(void);

/* exact expansion primitives exported for tests */
int orc_expansion_sign(int n, const double* e);
int orc_exp_product(int elen, const double* e, int flen, const double* f, double* h);

/* -------------------------------------------------------- triangulation */
typedef struct orc_tri orc_tri;
enum { ORC_PRED_REFERENCE = 0, ORC_PRED_EXACT = 1 };

/* HostTriangulation() -- Host_Triangulation.h:107-137 (4 corners at +-1e7) */
orc_tri* orc_tri_create(int pred_mode);
void orc_tri_destroy(orc_tri* t);
int orc_tri_sizes(const orc_tri* t, int* out4); /* pos, v, he, t */
/* addDelaunayPoints -- Host_Triangulation.h:296-300.  sort_mode 0 keeps the
 * given order (vertex id = 4 + index); 1 sorts by the reference's Morton key
 * of the coordinates truncated to integers (stable).  Returns the number of
 * points inserted, or -(1+i) if point i could not be inserted (duplicate). */
int orc_tri_add_delaunay_points(orc_tri* t, const float* xy, int n, int sort_mode);
/* copy out / load the four reference arrays (AoS as in the reference) */
void orc_tri_export(const orc_tri* t, float* pos, int* v2, int* he4, int* t1);
int orc_tri_import(orc_tri* t, const float* pos, int npos, const int* v2, int nv,
                   const int* he4, int nhe, const int* t1, int nt);
void orc_tri_set_positions(orc_tri* t, const float* pos, int npos);
/* Host_Triangulation.h:305-319: rescans from edge 0 after every flip */
int orc_tri_delonize(orc_tri* t, long long* flips);
/* same predicates and flips with a work stack (labelled restatement for 1M) */
int orc_tri_delonize_worklist(orc_tri* t, long long* flips);
/* Host_Triangulation.h:417-471: one pass of B/C steps; returns #flips */
int orc_tri_untangle(orc_tri* t);
int orc_tri_move_point(orc_tri* t, int i, float dx, float dy);
/* the mesh lives on the unit torus: predicates use minimum-image differences */
void orc_tri_set_periodic(orc_tri* t, int periodic);
/* counts per Host diagnostics (moving_device_mesh.cu:212-242) with the
 * current predicate mode: [non-Delaunay edges, creased, inverted edges,
 * non-CCW triangles] */
void orc_tri_diagnostics(const orc_tri* t, long long* out4);
/* oneRing -- Host_Triangulation.h:492-519 (CCW forward walk, then backward) */
int orc_tri_one_ring(const orc_tri* t, int v, int* out, int cap);
/* getFRNN -- Host_Triangulation.h:525-560 (DFS over one-rings) */
int orc_tri_get_frnn(const orc_tri* t, int v, float r, int* out, int cap);

/* structural invariants of a half-edge mesh given as SoA or AoS arrays.
 * Returns 0 if all hold, else an error code; msg gets a description. */
int orc_check_mesh(const float* pos, int npos, const int* he4, int nhe,
                   const int* t1, int nt, const int* v2, int nv, int periodic,
                   int require_ccw, char* msg, int msglen);
/* number of interior edges that fail the exact (SoS) Delaunay test */
long long orc_count_non_delaunay(const float* pos, const int* he4, int nhe, int periodic);
/* canonical triangle list: each row = 3 vertex ids rotated so the smallest
 * id comes first (orientation kept), rows sorted.  Returns #triangles. */
int orc_canonical_triangles(const int* he4, int nhe, const int* t1, int nt, int* out3);

/* from-scratch periodic Delaunay triangulation on the unit torus (no corner
 * vertices): points replicated across the boundary within `margin`, planar
 * exact DT, then one triangle kept per orbit.  Outputs (caller-allocated)
 * tri3 = 3 vertex ids (CCW) per triangle, 2n triangles.  Returns #triangles
 * or a negative error. */
int orc_torus_delaunay(const float* xy, int n, float margin, int* tri3);

/* ----------------------------------------------------------- grid FRNN */
/* gridCount2d semantics: cell (int)((p-min)/cs) % nc, row-major key
 * (y*nx+x)%tot, 3x3 query with the reference bound test, FP32 distance
 * r2 = dx*dx + dy*dy (no contraction), accept r2 <= rad*rad, j != i.
 * Outputs per point (input order): counts[i] and neighbours as ORIGINAL ids
 * in nbr[i*cap + k] (first cap kept).  periodic != 0: wrap cells and use
 * minimum-image differences (the torus grid has no reference counterpart). */
int orc_grid_frnn(const float* xy, int n, float minx, float miny, float csx,
                  float csy, int ncx, int ncy, float rad, int periodic,
                  int* counts, int* nbr, int cap);
/* brute force O(n^2) with the same FP32 predicate */
int orc_brute_frnn(const float* xy, int n, float rad, int periodic, int* counts,
                   int* nbr, int cap);
/* cell keys and the stable counting-sort permutation (sorted slot -> original
 * id) for the reference row-major key */
int orc_grid_sort(const float* xy, int n, float minx, float miny, float csx,
                  float csy, int ncx, int ncy, int* hash, int* perm, int* cell_start);

/* ---------------------------------------------------------------- rng */
typedef struct { uint32_t mt[624]; int idx; } orc_mt19937;
void orc_mt_seed(orc_mt19937* g, uint32_t seed);
uint32_t orc_mt_next(orc_mt19937* g);
float orc_uniform_float(orc_mt19937* g, float a, float b);   /* libstdc++ */
double orc_uniform_double(orc_mt19937* g, double a, double b);
/* fill out[0..n) with uniform_real_distribution draws (float or double) */
void orc_mt_uniform_floats(uint32_t seed, float a, float b, int n, float* out);
void orc_mt_uniform_doubles(uint32_t seed, double a, double b, int n, double* out);
void orc_mt_raw(uint32_t seed, int n, uint32_t* out);

#ifdef __cplusplus
}
#endif
