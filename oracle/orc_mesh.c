/*
 * orc_mesh.c -- mesh invariants, canonical triangle sets and the periodic
 * from-scratch triangulation (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Invariants follow the half-edge conventions of Common_Triangulation.h:42-57
 * (twin = i^1, HalfEdge{next, v, t, op}, t = -1 on the outer face) and the
 * outer CW cycle built at Host_Triangulation.h:119-133.
 */
#include "oracle.h"
#include "orc_internal.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

orc_tri* orc_tri_create(int pred_mode);
int orc_tri_add_ext_point(orc_tri* T, const orc_pt* p);
const orc_halfedge* orc_tri_he(const orc_tri* T);
const orc_triangle* orc_tri_t(const orc_tri* T);
const orc_pt* orc_tri_ext(const orc_tri* T);

#define NXT(h) he4[4 * (h)]
#define ORG(h) he4[4 * (h) + 1]
#define TRI(h) he4[4 * (h) + 2]
#define OPP(h) he4[4 * (h) + 3]

static int fail(char* msg, int len, int code, const char* s, long a, long b) {
    if (msg && len > 0) snprintf(msg, (size_t)len, "%s (%ld, %ld)", s, a, b);
    return code;
}

static orc_pt pt_of(const float* pos, int v) {
    orc_pt p = {pos[2 * v], pos[2 * v + 1], 0, 0};
    return p;
}

static int orient_ids(const float* pos, int a, int b, int c, int periodic) {
    double pa[2] = {pos[2 * a], pos[2 * a + 1]}, pb[2] = {pos[2 * b], pos[2 * b + 1]},
           pc[2] = {pos[2 * c], pos[2 * c + 1]};
    return orc_orient2d_exact(pa, pb, pc, periodic);
}

int orc_check_mesh(const float* pos, int npos, const int* he4, int nhe, const int* t1, int nt,
                   const int* v2, int nv, int periodic, int require_ccw, char* msg, int msglen) {
    if (nhe % 2) return fail(msg, msglen, 1, "odd half-edge count", nhe, 0);
    for (int h = 0; h < nhe; h++) {
        int n = NXT(h), v = ORG(h), t = TRI(h);
        if (n < 0 || n >= nhe) return fail(msg, msglen, 2, "next out of range", h, n);
        if (v < 0 || v >= npos) return fail(msg, msglen, 3, "origin out of range", h, v);
        if (t < -1 || t >= nt) return fail(msg, msglen, 4, "triangle out of range", h, t);
        /* dest(h) == origin(next(h)) == origin(twin(h)) */
        if (ORG(h ^ 1) != ORG(n)) return fail(msg, msglen, 5, "twin/next origin mismatch", h, n);
        if (t >= 0) {
            int n2 = NXT(n);
            if (NXT(n2) != h) return fail(msg, msglen, 6, "face is not a 3-cycle", h, t);
            if (TRI(n) != t || TRI(n2) != t) return fail(msg, msglen, 7, "face ids differ on cycle", h, t);
            if (OPP(h) != ORG(n2)) return fail(msg, msglen, 8, "op != origin(next(next))", h, OPP(h));
        }
        if (ORG(h) == ORG(h ^ 1)) return fail(msg, msglen, 9, "degenerate edge", h, ORG(h));
    }
    for (int t = 0; t < nt; t++) {
        int h = t1[t];
        if (h < 0 || h >= nhe || TRI(h) != t) return fail(msg, msglen, 10, "t.he not in triangle", t, h);
        if (require_ccw) {
            int a = ORG(h), b = ORG(NXT(h)), c = ORG(NXT(NXT(h)));
            if (orient_ids(pos, a, b, c, periodic) <= 0)
                return fail(msg, msglen, 11, "triangle not strictly CCW", t, h);
        }
    }
    /* vertex rings: v.he originates at v and the rotation visits every
     * outgoing half-edge exactly once */
    int* outdeg = (int*)calloc((size_t)npos, sizeof(int));
    for (int h = 0; h < nhe; h++) outdeg[ORG(h)]++;
    int rc = 0;
    for (int v = 0; v < nv && !rc; v++) {
        int h0 = v2[2 * v + 1];
        if (h0 < 0 || h0 >= nhe || ORG(h0) != v) { rc = fail(msg, msglen, 12, "v.he does not start at v", v, h0); break; }
        int h = h0, k = 0;
        do {
            h = NXT(h ^ 1);
            if (++k > outdeg[v]) break;
        } while (h != h0);
        if (k != outdeg[v]) rc = fail(msg, msglen, 13, "vertex rotation != out-degree", v, k);
    }
    free(outdeg);
    if (rc) return rc;
    /* Euler characteristic: faces = triangles (+1 outer face in the plane) */
    long chi = (long)nv - nhe / 2 + nt + (periodic ? 0 : 1);
    if (chi != (periodic ? 0 : 2)) return fail(msg, msglen, 14, "Euler characteristic", chi, nv);
    return 0;
}

long long orc_count_non_delaunay(const float* pos, const int* he4, int nhe, int periodic) {
    long long bad = 0;
    for (int h = 0; h < nhe; h += 2) {
        if (TRI(h) < 0 || TRI(h ^ 1) < 0) continue;
        int v0 = ORG(h), v1 = OPP(h ^ 1), v2 = ORG(h ^ 1), v3 = OPP(h);
        double a[2] = {pos[2 * v0], pos[2 * v0 + 1]}, b[2] = {pos[2 * v1], pos[2 * v1 + 1]},
               c[2] = {pos[2 * v2], pos[2 * v2 + 1]}, d[2] = {pos[2 * v3], pos[2 * v3 + 1]};
        if (orc_incircle_sos(a, b, c, d, periodic) > 0) bad++;
    }
    return bad;
}

static int cmp3(const void* a, const void* b) {
    const int* p = (const int*)a; const int* q = (const int*)b;
    for (int i = 0; i < 3; i++) if (p[i] != q[i]) return p[i] < q[i] ? -1 : 1;
    return 0;
}
static void rot_min(int* r) {
    while (r[0] > r[1] || r[0] > r[2]) { int x = r[0]; r[0] = r[1]; r[1] = r[2]; r[2] = x; }
}

int orc_canonical_triangles(const int* he4, int nhe, const int* t1, int nt, int* out3) {
    (void)nhe;
    for (int t = 0; t < nt; t++) {
        int h = t1[t];
        int* r = out3 + 3 * t;
        r[0] = ORG(h); r[1] = ORG(NXT(h)); r[2] = ORG(NXT(NXT(h)));
        rot_min(r);
    }
    qsort(out3, (size_t)nt, sizeof(int) * 3, cmp3);
    return nt;
}

/*
 * Periodic Delaunay triangulation of n points on the unit torus.  Copies of
 * the points translated by the 8 lattice neighbours are added when they fall
 * within `margin` of the unit square; the planar exact DT of all of them is
 * built; a torus triangle is kept from the copy whose smallest canonical
 * vertex id is the untranslated one (each orbit has exactly one such copy
 * among triangles whose extent is below the margin).  The perturbation ranks
 * copies by canonical coordinates, matching a min-image predicate.
 */
int orc_torus_delaunay(const float* xy, int n, float margin, int* tri3) {
    orc_tri* T = orc_tri_create(ORC_PRED_EXACT);
    int cap = n * 9 + 16, m = 0;
    orc_pt* pts = (orc_pt*)malloc(sizeof(orc_pt) * (size_t)cap);
    int* vid = (int*)malloc(sizeof(int) * (size_t)cap);
    for (int i = 0; i < n; i++) { pts[m] = (orc_pt){xy[2 * i], xy[2 * i + 1], 0, 0}; vid[m++] = i; }
    for (int sx = -1; sx <= 1; sx++)
        for (int sy = -1; sy <= 1; sy++) {
            if (!sx && !sy) continue;
            for (int i = 0; i < n; i++) {
                double x = (double)xy[2 * i] + sx, y = (double)xy[2 * i + 1] + sy;
                if (x < -margin || x >= 1.0 + margin || y < -margin || y >= 1.0 + margin) continue;
                pts[m] = (orc_pt){xy[2 * i], xy[2 * i + 1], sx, sy};
                vid[m++] = i;
            }
        }
    /* insertion order: a coarse grid sweep keeps the walks short */
    int* order = (int*)malloc(sizeof(int) * (size_t)m);
    {
        int G = 1; while ((long)G * G * 4 < m) G *= 2;
        long long* key = (long long*)malloc(sizeof(long long) * (size_t)m);
        for (int k = 0; k < m; k++) {
            double x = (pts[k].x + pts[k].sx + margin) / (1.0 + 2 * margin);
            double y = (pts[k].y + pts[k].sy + margin) / (1.0 + 2 * margin);
            int gx = (int)(x * G), gy = (int)(y * G);
            if (gx < 0) gx = 0; if (gx >= G) gx = G - 1;
            if (gy < 0) gy = 0; if (gy >= G) gy = G - 1;
            int row = gy, col = (gy & 1) ? (G - 1 - gx) : gx; /* serpentine */
            key[k] = ((long long)(row * G + col) << 32) | k;
        }
        for (int k = 0; k < m; k++) order[k] = k;
        /* simple sort on keys */
        long long* tmpk = key;
        for (int k = 0; k < m; k++) tmpk[k] = key[k];
        int cmpll(const void* a, const void* b);
        qsort(tmpk, (size_t)m, sizeof(long long), cmpll);
        for (int k = 0; k < m; k++) order[k] = (int)(tmpk[k] & 0xffffffffLL);
        free(key);
    }
    int* id_of_vertex = (int*)malloc(sizeof(int) * (size_t)(m + 4));
    int* pid_of_vertex = (int*)malloc(sizeof(int) * (size_t)(m + 4));
    for (int k = 0; k < 4; k++) id_of_vertex[k] = -1, pid_of_vertex[k] = -1;
    int nvtx = 4;
    for (int k = 0; k < m; k++) {
        int j = order[k];
        if (!orc_tri_add_ext_point(T, &pts[j])) { free(pts); free(vid); free(order); free(id_of_vertex); free(pid_of_vertex); orc_tri_destroy(T); return -2; }
        id_of_vertex[nvtx] = vid[j];
        pid_of_vertex[nvtx] = j;
        nvtx++;
    }
    int sz[4];
    orc_tri_sizes(T, sz);
    const orc_halfedge* he = orc_tri_he(T);
    const orc_triangle* tt = orc_tri_t(T);
    int out = 0;
    for (int t = 0; t < sz[3]; t++) {
        int h = tt[t].he;
        int vs[3] = {he[h].v, he[he[h].next].v, he[he[he[h].next].next].v};
        int bad = 0, kmin = 0;
        for (int q = 0; q < 3; q++) if (vs[q] < 4) bad = 1;
        if (bad) continue;
        int ids[3];
        for (int q = 0; q < 3; q++) ids[q] = id_of_vertex[vs[q]];
        for (int q = 1; q < 3; q++) if (ids[q] < ids[kmin]) kmin = q;
        const orc_pt* pm = &pts[pid_of_vertex[vs[kmin]]];
        if (pm->sx != 0 || pm->sy != 0) continue;
        /* all other vertices must be within the margin band (kept triangle
         * lies near the base cell); the min-id vertex is a base copy */
        if (out >= 2 * n) { out = -3; break; }
        tri3[3 * out] = ids[0]; tri3[3 * out + 1] = ids[1]; tri3[3 * out + 2] = ids[2];
        out++;
    }
    free(pts); free(vid); free(order); free(id_of_vertex); free(pid_of_vertex);
    orc_tri_destroy(T);
    return out;
}

int cmpll(const void* a, const void* b) {
    long long x = *(const long long*)a, y = *(const long long*)b;
    return x < y ? -1 : (x > y);
}
