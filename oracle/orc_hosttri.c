/*
 * orc_hosttri.c -- restatement of the reference host triangulation
 * (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Follows Host_Triangulation.h:107-560 and Common_Triangulation.h:69-209.
 * Two predicate modes:
 *   ORC_PRED_REFERENCE  verbatim reference predicates (FP64 / FP32-sub with
 *                       EPS dead zone) and the reference's centroid walk;
 *   ORC_PRED_EXACT      exact orientation, exact incircle with the
 *                       symbolic-perturbation tie rule, a visibility walk,
 *                       and point-on-edge insertion (2->4 split), which the
 *                       reference leaves as a stub (Common_Triangulation.h:211).
 */
#include "oracle.h"
#include "orc_internal.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct orc_tri {
    float* pos; int npos, cpos;     /* xy pairs */
    orc_pt* ext;                    /* optional exact coordinates (torus build) */
    int* rank_id;                   /* optional: canonical vertex id for ext points */
    orc_vertex* v; int nv, cv;
    orc_halfedge* he; int nhe, che;
    orc_triangle* t; int nt, ct;
    int mode;
    int periodic;
};

#define GROW(arr, n, cap, type)                                             \
    do {                                                                    \
        if ((n) >= (cap)) {                                                 \
            (cap) = (cap) ? (cap) * 2 : 64;                                 \
            (arr) = (type*)realloc((arr), sizeof(type) * (size_t)(cap));    \
        }                                                                   \
    } while (0)

static void push_pos(orc_tri* T, float x, float y) {
    if (T->npos >= T->cpos) {
        T->cpos = T->cpos ? T->cpos * 2 : 64;
        T->pos = (float*)realloc(T->pos, sizeof(float) * 2 * (size_t)T->cpos);
    }
    T->pos[2 * T->npos] = x;
    T->pos[2 * T->npos + 1] = y;
    T->npos++;
}
static void push_v(orc_tri* T, orc_vertex v) { GROW(T->v, T->nv, T->cv, orc_vertex); T->v[T->nv++] = v; }
static void push_he(orc_tri* T, orc_halfedge h) { GROW(T->he, T->nhe, T->che, orc_halfedge); T->he[T->nhe++] = h; }
static void push_t(orc_tri* T, orc_triangle t) { GROW(T->t, T->nt, T->ct, orc_triangle); T->t[T->nt++] = t; }

/* --------------------------------------------------------- predicates */
static orc_pt P(const orc_tri* T, int v) {
    if (T->ext) return T->ext[v];
    orc_pt p = {T->pos[2 * v], T->pos[2 * v + 1], 0, 0};
    return p;
}
static void dv(const orc_tri* T, int v, double* out) {
    out[0] = T->pos[2 * v];
    out[1] = T->pos[2 * v + 1];
}
static int orient_v(const orc_tri* T, int a, int b, int c) {
    if (T->mode == ORC_PRED_REFERENCE)
        return orc_orient2d_ref(T->pos[2 * a], T->pos[2 * a + 1], T->pos[2 * b], T->pos[2 * b + 1],
                                T->pos[2 * c], T->pos[2 * c + 1]);
    if (T->periodic) {  /* unit torus, minimum-image differences */
        double pa[2], pb[2], pc[2];
        dv(T, a, pa); dv(T, b, pb); dv(T, c, pc);
        return orc_orient2d_exact(pa, pb, pc, 1);
    }
    orc_pt pa = P(T, a), pb = P(T, b), pc = P(T, c);
    return orc_orient_pt(&pa, &pb, &pc);
}
static int orient_vp(const orc_tri* T, int a, int b, const orc_pt* p) {
    if (T->mode == ORC_PRED_REFERENCE)
        return orc_orient2d_ref(T->pos[2 * a], T->pos[2 * a + 1], T->pos[2 * b], T->pos[2 * b + 1],
                                (float)p->x, (float)p->y);
    orc_pt pa = P(T, a), pb = P(T, b);
    return orc_orient_pt(&pa, &pb, p);
}
static int incircle_v(const orc_tri* T, int a, int b, int c, int d) {
    if (T->mode == ORC_PRED_REFERENCE)
        return orc_incircle_ref(T->pos[2 * a], T->pos[2 * a + 1], T->pos[2 * b], T->pos[2 * b + 1],
                                T->pos[2 * c], T->pos[2 * c + 1], T->pos[2 * d], T->pos[2 * d + 1]);
    if (T->periodic) {
        double pa[2], pb[2], pc[2], pd[2];
        dv(T, a, pa); dv(T, b, pb); dv(T, c, pc); dv(T, d, pd);
        return orc_incircle_sos(pa, pb, pc, pd, 1);
    }
    orc_pt pa = P(T, a), pb = P(T, b), pc = P(T, c), pd = P(T, d);
    return orc_incircle_sos_pt(&pa, &pb, &pc, &pd);
}

/* the imported mesh lives on the unit torus (no corners; minimum image) */
void orc_tri_set_periodic(orc_tri* T, int periodic) { T->periodic = periodic ? 1 : 0; }

/* ----------------------------------------------- Common_Triangulation.h */
/* f2to2 -- Common_Triangulation.h:69-127 */
static void f2to2(orc_triangle* m_t, orc_halfedge* m_he, orc_vertex* m_v, int he_index) {
    int v[4], he[6], t[2];
    he[0] = he_index;
    he[1] = he_index ^ 1;
    he[2] = m_he[he[0]].next;
    he[3] = m_he[he[2]].next;
    he[4] = m_he[he[1]].next;
    he[5] = m_he[he[4]].next;
    t[0] = m_he[he[0]].t;
    t[1] = m_he[he[1]].t;
    v[0] = m_he[he[0]].v;
    v[1] = m_he[he[1]].v;
    v[2] = m_he[he[3]].v;
    v[3] = m_he[he[5]].v;
    m_t[t[0]].he = he[0];
    m_t[t[1]].he = he[1];
    m_v[v[1]].he = he[2];
    m_v[v[0]].he = he[4];
    /* t0 = (v2 -> v3 -> v1) through he0, he5, he2 */
    m_he[he[0]].next = he[5]; m_he[he[5]].next = he[2]; m_he[he[2]].next = he[0];
    m_he[he[0]].v = v[2];     m_he[he[5]].v = v[3];     m_he[he[2]].v = v[1];
    m_he[he[0]].t = t[0];     m_he[he[5]].t = t[0];     m_he[he[2]].t = t[0];
    m_he[he[0]].op = v[1];    m_he[he[5]].op = v[2];    m_he[he[2]].op = v[3];
    /* t1 = (v3 -> v2 -> v0) through he1, he3, he4 */
    m_he[he[1]].next = he[3]; m_he[he[3]].next = he[4]; m_he[he[4]].next = he[1];
    m_he[he[1]].v = v[3];     m_he[he[3]].v = v[2];     m_he[he[4]].v = v[0];
    m_he[he[1]].t = t[1];     m_he[he[3]].t = t[1];     m_he[he[4]].t = t[1];
    m_he[he[1]].op = v[0];    m_he[he[3]].op = v[3];    m_he[he[4]].op = v[2];
}

typedef struct { int v_index, t0_index, t1_index, he0_index, he2_index, he4_index; } f3to1_info;

/* f1to3 -- Common_Triangulation.h:129-209 */
static void f1to3(orc_triangle* m_t, orc_halfedge* m_he, orc_vertex* m_v, f3to1_info fi, int t_index) {
    int he[9], t[3], v[4];
    t[0] = t_index; t[1] = fi.t0_index; t[2] = fi.t1_index;
    he[0] = m_t[t_index].he;
    he[1] = m_he[he[0]].next;
    he[2] = m_he[he[1]].next;
    he[8] = fi.he0_index; he[7] = fi.he0_index ^ 1;
    he[6] = fi.he2_index; he[5] = fi.he2_index ^ 1;
    he[4] = fi.he4_index; he[3] = fi.he4_index ^ 1;
    v[0] = m_he[he[0]].v; v[1] = m_he[he[1]].v; v[2] = m_he[he[2]].v; v[3] = fi.v_index;
    m_t[t[0]].he = he[0]; m_t[t[1]].he = he[1]; m_t[t[2]].he = he[2];
    m_v[v[3]].he = he[4];
    /* t0 */
    m_he[he[0]].next = he[3]; m_he[he[3]].next = he[5]; m_he[he[5]].next = he[0];
    m_he[he[0]].v = v[0];     m_he[he[3]].v = v[1];     m_he[he[5]].v = v[3];
    m_he[he[0]].t = t[0];     m_he[he[3]].t = t[0];     m_he[he[5]].t = t[0];
    m_he[he[0]].op = v[3];    m_he[he[3]].op = v[0];    m_he[he[5]].op = v[1];
    /* t1 */
    m_he[he[1]].next = he[7]; m_he[he[7]].next = he[4]; m_he[he[4]].next = he[1];
    m_he[he[1]].v = v[1];     m_he[he[7]].v = v[2];     m_he[he[4]].v = v[3];
    m_he[he[1]].t = t[1];     m_he[he[7]].t = t[1];     m_he[he[4]].t = t[1];
    m_he[he[1]].op = v[3];    m_he[he[7]].op = v[1];    m_he[he[4]].op = v[2];
    /* t2 */
    m_he[he[2]].next = he[6]; m_he[he[6]].next = he[8]; m_he[he[8]].next = he[2];
    m_he[he[2]].v = v[2];     m_he[he[6]].v = v[0];     m_he[he[8]].v = v[3];
    m_he[he[2]].t = t[2];     m_he[he[6]].t = t[2];     m_he[he[8]].t = t[2];
    m_he[he[2]].op = v[3];    m_he[he[6]].op = v[2];    m_he[he[8]].op = v[0];
}

/*
 * f2to4: insert vertex p on the interior edge he0 = (a -> b) shared by
 * t0 = (a, b, c) and t1 = (b, a, d).  Result: t0 = (a, p, c), tA = (p, b, c),
 * t1 = (p, a, d), tB = (b, p, d).  he0 becomes a -> p and he1 p -> a; the new
 * pairs are n0 = (p -> b | b -> p), n1 = (p -> c | c -> p), n2 = (p -> d | d -> p).
 */
static void f2to4(orc_triangle* m_t, orc_halfedge* m_he, orc_vertex* m_v, int p, int he0,
                  int tA, int tB, int n0, int n1, int n2) {
    int he1 = he0 ^ 1;
    int he2 = m_he[he0].next, he3 = m_he[he2].next; /* b->c, c->a */
    int he4 = m_he[he1].next, he5 = m_he[he4].next; /* a->d, d->b */
    int t0 = m_he[he0].t, t1 = m_he[he1].t;
    int a = m_he[he0].v, b = m_he[he1].v, c = m_he[he3].v, d = m_he[he5].v;
#define SET(h, nx, vv, tt, oo) do { m_he[h].next = nx; m_he[h].v = vv; m_he[h].t = tt; m_he[h].op = oo; } while (0)
    /* t0 = (a, p, c): he0 a->p, n1 p->c, he3 c->a */
    SET(he0, n1, a, t0, c); SET(n1, he3, p, t0, a); SET(he3, he0, c, t0, p);
    /* tA = (p, b, c): n0 p->b, he2 b->c, n1^1 c->p */
    SET(n0, he2, p, tA, c); SET(he2, n1 ^ 1, b, tA, p); SET(n1 ^ 1, n0, c, tA, b);
    /* t1 = (p, a, d): he1 p->a, he4 a->d, n2^1 d->p */
    SET(he1, he4, p, t1, d); SET(he4, n2 ^ 1, a, t1, p); SET(n2 ^ 1, he1, d, t1, a);
    /* tB = (b, p, d): n0^1 b->p, n2 p->d, he5 d->b */
    SET(n0 ^ 1, n2, b, tB, d); SET(n2, he5, p, tB, b); SET(he5, n0 ^ 1, d, tB, p);
#undef SET
    m_t[t0].he = he0; m_t[tA].he = n0; m_t[t1].he = he1; m_t[tB].he = n0 ^ 1;
    m_v[p].he = n0; m_v[a].he = he0; m_v[b].he = he2;
}

/* ---------------------------------------- HostTriangulation() :107-137 */
orc_tri* orc_tri_create(int pred_mode) {
    orc_tri* T = (orc_tri*)calloc(1, sizeof(orc_tri));
    T->mode = pred_mode;
    const float cmin = -10000000.f, cmax = 10000000.f; /* Common_Triangulation.h:12-13 */
    push_pos(T, cmin, cmin); push_pos(T, cmax, cmin); push_pos(T, cmax, cmax); push_pos(T, cmin, cmax);
    push_v(T, (orc_vertex){0, 0}); push_v(T, (orc_vertex){1, 2});
    push_v(T, (orc_vertex){2, 6}); push_v(T, (orc_vertex){3, 8});
    /* next, v, t, op */
    push_he(T, (orc_halfedge){2, 0, 0, 2}); push_he(T, (orc_halfedge){9, 1, -1, -1});
    push_he(T, (orc_halfedge){4, 1, 0, 0}); push_he(T, (orc_halfedge){1, 2, -1, -1});
    push_he(T, (orc_halfedge){0, 2, 0, 1}); push_he(T, (orc_halfedge){6, 0, 1, 3});
    push_he(T, (orc_halfedge){8, 2, 1, 0}); push_he(T, (orc_halfedge){3, 3, -1, -1});
    push_he(T, (orc_halfedge){5, 3, 1, 2}); push_he(T, (orc_halfedge){7, 0, -1, -1});
    push_t(T, (orc_triangle){4}); push_t(T, (orc_triangle){5});
    return T;
}

void orc_tri_destroy(orc_tri* T) {
    if (!T) return;
    free(T->pos); free(T->ext); free(T->rank_id); free(T->v); free(T->he); free(T->t);
    free(T);
}

int orc_tri_sizes(const orc_tri* T, int* o) {
    o[0] = T->npos; o[1] = T->nv; o[2] = T->nhe; o[3] = T->nt;
    return 0;
}

/* ------------------------------------------------- point location */
/* findContainingTriangleIndexWalking -- Host_Triangulation.h:148-186 */
static int walk_reference(const orc_tri* T, float px, float py, int t_guess) {
    int t_prev = -1;
    long long steps = 0, cap = 4LL * T->nt + 100;
    orc_pt p = {px, py, 0, 0};
keep_walking:
    if (++steps > cap) return -1;
    {
        int he0 = T->t[t_guess].he;
        int vv[3] = {T->he[he0].v, T->he[he0 ^ 1].v, T->he[he0].op};
        int he[3] = {he0, T->he[he0].next, 0};
        he[2] = T->he[he[1]].next;
        float pe[3][2];
        for (int i = 0; i < 3; i++) { pe[i][0] = T->pos[2 * vv[i]]; pe[i][1] = T->pos[2 * vv[i] + 1]; }
        /* centroid = (pe0 + pe1 + pe2) / 3.f in FP32 */
        float cx = ((pe[0][0] + pe[1][0]) + pe[2][0]) / 3.f;
        float cy = ((pe[0][1] + pe[1][1]) + pe[2][1]) / 3.f;
        for (int i = 0; i < 3; i++) {
            if (T->he[he[i] ^ 1].t != t_prev) {
                int j = (i + 1) % 3;
                int o1 = orc_orient2d_ref(cx, cy, px, py, pe[i][0], pe[i][1]);
                int o2 = orc_orient2d_ref(cx, cy, px, py, pe[j][0], pe[j][1]);
                int o3 = orc_orient2d_ref(pe[i][0], pe[i][1], pe[j][0], pe[j][1], cx, cy);
                int o4 = orc_orient2d_ref(pe[i][0], pe[i][1], pe[j][0], pe[j][1], px, py);
                if (o1 != o2 && o3 != o4) {
                    t_prev = t_guess;
                    t_guess = T->he[he[i] ^ 1].t;
                    if (t_guess < 0) return -1;
                    goto keep_walking;
                }
            }
        }
    }
    (void)p;
    return t_guess;
}

/* visibility walk with exact orientation.  Returns the triangle, and in
 * *on_edge the half-edge of that triangle p lies on (or -1). */
static int walk_exact(const orc_tri* T, const orc_pt* p, int t_guess, int* on_edge) {
    long long steps = 0, cap = 4LL * T->nt + 100;
    int t = t_guess;
    for (;;) {
        if (++steps > cap) break;
        int he0 = T->t[t].he, moved = 0;
        int hs[3] = {he0, T->he[he0].next, 0};
        hs[2] = T->he[hs[1]].next;
        int zero = -1;
        for (int k = 0; k < 3; k++) {
            int h = hs[(k + (int)(steps % 3)) % 3];
            int a = T->he[h].v, b = T->he[h ^ 1].v;
            int o = orient_vp(T, a, b, p);
            if (o < 0) {
                int nt = T->he[h ^ 1].t;
                if (nt < 0) return -1;
                t = nt; moved = 1; break;
            }
            if (o == 0) zero = h;
        }
        if (!moved) { *on_edge = zero; return t; }
    }
    /* fall back to a scan (only if the walk cycled) */
    for (int tt = 0; tt < T->nt; tt++) {
        int he0 = T->t[tt].he;
        int hs[3] = {he0, T->he[he0].next, 0};
        hs[2] = T->he[hs[1]].next;
        int ok = 1, zero = -1;
        for (int k = 0; k < 3 && ok; k++) {
            int o = orient_vp(T, T->he[hs[k]].v, T->he[hs[k] ^ 1].v, p);
            if (o < 0) ok = 0;
            else if (o == 0) zero = hs[k];
        }
        if (ok) { *on_edge = zero; return tt; }
    }
    return -1;
}

/* --------------------------------------------------------- delonizing */
/* delonizeEdge -- Host_Triangulation.h:323-346 */
static int delonize_edge(orc_tri* T, int he_index) {
    if (T->he[he_index].t == -1) return 0;
    if (T->he[he_index ^ 1].t == -1) return 0;
    int v0 = T->he[he_index].v, v1 = T->he[he_index ^ 1].op;
    int v2 = T->he[he_index ^ 1].v, v3 = T->he[he_index].op;
    if (incircle_v(T, v0, v1, v2, v3) > 0) {
        f2to2(T->t, T->he, T->v, he_index);
        return 1;
    }
    return 0;
}

/* delonizeVertex -- Host_Triangulation.h:368-395 */
static void delonize_vertex(orc_tri* T, int v_index) {
    int initial = T->v[v_index].he;
    int cur = initial;
    long long guard = 0;
    do {
        int link = T->he[cur].next;
        while (delonize_edge(T, link)) {
            link = T->he[cur].next;
            if (++guard > 100000000LL) return;
        }
        cur = T->he[T->he[cur].next].next ^ 1; /* NEXT_OUTGOING */
    } while (cur != initial);
}

static int dup_of(const orc_tri* T, int tri, const orc_pt* p) {
    int he0 = T->t[tri].he;
    int vs[3] = {T->he[he0].v, T->he[he0 ^ 1].v, T->he[he0].op};
    for (int k = 0; k < 3; k++) {
        orc_pt q = P(T, vs[k]);
        if (q.x == p->x && q.y == p->y && q.sx == p->sx && q.sy == p->sy) return 1;
    }
    return 0;
}

/* addDelaunayPoint -- Host_Triangulation.h:248-280 (exact mode adds the
 * point-on-edge case).  Returns 1 on success, 0 on failure. */
static int add_delaunay_point(orc_tri* T, float x, float y, const orc_pt* ext) {
    int on_edge = -1, t_index;
    orc_pt p = ext ? *ext : (orc_pt){x, y, 0, 0};
    if (T->mode == ORC_PRED_REFERENCE) t_index = walk_reference(T, x, y, T->nt - 1);
    else t_index = walk_exact(T, &p, T->nt - 1, &on_edge);
    if (t_index == -1) return 0;
    if (T->mode == ORC_PRED_EXACT && dup_of(T, t_index, &p)) return 0;
    push_pos(T, x, y);
    if (T->ext || ext) {
        int had = T->ext != NULL;
        T->ext = (orc_pt*)realloc(T->ext, sizeof(orc_pt) * (size_t)T->cpos);
        if (!had)
            for (int i = 0; i < T->npos - 1; i++)
                T->ext[i] = (orc_pt){T->pos[2 * i], T->pos[2 * i + 1], 0, 0};
        T->ext[T->npos - 1] = p;
    }
    push_v(T, (orc_vertex){T->npos - 1, 0});
    push_t(T, (orc_triangle){0});
    push_t(T, (orc_triangle){0});
    for (int k = 0; k < 6; k++) push_he(T, (orc_halfedge){0, 0, 0, 0});
    int vnew = T->nv - 1;
    if (on_edge < 0) {
        f3to1_info fi = {vnew, T->nt - 1, T->nt - 2, T->nhe - 1, T->nhe - 3, T->nhe - 5};
        f1to3(T->t, T->he, T->v, fi, t_index);
    } else {
        if (T->he[on_edge ^ 1].t < 0) return 0;
        f2to4(T->t, T->he, T->v, vnew, on_edge, T->nt - 2, T->nt - 1, T->nhe - 6, T->nhe - 4,
              T->nhe - 2);
    }
    delonize_vertex(T, vnew);
    return 1;
}

/* libmorton m2D_e_LUT of the coordinates truncated to unsigned ints
 * (Host_Triangulation.h:287-294); equivalent bit interleave */
static uint64_t morton_key(float x, float y) {
    uint32_t ux = (uint32_t)(int64_t)x, uy = (uint32_t)(int64_t)y;
    uint64_t k = 0;
    for (int b = 0; b < 32; b++) {
        k |= (uint64_t)((ux >> b) & 1u) << (2 * b);
        k |= (uint64_t)((uy >> b) & 1u) << (2 * b + 1);
    }
    return k;
}
typedef struct { uint64_t key; int idx; } kidx;
static int cmp_kidx(const void* a, const void* b) {
    const kidx* p = (const kidx*)a; const kidx* q = (const kidx*)b;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    return p->idx - q->idx;
}

int orc_tri_add_delaunay_points(orc_tri* T, const float* xy, int n, int sort_mode) {
    int* order = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) order[i] = i;
    if (sort_mode == 1) {
        kidx* k = (kidx*)malloc(sizeof(kidx) * (size_t)(n > 0 ? n : 1));
        for (int i = 0; i < n; i++) { k[i].key = morton_key(xy[2 * i], xy[2 * i + 1]); k[i].idx = i; }
        qsort(k, (size_t)n, sizeof(kidx), cmp_kidx);
        for (int i = 0; i < n; i++) order[i] = k[i].idx;
        free(k);
    }
    int done = 0;
    for (int i = 0; i < n; i++) {
        int j = order[i];
        if (!add_delaunay_point(T, xy[2 * j], xy[2 * j + 1], NULL)) { free(order); return -(1 + j); }
        done++;
    }
    free(order);
    return done;
}

/* used by the torus builder: insert a point with exact extended coordinates */
int orc_tri_add_ext_point(orc_tri* T, const orc_pt* p) {
    return add_delaunay_point(T, (float)p->x, (float)p->y, p);
}
const orc_halfedge* orc_tri_he(const orc_tri* T) { return T->he; }
const orc_triangle* orc_tri_t(const orc_tri* T) { return T->t; }
const orc_pt* orc_tri_ext(const orc_tri* T) { return T->ext; }

/* ----------------------------------------------------- import / export */
void orc_tri_export(const orc_tri* T, float* pos, int* v2, int* he4, int* t1) {
    if (pos) memcpy(pos, T->pos, sizeof(float) * 2 * (size_t)T->npos);
    if (v2) memcpy(v2, T->v, sizeof(orc_vertex) * (size_t)T->nv);
    if (he4) memcpy(he4, T->he, sizeof(orc_halfedge) * (size_t)T->nhe);
    if (t1) memcpy(t1, T->t, sizeof(orc_triangle) * (size_t)T->nt);
}

int orc_tri_import(orc_tri* T, const float* pos, int npos, const int* v2, int nv, const int* he4,
                   int nhe, const int* t1, int nt) {
    T->npos = T->nv = T->nhe = T->nt = 0;
    for (int i = 0; i < npos; i++) push_pos(T, pos[2 * i], pos[2 * i + 1]);
    for (int i = 0; i < nv; i++) push_v(T, (orc_vertex){v2[2 * i], v2[2 * i + 1]});
    for (int i = 0; i < nhe; i++)
        push_he(T, (orc_halfedge){he4[4 * i], he4[4 * i + 1], he4[4 * i + 2], he4[4 * i + 3]});
    for (int i = 0; i < nt; i++) push_t(T, (orc_triangle){t1[i]});
    free(T->ext); T->ext = NULL;
    return 0;
}

void orc_tri_set_positions(orc_tri* T, const float* pos, int npos) {
    int n = npos < T->npos ? npos : T->npos;
    memcpy(T->pos, pos, sizeof(float) * 2 * (size_t)n);
}

/* delonize -- Host_Triangulation.h:305-319 (restart the scan after every flip) */
int orc_tri_delonize(orc_tri* T, long long* flips) {
    int flipped = 0, ready = 0;
    long long f = 0;
    while (!ready) {
        ready = 1;
        for (int i = 0; i < T->nhe; i += 2) {
            if (delonize_edge(T, i)) { ready = 0; flipped = 1; f++; break; }
        }
        if (f > 50LL * T->nhe + 1000) break; /* inexact predicates can cycle */
    }
    if (flips) *flips = f;
    return flipped;
}

/* same predicates and flips, driven by a work stack */
int orc_tri_delonize_worklist(orc_tri* T, long long* flips) {
    int ne = T->nhe / 2;
    int* stack = (int*)malloc(sizeof(int) * (size_t)(ne + 8) * 4);
    char* inq = (char*)calloc((size_t)ne, 1);
    int top = 0;
    for (int e = ne - 1; e >= 0; e--) { stack[top++] = e; inq[e] = 1; }
    long long f = 0, cap = 50LL * T->nhe + 1000;
    while (top > 0 && f <= cap) {
        int e = stack[--top];
        inq[e] = 0;
        int h = 2 * e;
        if (delonize_edge(T, h)) {
            f++;
            int nb[4];
            nb[0] = T->he[h].next; nb[1] = T->he[nb[0]].next;
            nb[2] = T->he[h ^ 1].next; nb[3] = T->he[nb[2]].next;
            for (int k = 0; k < 4; k++) {
                int ee = nb[k] >> 1;
                if (!inq[ee]) {
                    if (top >= (ne + 8) * 4) continue;
                    stack[top++] = ee; inq[ee] = 1;
                }
            }
        }
    }
    free(stack); free(inq);
    if (flips) *flips = f;
    return f > 0;
}

/* ------------------------------------------------------------ untangle */
/* isCCW / isCreased / isInside / isInsideInverted -- Helpers_Triangulation.h:163-216 */
static int tri_verts(const orc_tri* T, int t, int* vs) {
    int he = T->t[t].he;
    vs[0] = T->he[he].v; vs[1] = T->he[he ^ 1].v; vs[2] = T->he[he].op;
    return 0;
}
static int is_ccw(const orc_tri* T, int t) {
    int vs[3]; tri_verts(T, t, vs);
    return orient_v(T, vs[0], vs[1], vs[2]) >= 0;
}
static int is_creased(const orc_tri* T, int he) {
    if (T->he[he].t == -1 || T->he[he ^ 1].t == -1) return 0;
    return is_ccw(T, T->he[he ^ 1].t) != is_ccw(T, T->he[he].t);
}
static int is_inverted_edge(const orc_tri* T, int he) {
    if (T->he[he].t == -1 || T->he[he ^ 1].t == -1) return 0;
    return !is_ccw(T, T->he[he ^ 1].t) && !is_ccw(T, T->he[he].t);
}
static int is_inside(const orc_tri* T, int pv, int t) {
    int vs[3]; tri_verts(T, t, vs);
    for (int i = 0; i < 3; i++)
        if (orient_v(T, vs[i], vs[(i + 1) % 3], pv) <= 0) return 0;
    return 1;
}
static int is_inside_inverted(const orc_tri* T, int pv, int t) {
    int vs[3]; tri_verts(T, t, vs);
    for (int i = 0; i < 3; i++)
        if (orient_v(T, vs[i], vs[(i + 1) % 3], pv) >= 0) return 0;
    return 1;
}
/* untangleEdge -- Host_Triangulation.h:425-471 (B and C steps) */
static int untangle_edge(orc_tri* T, int he_index) {
    if (!is_creased(T, he_index)) return 0;
    int up = he_index, inv = he_index ^ 1;
    if (!is_ccw(T, T->he[up].t)) { int x = up; up = inv; inv = x; }
    int op_up = T->he[T->he[T->he[up].next].next].v;
    int op_inv = T->he[T->he[T->he[inv].next].next].v;
    int t_up = T->he[up].t, t_inv = T->he[inv].t;
    if (is_inside(T, op_inv, t_up)) { f2to2(T->t, T->he, T->v, he_index); return 1; }
    if (is_inside_inverted(T, op_up, t_inv)) { f2to2(T->t, T->he, T->v, he_index); return 1; }
    return 0;
}
/* untangle -- Host_Triangulation.h:417-423 (one pass over even half-edges) */
int orc_tri_untangle(orc_tri* T) {
    int n = 0;
    for (int i = 0; i < T->nhe; i += 2) n += untangle_edge(T, i);
    return n;
}

/* movePoint -- Host_Triangulation.h:403-406 */
int orc_tri_move_point(orc_tri* T, int i, float dx, float dy) {
    T->pos[2 * i] += dx;
    T->pos[2 * i + 1] += dy;
    return 1;
}

/* moving_device_mesh.cu:212-242 diagnostics, plus the number of triangles
 * that are not strictly CCW (exact) */
void orc_tri_diagnostics(const orc_tri* T, long long* out) {
    long long nd = 0, cr = 0, iv = 0, nccw = 0;
    for (int i = 0; i < T->nhe / 2; i++) {
        int h = 2 * i;
        if (T->he[h].t == -1 || T->he[h ^ 1].t == -1) continue;
        int v0 = T->he[h].v, v1 = T->he[h ^ 1].op, v2 = T->he[h ^ 1].v, v3 = T->he[h].op;
        if (incircle_v(T, v0, v1, v2, v3) > 0) nd++;
        if (is_creased(T, h)) cr++;
        if (is_inverted_edge(T, h)) iv++;
    }
    for (int t = 0; t < T->nt; t++) {
        int vs[3]; tri_verts(T, t, vs);
        if (T->periodic) {
            double a[2], b[2], c[2];
            dv(T, vs[0], a); dv(T, vs[1], b); dv(T, vs[2], c);
            if (orc_orient2d_exact(a, b, c, 1) <= 0) nccw++;
            continue;
        }
        orc_pt a = P(T, vs[0]), b = P(T, vs[1]), c = P(T, vs[2]);
        if (orc_orient_pt(&a, &b, &c) <= 0) nccw++;
    }
    out[0] = nd; out[1] = cr; out[2] = iv; out[3] = nccw;
}

/* ------------------------------------------------------------ queries */
/* oneRing -- Host_Triangulation.h:492-519 */
int orc_tri_one_ring(const orc_tri* T, int v_index, int* out, int cap) {
    int n = 0;
    int initial = T->v[v_index].he;
    if (T->he[initial].t == -1) initial ^= 1;
    int cur = T->he[T->he[initial].next].next ^ 1; /* NEXT_OUTGOING */
    int adj = T->he[initial ^ 1].v;
    if (n < cap) out[n] = adj;
    n++;
    while (cur != initial) {
        if (T->he[cur].t == -1) break;
        adj = T->he[cur ^ 1].v;
        if (n < cap) out[n] = adj;
        n++;
        cur = T->he[T->he[cur].next].next ^ 1;
    }
    cur = T->he[initial ^ 1].next; /* PREV_OUTGOING */
    while (cur != initial) {
        if (T->he[cur].t == -1) break;
        adj = T->he[cur ^ 1].v;
        if (n < cap) out[n] = adj;
        n++;
        cur = T->he[cur ^ 1].next;
    }
    return n;
}

/* getFRNN -- Host_Triangulation.h:525-560 */
int orc_tri_get_frnn(const orc_tri* T, int v_index, float r, int* out, int cap) {
    char* visited = (char*)calloc((size_t)T->npos, 1);
    int* stack = (int*)malloc(sizeof(int) * (size_t)T->npos + 16);
    int top = 0, n = 0;
    float cx = T->pos[2 * v_index], cy = T->pos[2 * v_index + 1];
    float r2 = r * r;
    visited[v_index] = 1;
    stack[top++] = v_index;
    int ring[4096];
    while (top > 0) {
        int s = stack[--top];
        int m = orc_tri_one_ring(T, s, ring, 4096);
        if (m > 4096) m = 4096;
        for (int k = 0; k < m; k++) {
            int adj = ring[k];
            if (!visited[adj]) {
                visited[adj] = 1;
                float dx = cx - T->pos[2 * adj], dy = cy - T->pos[2 * adj + 1];
                float d2 = dx * dx + dy * dy;
                if (d2 <= r2) {
                    if (n < cap) out[n] = adj;
                    n++;
                    stack[top++] = adj;
                }
            }
        }
    }
    free(visited); free(stack);
    return n;
}
