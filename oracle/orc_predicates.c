/*
 * orc_predicates.c -- oracle predicates (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * 1. Verbatim restatement of the reference predicates:
 *      orient2d  Helpers_Triangulation.h:86-122  (FP64 of FP32->FP64 differences,
 *                                                 |det| < EPS -> 0)
 *      inCircle  Helpers_Triangulation.h:136-151 (FP32 differences widened to
 *                                                 FP64, FP64 3x3 determinant)
 *    with EPS = 1e-17f (Common_Triangulation.h:14).
 * 2. Exact signs by expansion arithmetic, written from the algorithms in
 *    J. R. Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
 *    Robust Geometric Predicates", CMU-CS-96-140 (1996): Two-Sum, Two-Diff,
 *    Two-Product (here via fma), Grow-Expansion (Thm 10), Expansion-Sum as
 *    repeated Grow-Expansion (Thm 12), Scale-Expansion (Thm 19).  A point is
 *    (x + sx, y + sy) with double x, y and integer lattice shifts sx, sy, so
 *    periodic copies are represented exactly; every coordinate difference is
 *    formed exactly as an expansion.  A floating-point filter with 4x
 *    Shewchuk's error bound A decides clear cases first.
 * 3. Symbolic perturbation for exact cocircularity (orc_incircle_sos).
 *
 * Must be compiled without FP contraction and without fast-math.
 */
#include "oracle.h"
#include "orc_internal.h"
#include <math.h>
#include <string.h>

static long long g_stats[4]; /* orient filter, orient exact, incircle filter, incircle exact */
void orc_pred_stats(long long* out4) { memcpy(out4, g_stats, sizeof g_stats); }
void orc_pred_stats_reset(void) { memset(g_stats, 0, sizeof g_stats); }

/* ------------------------------------------------ reference verbatim forms */
static const float REF_EPS = 0.00000000000000001f; /* Common_Triangulation.h:14 */

int orc_orient2d_ref(float ax, float ay, float bx, float by, float cx, float cy) {
    double l = ((double)bx - (double)ax) * ((double)cy - (double)ay);
    double r = ((double)by - (double)ay) * ((double)cx - (double)ax);
    double det = l - r;
    return fabs(det) < REF_EPS ? 0 : (det > 0 ? 1 : -1);
}

int orc_incircle_ref(float ax, float ay, float bx, float by, float cx, float cy,
                     float dx, float dy) {
    /* the six differences are glm::vec2 (FP32) subtractions */
    const float f00 = ax - dx, f01 = ay - dy, f10 = bx - dx, f11 = by - dy;
    const float f20 = cx - dx, f21 = cy - dy;
    const double a00 = f00, a01 = f01, a10 = f10, a11 = f11, a20 = f20, a21 = f21;
    const double a02 = a00 * a00 + a01 * a01;
    const double a12 = a10 * a10 + a11 * a11;
    const double a22 = a20 * a20 + a21 * a21;
    double det = (a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) +
                  a02 * (a10 * a21 - a11 * a20));
    return fabs(det) < REF_EPS ? 0 : (det > 0 ? 1 : -1);
}

/* ------------------------------------------------- error-free transforms */
static inline void two_sum(double a, double b, double* s, double* e) {
    double x = a + b;
    double bv = x - a;
    double av = x - bv;
    double br = b - bv;
    double ar = a - av;
    *s = x;
    *e = ar + br;
}
static inline void two_diff(double a, double b, double* s, double* e) {
    double x = a - b;
    double bv = a - x;
    double av = x + bv;
    double br = bv - b;
    double ar = a - av;
    *s = x;
    *e = ar + br;
}
static inline void two_prod(double a, double b, double* p, double* e) {
    double x = a * b;
    *p = x;
    *e = fma(a, b, -x);
}

/* h = e + b (Grow-Expansion with zero elimination); e nonoverlapping and
 * increasing in magnitude.  Returns length of h (>= 1). */
static int grow(int elen, const double* e, double b, double* h) {
    double q = b, hh;
    int k = 0;
    for (int i = 0; i < elen; i++) {
        two_sum(q, e[i], &q, &hh);
        if (hh != 0.0) h[k++] = hh;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}

/* h = e + f by repeated Grow-Expansion; tmp needs elen+flen+1 */
static int esum(int elen, const double* e, int flen, const double* f, double* h, double* tmp) {
    int n = elen;
    memcpy(h, e, sizeof(double) * (size_t)elen);
    for (int j = 0; j < flen; j++) {
        if (f[j] == 0.0) continue;
        int m = grow(n, h, f[j], tmp);
        memcpy(h, tmp, sizeof(double) * (size_t)m);
        n = m;
    }
    if (n == 0) { h[0] = 0.0; n = 1; }
    return n;
}

/* h = e * b (Scale-Expansion with zero elimination) */
static int scale(int elen, const double* e, double b, double* h) {
    double q, hh, hi, lo, s;
    int k = 0;
    two_prod(e[0], b, &q, &hh);
    if (hh != 0.0) h[k++] = hh;
    for (int i = 1; i < elen; i++) {
        two_prod(e[i], b, &hi, &lo);
        two_sum(q, lo, &s, &hh);
        if (hh != 0.0) h[k++] = hh;
        two_sum(hi, s, &q, &hh);
        if (hh != 0.0) h[k++] = hh;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}

int orc_expansion_sign(int n, const double* e) {
    for (int i = n - 1; i >= 0; i--)
        if (e[i] != 0.0) return e[i] > 0 ? 1 : -1;
    return 0;
}

#define MAXE 8192
/* h = e * f ; returns length */
int orc_exp_product(int elen, const double* e, int flen, const double* f, double* h) {
    double acc[MAXE], part[MAXE], tmp[MAXE];
    int n = 1;
    acc[0] = 0.0;
    for (int j = 0; j < flen; j++) {
        if (f[j] == 0.0) continue;
        int m = scale(elen, e, f[j], part);
        n = esum(n, acc, m, part, h, tmp);
        memcpy(acc, h, sizeof(double) * (size_t)n);
    }
    memcpy(h, acc, sizeof(double) * (size_t)n);
    return n;
}

/* exact (b + kb) - (a + ka) as an expansion; returns length (<= 3) */
static int diff_exp(double b, double a, int k, double* out) {
    double hi, lo, e[2], h[3];
    two_diff(b, a, &hi, &lo);
    int n = 0;
    if (lo != 0.0) e[n++] = lo;
    if (hi != 0.0 || n == 0) e[n++] = hi;
    if (k == 0) { memcpy(out, e, sizeof(double) * (size_t)n); return n; }
    int m = grow(n, e, (double)k, h);
    memcpy(out, h, sizeof(double) * (size_t)m);
    return m;
}
static inline double exp_value(int n, const double* e) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += e[i];
    return s;
}

/* lattice shift that brings b to the minimum image relative to a */
int orc_minimg_shift(double b, double a) {
    double hi, lo;
    two_diff(b, a, &hi, &lo);
    if (hi > 0.5) return -1;
    if (hi < -0.5) return 1;
    return 0;
}

static const double EPS64 = 1.1102230246251565e-16; /* 2^-53 */

/* ---------------------------------------------------------------- orient */
int orc_orient_pt(const orc_pt* a, const orc_pt* b, const orc_pt* c) {
    double ex[3], ey[3], fx[3], fy[3];
    int nex = diff_exp(a->x, c->x, a->sx - c->sx, ex);
    int ney = diff_exp(a->y, c->y, a->sy - c->sy, ey);
    int nfx = diff_exp(b->x, c->x, b->sx - c->sx, fx);
    int nfy = diff_exp(b->y, c->y, b->sy - c->sy, fy);
    double acx = exp_value(nex, ex), acy = exp_value(ney, ey);
    double bcx = exp_value(nfx, fx), bcy = exp_value(nfy, fy);
    double dl = acx * bcy, dr = acy * bcx, det = dl - dr;
    double detsum = fabs(dl) + fabs(dr);
    double bound = 4.0 * (3.0 + 16.0 * EPS64) * EPS64 * detsum;
    if (detsum > 1e-250 && (det > bound || -det > bound)) {
        g_stats[0]++;
        return det > 0 ? 1 : -1;
    }
    g_stats[1]++;
    double l[64], r[64], h[128], tmp[130];
    int nl = orc_exp_product(nex, ex, nfy, fy, l);
    int nr = orc_exp_product(ney, ey, nfx, fx, r);
    for (int i = 0; i < nr; i++) r[i] = -r[i];
    int n = esum(nl, l, nr, r, h, tmp);
    return orc_expansion_sign(n, h);
}

/* -------------------------------------------------------------- incircle */
typedef struct {
    int s_det;                   /* sign of the determinant */
    int s_bc, s_ca, s_ab, s_abc; /* cofactor signs (for the perturbation) */
} incircle_signs;

static int minor_exp(int na, const double* ax, int nb, const double* by, int nc,
                     const double* cx, int nd, const double* dy, double* out) {
    /* out = ax*by - cx*dy */
    double p1[64], p2[64], tmp[160];
    int n1 = orc_exp_product(na, ax, nb, by, p1);
    int n2 = orc_exp_product(nc, cx, nd, dy, p2);
    for (int i = 0; i < n2; i++) p2[i] = -p2[i];
    return esum(n1, p1, n2, p2, out, tmp);
}

static void incircle_exact_full(const orc_pt* a, const orc_pt* b, const orc_pt* c,
                                const orc_pt* d, incircle_signs* out) {
    double adx[3], ady[3], bdx[3], bdy[3], cdx[3], cdy[3];
    int nadx = diff_exp(a->x, d->x, a->sx - d->sx, adx);
    int nady = diff_exp(a->y, d->y, a->sy - d->sy, ady);
    int nbdx = diff_exp(b->x, d->x, b->sx - d->sx, bdx);
    int nbdy = diff_exp(b->y, d->y, b->sy - d->sy, bdy);
    int ncdx = diff_exp(c->x, d->x, c->sx - d->sx, cdx);
    int ncdy = diff_exp(c->y, d->y, c->sy - d->sy, cdy);

    double bc[160], ca[160], ab[160], alift[160], blift[160], clift[160];
    double tmp[MAXE];
    int nbc = minor_exp(nbdx, bdx, ncdy, cdy, ncdx, cdx, nbdy, bdy, bc);
    int nca = minor_exp(ncdx, cdx, nady, ady, nadx, adx, ncdy, cdy, ca);
    int nab = minor_exp(nadx, adx, nbdy, bdy, nbdx, bdx, nady, ady, ab);
    double p1[64], p2[64];
    int n1, n2;
    n1 = orc_exp_product(nadx, adx, nadx, adx, p1);
    n2 = orc_exp_product(nady, ady, nady, ady, p2);
    int nal = esum(n1, p1, n2, p2, alift, tmp);
    n1 = orc_exp_product(nbdx, bdx, nbdx, bdx, p1);
    n2 = orc_exp_product(nbdy, bdy, nbdy, bdy, p2);
    int nbl = esum(n1, p1, n2, p2, blift, tmp);
    n1 = orc_exp_product(ncdx, cdx, ncdx, cdx, p1);
    n2 = orc_exp_product(ncdy, cdy, ncdy, cdy, p2);
    int ncl = esum(n1, p1, n2, p2, clift, tmp);

    double t1[MAXE], t2[MAXE], t3[MAXE], s12[MAXE], det[MAXE];
    int m1 = orc_exp_product(nal, alift, nbc, bc, t1);
    int m2 = orc_exp_product(nbl, blift, nca, ca, t2);
    int m3 = orc_exp_product(ncl, clift, nab, ab, t3);
    int m12 = esum(m1, t1, m2, t2, s12, tmp);
    int md = esum(m12, s12, m3, t3, det, tmp);
    out->s_det = orc_expansion_sign(md, det);
    out->s_bc = orc_expansion_sign(nbc, bc);
    out->s_ca = orc_expansion_sign(nca, ca);
    out->s_ab = orc_expansion_sign(nab, ab);
    double s[400], s2[600];
    int ns = esum(nbc, bc, nca, ca, s, tmp);
    int ns2 = esum(ns, s, nab, ab, s2, tmp);
    out->s_abc = orc_expansion_sign(ns2, s2); /* orient(a,b,c) = bc + ca + ab */
}

/* filtered determinant sign; returns 2 when undecided */
static int incircle_filter(const orc_pt* a, const orc_pt* b, const orc_pt* c, const orc_pt* d) {
    double e[3];
    int n;
    n = diff_exp(a->x, d->x, a->sx - d->sx, e); double adx = exp_value(n, e);
    n = diff_exp(a->y, d->y, a->sy - d->sy, e); double ady = exp_value(n, e);
    n = diff_exp(b->x, d->x, b->sx - d->sx, e); double bdx = exp_value(n, e);
    n = diff_exp(b->y, d->y, b->sy - d->sy, e); double bdy = exp_value(n, e);
    n = diff_exp(c->x, d->x, c->sx - d->sx, e); double cdx = exp_value(n, e);
    n = diff_exp(c->y, d->y, c->sy - d->sy, e); double cdy = exp_value(n, e);
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    double alift = adx * adx + ady * ady;
    double cdxady = cdx * ady, adxcdy = adx * cdy;
    double blift = bdx * bdx + bdy * bdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    double clift = cdx * cdx + cdy * cdy;
    double det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) + clift * (adxbdy - bdxady);
    double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * alift + (fabs(cdxady) + fabs(adxcdy)) * blift +
                  (fabs(adxbdy) + fabs(bdxady)) * clift;
    double bound = 4.0 * (10.0 + 96.0 * EPS64) * EPS64 * perm;
    if (perm > 1e-250 && (det > bound || -det > bound)) return det > 0 ? 1 : -1;
    return 2;
}

int orc_incircle_pt(const orc_pt* a, const orc_pt* b, const orc_pt* c, const orc_pt* d) {
    int s = incircle_filter(a, b, c, d);
    if (s != 2) { g_stats[2]++; return s; }
    g_stats[3]++;
    incircle_signs sg;
    incircle_exact_full(a, b, c, d, &sg);
    return sg.s_det;
}

/* rank order for the perturbation: canonical coordinates, then shift */
static int rank_greater(const orc_pt* p, const orc_pt* q) {
    if (p->x != q->x) return p->x > q->x;
    if (p->y != q->y) return p->y > q->y;
    if (p->sx != q->sx) return p->sx > q->sx;
    return p->sy > q->sy;
}

/*
 * Symbolic perturbation.  The incircle determinant equals the 4x4
 * determinant with rows (x, y, x^2+y^2, 1), linear in each lift z_i with
 * cofactors C_a = orient(b,c,d) = bc, C_b = -orient(a,c,d) = ca,
 * C_c = orient(a,b,d) = ab, C_d = -orient(a,b,c).  Raising z_i by eps^(k_i),
 * with smaller k_i for lexicographically larger canonical coordinates, makes
 * the sign of an exactly-zero determinant the sign of the first nonzero
 * cofactor in decreasing lexicographic order of the points.  C_d != 0
 * whenever a, b, c is a proper triangle, so the rule always decides.
 */
int orc_incircle_sos_pt(const orc_pt* a, const orc_pt* b, const orc_pt* c, const orc_pt* d) {
    int s = incircle_filter(a, b, c, d);
    if (s != 2) { g_stats[2]++; return s; }
    g_stats[3]++;
    incircle_signs sg;
    incircle_exact_full(a, b, c, d, &sg);
    if (sg.s_det != 0) return sg.s_det;
    const orc_pt* pts[4] = {a, b, c, d};
    int cof[4] = {sg.s_bc, sg.s_ca, sg.s_ab, -sg.s_abc};
    int order[4] = {0, 1, 2, 3};
    for (int i = 0; i < 4; i++) /* sort decreasing rank */
        for (int j = i + 1; j < 4; j++)
            if (rank_greater(pts[order[j]], pts[order[i]])) {
                int t = order[i]; order[i] = order[j]; order[j] = t;
            }
    for (int i = 0; i < 4; i++)
        if (cof[order[i]] != 0) return cof[order[i]];
    return 0; /* all four points collinear */
}

/* ------------------------------------------- plain / min-image front ends */
static void mk_pt(orc_pt* p, const double* c) { p->x = c[0]; p->y = c[1]; p->sx = 0; p->sy = 0; }
static void mk_rel(orc_pt* p, const double* c, const double* base, int periodic) {
    mk_pt(p, c);
    if (periodic) { p->sx = orc_minimg_shift(c[0], base[0]); p->sy = orc_minimg_shift(c[1], base[1]); }
}

int orc_orient2d_exact(const double* a, const double* b, const double* c, int periodic) {
    orc_pt pa, pb, pc;
    mk_rel(&pa, a, c, periodic); mk_rel(&pb, b, c, periodic); mk_pt(&pc, c);
    return orc_orient_pt(&pa, &pb, &pc);
}
int orc_incircle_exact(const double* a, const double* b, const double* c, const double* d,
                       int periodic) {
    orc_pt pa, pb, pc, pd;
    mk_rel(&pa, a, d, periodic); mk_rel(&pb, b, d, periodic); mk_rel(&pc, c, d, periodic);
    mk_pt(&pd, d);
    return orc_incircle_pt(&pa, &pb, &pc, &pd);
}
int orc_incircle_sos(const double* a, const double* b, const double* c, const double* d,
                     int periodic) {
    orc_pt pa, pb, pc, pd;
    mk_rel(&pa, a, d, periodic); mk_rel(&pb, b, d, periodic); mk_rel(&pc, c, d, periodic);
    mk_pt(&pd, d);
    return orc_incircle_sos_pt(&pa, &pb, &pc, &pd);
}
