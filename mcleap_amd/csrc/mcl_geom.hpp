// mcl_geom.hpp -- geometric predicates for the kinetic Delaunay engine.
//
// Every predicate sign is exact.  Evaluation runs as a cascade:
//   stage 0  FP32: differences rounded once to FP32, Shewchuk's semi-static
//            error bound A with eps = 2^-24 (valid while no term underflows;
//            the permanent must exceed 2^-100 to use this stage);
//   stage 1  FP64: the same with eps = 2^-53;
//   stage 2  exact expansion arithmetic when every coordinate difference is a
//            single double ("small": <= 96 components);
//   stage 3  exact expansion arithmetic for differences that need two or
//            three components ("big": <= 7776 components).  Only kernels that
//            are launched with few waves instantiate this stage; elsewhere it
//            is reported as undecided (kNeedBig) and retried there.
// The expansion algorithms are Two-Sum, Two-Diff, Two-Product (fma),
// Grow-Expansion, Expansion-Sum (as repeated growth) and Scale-Expansion from
// J. R. Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
// Robust Geometric Predicates" (CMU-CS-96-140).
//
// Semantics match the reference's calls: orient2d(a, b, c) > 0 for CCW
// (Helpers_Triangulation.h:86-122); inCircle(a, b, c, d) > 0 when d lies
// inside the circle through CCW a, b, c (Helpers_Triangulation.h:136-151).
// Exactly cocircular quads are decided by a symbolic perturbation of the
// lifts ordered by the points' lexicographic (x, y) order (largest point,
// largest perturbation), so the Delaunay triangulation is unique for any
// set of distinct points and any vertex numbering.
//
// Periodic (unit torus) mode forms each difference as the minimum image,
// exactly: Two-Diff, then a +-1 shift of the high part (exact by Sterbenz).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define MCL_HD __host__ __device__ __forceinline__

namespace mcl {

enum PredStage : int { kStageF32 = 0, kStageF64 = 1, kStageExactSmall = 2, kStageExactBig = 3 };
constexpr int kNeedBig = 2;  // sign value meaning "undecided without the big exact stage"

// ----------------------------------------------------------- error-free ops
MCL_HD void two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
    double x = a + b;
    double bv = x - a;
    double av = x - bv;
    double br = b - bv;
    double ar = a - av;
    s = x;
    e = ar + br;
}
MCL_HD void two_diff(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
    double x = a - b;
    double bv = a - x;
    double av = x + bv;
    double br = bv - b;
    double ar = a - av;
    s = x;
    e = ar + br;
}
MCL_HD void two_prod(double a, double b, double& p, double& e) {
#pragma clang fp contract(off)
    double x = a * b;
    p = x;
    e = fma(a, b, -x);
}

// h = h + b in place (Grow-Expansion with zero elimination: the write index
// never passes the read index); h needs room for n + 1; returns length >= 1
MCL_HD int exp_grow(double* h, int n, double b) {
    double q = b, hh;
    int k = 0;
    for (int i = 0; i < n; i++) {
        two_sum(q, h[i], q, hh);
        if (hh != 0.0) h[k++] = hh;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}
// h = h + f in place by repeated growth (Expansion-Sum); room for n + nf
MCL_HD int exp_add(double* h, int n, const double* f, int nf) {
    for (int j = 0; j < nf; j++)
        if (f[j] != 0.0) n = exp_grow(h, n, f[j]);
    return n;
}
// h = e * b, zero-eliminated (Scale-Expansion); room for 2 * n
MCL_HD int exp_scale(const double* e, int n, double b, double* h) {
    double q, hh, hi, lo, s;
    int k = 0;
    two_prod(e[0], b, q, hh);
    if (hh != 0.0) h[k++] = hh;
    for (int i = 1; i < n; i++) {
        two_prod(e[i], b, hi, lo);
        two_sum(q, lo, s, hh);
        if (hh != 0.0) h[k++] = hh;
        two_sum(hi, s, q, hh);
        if (hh != 0.0) h[k++] = hh;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}
MCL_HD int exp_sign(const double* e, int n) {
    for (int i = n - 1; i >= 0; i--)
        if (e[i] != 0.0) return e[i] > 0.0 ? 1 : -1;
    return 0;
}
// h += e * f ; sc needs room for 2 * ne, h for the final length
MCL_HD int exp_addmul(double* h, int n, const double* e, int ne, const double* f, int nf, double* sc) {
    for (int j = 0; j < nf; j++) {
        if (f[j] == 0.0) continue;
        int m = exp_scale(e, ne, f[j], sc);
        n = exp_add(h, n, sc, m);
    }
    return n;
}
// h = e * f
MCL_HD int exp_mul(const double* e, int ne, const double* f, int nf, double* h, double* sc) {
    h[0] = 0.0;
    return exp_addmul(h, 1, e, ne, f, nf, sc);
}
MCL_HD int exp_neg(double* h, int n) {
    for (int i = 0; i < n; i++) h[i] = -h[i];
    return n;
}

// ------------------------------------------------------------- differences
// exact b - a (+ integer shift k), minimum image when periodic; <= 3 components
struct Dexp {
    double c[3];
    int n;
};
MCL_HD Dexp diff_exact(double b, double a, int k, bool periodic) {
    double hi, lo;
    two_diff(b, a, hi, lo);
    if (periodic) {
        if (hi > 0.5) hi -= 1.0;
        else if (hi < -0.5) hi += 1.0;
    }
    Dexp d;
    double e[2];
    int n = 0;
    if (lo != 0.0) e[n++] = lo;
    if (hi != 0.0 || n == 0) e[n++] = hi;
    if (k == 0) {
        d.n = n;
        d.c[0] = e[0];
        d.c[1] = n > 1 ? e[1] : 0.0;
        d.c[2] = 0.0;
        return d;
    }
    d.c[0] = e[0];
    d.c[1] = n > 1 ? e[1] : 0.0;
    d.n = exp_grow(d.c, n, (double)k);
    return d;
}
// the components summed in order; constant indices keep a Dexp in registers
// (a loop over d.n would index the array at run time and put it in scratch)
MCL_HD double dexp_value(const Dexp& d) {
    double s = 0.0;
    if (d.n > 0) s += d.c[0];
    if (d.n > 1) s += d.c[1];
    if (d.n > 2) s += d.c[2];
    return s;
}
MCL_HD float min_image_f(float d) {
    if (d > 0.5f) d -= 1.0f;
    else if (d < -0.5f) d += 1.0f;
    return d;
}

// --------------------------------------------------------------- orient2d
// exact sign from difference expansions ac = a - c, bc = b - c
MCL_HD int orient_exact_from(const Dexp& acx, const Dexp& acy, const Dexp& bcx, const Dexp& bcy) {
    double h[40], r[20], sc[8];
    int n = exp_mul(acx.c, acx.n, bcy.c, bcy.n, h, sc);
    int nr = exp_neg(r, exp_mul(acy.c, acy.n, bcx.c, bcx.n, r, sc));
    n = exp_add(h, n, r, nr);
    return exp_sign(h, n);
}

// orient2d(a, b, c) with the full cascade; *stage reports which stage decided
template <bool kF32>
MCL_HD int orient(float2 a, float2 b, float2 c, bool periodic, int* stage) {
#pragma clang fp contract(off)
    if (kF32 && !periodic) {
        const float e32 = 5.9604644775390625e-08f;  // 2^-24
        float acx = a.x - c.x, acy = a.y - c.y, bcx = b.x - c.x, bcy = b.y - c.y;
        float dl = acx * bcy, dr = acy * bcx, det = dl - dr;
        float ds = fabsf(dl) + fabsf(dr);
        float bound = (3.0f + 16.0f * e32) * e32 * ds;
        if (ds > 7.9e-31f && (det > bound || -det > bound)) { *stage = kStageF32; return det > 0 ? 1 : -1; }
    }
    Dexp acx = diff_exact(a.x, c.x, 0, periodic), acy = diff_exact(a.y, c.y, 0, periodic);
    Dexp bcx = diff_exact(b.x, c.x, 0, periodic), bcy = diff_exact(b.y, c.y, 0, periodic);
    if (kF32 && periodic) {
        const float e32 = 5.9604644775390625e-08f;
        float ax = (float)dexp_value(acx), ay = (float)dexp_value(acy);
        float bx = (float)dexp_value(bcx), by = (float)dexp_value(bcy);
        float dl = ax * by, dr = ay * bx, det = dl - dr;
        float ds = fabsf(dl) + fabsf(dr);
        float bound = (3.0f + 16.0f * e32) * e32 * ds * 1.0001f;
        if (ds > 7.9e-31f && (det > bound || -det > bound)) { *stage = kStageF32; return det > 0 ? 1 : -1; }
    }
    {
        const double e64 = 1.1102230246251565e-16;  // 2^-53
        double ax = dexp_value(acx), ay = dexp_value(acy), bx = dexp_value(bcx), by = dexp_value(bcy);
        double dl = ax * by, dr = ay * bx, det = dl - dr;
        double ds = fabs(dl) + fabs(dr);
        double bound = (3.0 + 16.0 * e64) * e64 * ds;
        if (ds > 1e-280 && (det > bound || -det > bound)) { *stage = kStageF64; return det > 0 ? 1 : -1; }
    }
    *stage = (acx.n | acy.n | bcx.n | bcy.n) > 1 ? kStageExactBig : kStageExactSmall;
    return orient_exact_from(acx, acy, bcx, bcy);
}

// --------------------------------------------------------------- incircle
struct InSigns {
    int det, bc, ca, ab, abc;
};

// exact determinant and cofactor signs.  Capacities: MC >= length of a
// minor or lift (2*d*d with d-component differences), DC >= 3*2*MC*MC.
template <int MC, int DC>
MCL_HD void incircle_exact_from(const Dexp& adx, const Dexp& ady, const Dexp& bdx, const Dexp& bdy,
                                const Dexp& cdx, const Dexp& cdy, InSigns& s) {
    double bc[MC], ca[MC], ab[MC], al[MC], bl[MC], cl[MC], q[MC], sc[2 * MC + 2];
    int nbc, nca, nab, nal, nbl, ncl, nq;
    // bc = bdx*cdy - cdx*bdy ; ca = cdx*ady - adx*cdy ; ab = adx*bdy - bdx*ady
    nbc = exp_mul(bdx.c, bdx.n, cdy.c, cdy.n, bc, sc);
    nq = exp_neg(q, exp_mul(cdx.c, cdx.n, bdy.c, bdy.n, q, sc));
    nbc = exp_add(bc, nbc, q, nq);
    nca = exp_mul(cdx.c, cdx.n, ady.c, ady.n, ca, sc);
    nq = exp_neg(q, exp_mul(adx.c, adx.n, cdy.c, cdy.n, q, sc));
    nca = exp_add(ca, nca, q, nq);
    nab = exp_mul(adx.c, adx.n, bdy.c, bdy.n, ab, sc);
    nq = exp_neg(q, exp_mul(bdx.c, bdx.n, ady.c, ady.n, q, sc));
    nab = exp_add(ab, nab, q, nq);
    // lifts
    nal = exp_mul(adx.c, adx.n, adx.c, adx.n, al, sc);
    nal = exp_addmul(al, nal, ady.c, ady.n, ady.c, ady.n, sc);
    nbl = exp_mul(bdx.c, bdx.n, bdx.c, bdx.n, bl, sc);
    nbl = exp_addmul(bl, nbl, bdy.c, bdy.n, bdy.c, bdy.n, sc);
    ncl = exp_mul(cdx.c, cdx.n, cdx.c, cdx.n, cl, sc);
    ncl = exp_addmul(cl, ncl, cdy.c, cdy.n, cdy.c, cdy.n, sc);
    // det = al*bc + bl*ca + cl*ab, accumulated in place
    double acc[DC];
    int nacc = 1;
    acc[0] = 0.0;
    nacc = exp_addmul(acc, nacc, al, nal, bc, nbc, sc);
    nacc = exp_addmul(acc, nacc, bl, nbl, ca, nca, sc);
    nacc = exp_addmul(acc, nacc, cl, ncl, ab, nab, sc);
    s.det = exp_sign(acc, nacc);
    s.bc = exp_sign(bc, nbc);
    s.ca = exp_sign(ca, nca);
    s.ab = exp_sign(ab, nab);
    // orient(a, b, c) = bc + ca + ab
    for (int i = 0; i < nbc; i++) acc[i] = bc[i];
    nacc = exp_add(acc, nbc, ca, nca);
    nacc = exp_add(acc, nacc, ab, nab);
    s.abc = exp_sign(acc, nacc);
}

MCL_HD bool lex_greater(float2 p, float2 q) { return p.x != q.x ? p.x > q.x : p.y > q.y; }

// sign of the perturbed determinant given the exact signs
MCL_HD int sos_resolve(const InSigns& s, float2 a, float2 b, float2 c, float2 d) {
    if (s.det != 0) return s.det;
    float2 P[4] = {a, b, c, d};
    int cof[4] = {s.bc, s.ca, s.ab, -s.abc};
    int o[4] = {0, 1, 2, 3};
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++)
            if (lex_greater(P[o[j]], P[o[i]])) { int t = o[i]; o[i] = o[j]; o[j] = t; }
    for (int i = 0; i < 4; i++)
        if (cof[o[i]] != 0) return cof[o[i]];
    return 0;
}

// The FP32 stage of incircle(a, b, c, d) on the plane (no minimum image):
// +-1, or 0 when its error bound cannot decide.
MCL_HD int incircle_f32_plane(float2 a, float2 b, float2 c, float2 d) {
#pragma clang fp contract(off)
    const float e = 5.9604644775390625e-08f;
    float adx = a.x - d.x, ady = a.y - d.y, bdx = b.x - d.x, bdy = b.y - d.y;
    float cdx = c.x - d.x, cdy = c.y - d.y;
    float bdxcdy = bdx * cdy, cdxbdy = cdx * bdy, alift = adx * adx + ady * ady;
    float cdxady = cdx * ady, adxcdy = adx * cdy, blift = bdx * bdx + bdy * bdy;
    float adxbdy = adx * bdy, bdxady = bdx * ady, clift = cdx * cdx + cdy * cdy;
    float det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) + clift * (adxbdy - bdxady);
    float perm = (fabsf(bdxcdy) + fabsf(cdxbdy)) * alift + (fabsf(cdxady) + fabsf(adxcdy)) * blift +
                 (fabsf(adxbdy) + fabsf(bdxady)) * clift;
    float bound = (10.0f + 96.0f * e) * e * perm;
    if (perm > 7.9e-31f && perm < 1e37f && (det > bound || -det > bound)) return det > 0 ? 1 : -1;
    return 0;
}

// incircle(a, b, c, d) with the symbolic-perturbation tie rule.  Returns
// +-1, or kNeedBig when kBig is false and only the big exact stage decides.
template <bool kF32, bool kBig>
MCL_HD int incircle(float2 a, float2 b, float2 c, float2 d, bool periodic, int* stage) {
#pragma clang fp contract(off)
    if (kF32 && !periodic) {
        const int sg = incircle_f32_plane(a, b, c, d);
        if (sg != 0) {
            *stage = kStageF32;
            return sg;
        }
    }
    Dexp adx = diff_exact(a.x, d.x, 0, periodic), ady = diff_exact(a.y, d.y, 0, periodic);
    Dexp bdx = diff_exact(b.x, d.x, 0, periodic), bdy = diff_exact(b.y, d.y, 0, periodic);
    Dexp cdx = diff_exact(c.x, d.x, 0, periodic), cdy = diff_exact(c.y, d.y, 0, periodic);
    if (kF32 && periodic) {
        const float e = 5.9604644775390625e-08f;
        float ax = (float)dexp_value(adx), ay = (float)dexp_value(ady);
        float bx = (float)dexp_value(bdx), by = (float)dexp_value(bdy);
        float cx = (float)dexp_value(cdx), cy = (float)dexp_value(cdy);
        float bdxcdy = bx * cy, cdxbdy = cx * by, alift = ax * ax + ay * ay;
        float cdxady = cx * ay, adxcdy = ax * cy, blift = bx * bx + by * by;
        float adxbdy = ax * by, bdxady = bx * ay, clift = cx * cx + cy * cy;
        float det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) + clift * (adxbdy - bdxady);
        float perm = (fabsf(bdxcdy) + fabsf(cdxbdy)) * alift + (fabsf(cdxady) + fabsf(adxcdy)) * blift +
                     (fabsf(adxbdy) + fabsf(bdxady)) * clift;
        float bound = (10.0f + 96.0f * e) * e * perm * 1.0001f;
        if (perm > 7.9e-31f && perm < 1e37f && (det > bound || -det > bound)) {
            *stage = kStageF32;
            return det > 0 ? 1 : -1;
        }
    }
    {
        const double e = 1.1102230246251565e-16;
        double ax = dexp_value(adx), ay = dexp_value(ady), bx = dexp_value(bdx), by = dexp_value(bdy);
        double cx = dexp_value(cdx), cy = dexp_value(cdy);
        double bdxcdy = bx * cy, cdxbdy = cx * by, alift = ax * ax + ay * ay;
        double cdxady = cx * ay, adxcdy = ax * cy, blift = bx * bx + by * by;
        double adxbdy = ax * by, bdxady = bx * ay, clift = cx * cx + cy * cy;
        double det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) + clift * (adxbdy - bdxady);
        double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * alift + (fabs(cdxady) + fabs(adxcdy)) * blift +
                      (fabs(adxbdy) + fabs(bdxady)) * clift;
        double bound = (10.0 + 96.0 * e) * e * perm;
        if (perm > 1e-280 && (det > bound || -det > bound)) {
            *stage = kStageF64;
            return det > 0 ? 1 : -1;
        }
    }
    InSigns s;
    bool single = (adx.n | ady.n | bdx.n | bdy.n | cdx.n | cdy.n) == 1;
    if (single) {
        *stage = kStageExactSmall;
        incircle_exact_from<4, 96>(adx, ady, bdx, bdy, cdx, cdy, s);
        return sos_resolve(s, a, b, c, d);
    }
    if (!kBig) return kNeedBig;
    *stage = kStageExactBig;
    incircle_exact_from<16, 1536>(adx, ady, bdx, bdy, cdx, cdy, s);
    return sos_resolve(s, a, b, c, d);
}

// reference-compatible FP64 forms (the "plain" predicate), kept for the
// plain path's accuracy report only: Helpers_Triangulation.h:86-151
MCL_HD int orient_ref64(float2 a, float2 b, float2 c) {
    double l = ((double)b.x - (double)a.x) * ((double)c.y - (double)a.y);
    double r = ((double)b.y - (double)a.y) * ((double)c.x - (double)a.x);
    double det = l - r;
    return fabs(det) < 1e-17f ? 0 : (det > 0 ? 1 : -1);
}

}  // namespace mcl
