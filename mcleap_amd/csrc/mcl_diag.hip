// mcl_diag.hip -- evaluation of the device predicates on caller arrays.
//
// The same device functions the flip, repair and relocation kernels call
// (mcl_geom.hpp), run one case per thread so that their signs and the stage
// that decided each sign can be compared with exact rational signs, and the
// plain FP64 forms (the sign of the determinant evaluated in double, no
// filter, no exact stage) can be compared with the exact cascade.
#include <algorithm>

#include "mcl_geom.hpp"
#include "mcl_internal.hpp"

namespace mcl {
namespace {

// kind: 0 orient (FP32 first), 1 incircle with tie rule (FP32 first),
//       2 orient (FP64 first), 3 incircle (FP64 first),
//       4 plain FP64 orient sign, 5 plain FP64 incircle sign,
//       6 incircle by exact expansions only (tie rule), 7 orient exact only
__global__ void __launch_bounds__(256) k_eval_pred(int kind, int n, const float2* __restrict__ pts, int periodic,
                                                   int* __restrict__ sign, int* __restrict__ stage) {
    const bool per = periodic != 0;
    // grid-stride over the cases: the exact stages' private arrays make a
    // resident wave large, so the launch keeps a bounded number of them
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int st = -1, s = 0;
        if (kind == 0 || kind == 2 || kind == 4) {
            const float2 a = pts[3 * i], b = pts[3 * i + 1], c = pts[3 * i + 2];
            if (kind == 0) s = orient<true>(a, b, c, per, &st);
            else if (kind == 2) s = orient<false>(a, b, c, per, &st);
            else {
                const Dexp ax = diff_exact(a.x, c.x, 0, per), ay = diff_exact(a.y, c.y, 0, per);
                const Dexp bx = diff_exact(b.x, c.x, 0, per), by = diff_exact(b.y, c.y, 0, per);
                const double det = dexp_value(ax) * dexp_value(by) - dexp_value(ay) * dexp_value(bx);
                s = (det > 0) - (det < 0);
            }
        } else if (kind == 7) {  // orientation by exact expansions alone
            const float2 a = pts[3 * i], b = pts[3 * i + 1], c = pts[3 * i + 2];
            s = orient_exact_from(diff_exact(a.x, c.x, 0, per), diff_exact(a.y, c.y, 0, per), diff_exact(b.x, c.x, 0, per),
                                  diff_exact(b.y, c.y, 0, per));
        } else if (kind == 6) {  // incircle by exact expansions alone, with the tie rule
            const float2 a = pts[4 * i], b = pts[4 * i + 1], c = pts[4 * i + 2], d = pts[4 * i + 3];
            InSigns sg;
            incircle_exact_from<16, 1536>(diff_exact(a.x, d.x, 0, per), diff_exact(a.y, d.y, 0, per),
                                          diff_exact(b.x, d.x, 0, per), diff_exact(b.y, d.y, 0, per),
                                          diff_exact(c.x, d.x, 0, per), diff_exact(c.y, d.y, 0, per), sg);
            s = sos_resolve(sg, a, b, c, d);
        } else {
            const float2 a = pts[4 * i], b = pts[4 * i + 1], c = pts[4 * i + 2], d = pts[4 * i + 3];
            if (kind == 1) s = incircle<true, true>(a, b, c, d, per, &st);
            else if (kind == 3) s = incircle<false, true>(a, b, c, d, per, &st);
            else {
                const double ax = dexp_value(diff_exact(a.x, d.x, 0, per)), ay = dexp_value(diff_exact(a.y, d.y, 0, per));
                const double bx = dexp_value(diff_exact(b.x, d.x, 0, per)), by = dexp_value(diff_exact(b.y, d.y, 0, per));
                const double cx = dexp_value(diff_exact(c.x, d.x, 0, per)), cy = dexp_value(diff_exact(c.y, d.y, 0, per));
                const double det = (ax * ax + ay * ay) * (bx * cy - cx * by) + (bx * bx + by * by) * (cx * ay - ax * cy) +
                                   (cx * cx + cy * cy) * (ax * by - bx * ay);
                s = (det > 0) - (det < 0);
            }
        }
        sign[i] = s;
        if (stage) stage[i] = st;
    }
}

}  // namespace
}  // namespace mcl

extern "C" int mcl_eval_predicates(int kind, int n, const void* d_pts, int periodic, void* d_sign, void* d_stage,
                                   void* stream) {
    if (kind < 0 || kind > 7 || n < 0 || (n > 0 && (!d_pts || !d_sign))) return MCL_ERR_INPUT;
    if (n == 0) return MCL_OK;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(mcl::k_eval_pred, dim3(std::min((n + 255) / 256, 1024)), dim3(256), 0, s, kind, n, (const float2*)d_pts,
                       periodic, (int*)d_sign, (int*)d_stage);
    MCL_HIP(hipGetLastError());
    MCL_HIP(hipStreamSynchronize(s));
    return MCL_OK;
}
