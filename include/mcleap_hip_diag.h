/*
 * mcleap_hip_diag.h -- verification entry points of libmcleap_hip.so.
 *
 * These have no counterpart in the reference: they expose the device
 * predicates of the flip, repair and relocation kernels so that callers can
 * measure them (sign agreement with an exact oracle, fraction decided at each
 * stage of the cascade, disagreements of a plain FP64 predicate).
 */
#ifndef MCLEAP_HIP_DIAG_H
#define MCLEAP_HIP_DIAG_H

#include "mcleap_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MCL_PRED_ORIENT 0          /* orient2d, FP32 filter -> FP64 filter -> exact */
#define MCL_PRED_INCIRCLE 1        /* incircle with the tie rule, FP32 -> FP64 -> exact */
#define MCL_PRED_ORIENT_F64 2      /* orient2d, FP64 filter -> exact (the plain form's cascade) */
#define MCL_PRED_INCIRCLE_F64 3    /* incircle, FP64 filter -> exact */
#define MCL_PRED_ORIENT_PLAIN 4    /* sign of the FP64 determinant, no filter, no exact stage */
#define MCL_PRED_INCIRCLE_PLAIN 5  /* sign of the FP64 lifted determinant, no filter */
#define MCL_PRED_INCIRCLE_EXACT 6  /* incircle by exact expansion arithmetic alone, tie rule (no filter) */
#define MCL_PRED_ORIENT_EXACT 7    /* orient2d by exact expansion arithmetic alone */

/* Evaluates n predicate cases on the device.  d_pts holds float2 points,
 * 3 per case for the orientation kinds (a, b, c: sign of (a-c) x (b-c)) and
 * 4 per case for the incircle kinds (a, b, c, d: positive when d is inside
 * the circle through CCW a, b, c).  periodic != 0 takes differences as the
 * minimum image on the unit torus.  d_sign receives -1/0/+1 per case;
 * d_stage (may be NULL) receives the deciding stage 0 FP32, 1 FP64,
 * 2 exact-small, 3 exact-big (-1 for the plain and exact-only kinds).
 * Synchronous. */
int mcl_eval_predicates(int kind, int n, const void* d_pts, int periodic, void* d_sign, void* d_stage,
                        void* stream);

#ifdef __cplusplus
}
#endif
#endif
