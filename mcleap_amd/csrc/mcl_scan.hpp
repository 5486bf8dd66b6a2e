// mcl_scan.hpp -- device-wide exclusive prefix sum (mcl_scan.hip)
#pragma once
#include <hip/hip_runtime.h>

namespace mcl {
// ints of scratch needed for a scan of n elements
int scan_scratch_ints(int n);
// d_out[0..n) = exclusive prefix sums of d_in, d_out[n] = total.
// d_in and d_out may alias.  Asynchronous on stream s.
int scan_exclusive(const int* d_in, int* d_out, int n, int* d_scratch, hipStream_t s);
}  // namespace mcl
