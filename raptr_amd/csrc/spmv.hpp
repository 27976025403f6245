// spmv.hpp -- internal SpMV helpers shared by the MLEM drivers.
#pragma once

#include "common.hpp"

namespace raptr {

// Lanes per row for a mean row length (4..64, powers of two).
int pick_lanes(double mean);

// y = A x for a CSR with 64-bit row pointers (FP32 accumulation).
int spmv64_f32(int64_t m, const int64_t* rp, const int* col, const float* val,
               const float* x, float* y, int lanes, hipStream_t st);

}  // namespace raptr
