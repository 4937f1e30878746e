// Launch wrappers of the small vector kernels (vec.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace emtd {

void gather(hipStream_t s, double* out, const double* in, const int* idx, int count, int nrhs, int64_t ld_out,
            int64_t ld_in);
void scatter(hipStream_t s, double* out, const double* in, const int* idx, int count, int nrhs, int64_t ld_out,
             int64_t ld_in);
void singletons(hipStream_t s, double* x, const double* b, const int* idx, const double* dinv, int count, int nrhs,
                int64_t ld);

}  // namespace emtd
