// Small vector kernels shared by the solver seam and the time-stepping driver.
#include <hip/hip_runtime.h>

#include "vec.hpp"

namespace emtd {
namespace {

// out[k + r*ld_out] = in[idx[k] + r*ld_in]
__global__ void k_gather(double* __restrict__ out, const double* __restrict__ in, const int* __restrict__ idx,
                         int count, int nrhs, int64_t ld_out, int64_t ld_in) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int src = idx[k];
    for (int r = 0; r < nrhs; ++r) out[k + r * ld_out] = in[src + r * ld_in];
}

// out[idx[k] + r*ld_out] = in[k + r*ld_in]
__global__ void k_scatter(double* __restrict__ out, const double* __restrict__ in, const int* __restrict__ idx,
                          int count, int nrhs, int64_t ld_out, int64_t ld_in) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int dst = idx[k];
    for (int r = 0; r < nrhs; ++r) out[dst + r * ld_out] = in[k + r * ld_in];
}

// x[idx[k]] = b[idx[k]] * dinv[k] for the singleton DOFs
__global__ void k_singletons(double* __restrict__ x, const double* __restrict__ b, const int* __restrict__ idx,
                             const double* __restrict__ dinv, int count, int nrhs, int64_t ld) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int i = idx[k];
    for (int r = 0; r < nrhs; ++r) x[i + r * ld] = b[i + r * ld] * dinv[k];
}

}  // namespace

void gather(hipStream_t s, double* out, const double* in, const int* idx, int count, int nrhs, int64_t ld_out,
            int64_t ld_in) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_gather, dim3((count + 255) / 256), dim3(256), 0, s, out, in, idx, count, nrhs, ld_out, ld_in);
}

void scatter(hipStream_t s, double* out, const double* in, const int* idx, int count, int nrhs, int64_t ld_out,
             int64_t ld_in) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_scatter, dim3((count + 255) / 256), dim3(256), 0, s, out, in, idx, count, nrhs, ld_out,
                       ld_in);
}

void singletons(hipStream_t s, double* x, const double* b, const int* idx, const double* dinv, int count, int nrhs,
                int64_t ld) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_singletons, dim3((count + 255) / 256), dim3(256), 0, s, x, b, idx, dinv, count, nrhs, ld);
}

}  // namespace emtd
