// GPU assembly of the curl-curl and mass matrices in lower RSF storage.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/emtd.h"
#include "common.hpp"

namespace emtd {

struct ElemData {
    double hx, hy, hz, sigma, mu;
};

// Contribution lists per slot (lower entries first, then the diagonal), in element order.
void build_contributions(int kpar, const int* ed, int n, const int* ig, const int* jg, std::vector<int64_t>& cptr,
                         std::vector<uint32_t>& codes);

class Assembler {
  public:
    Assembler(const emtd_mesh& m, int n, const int* ig, const int* jg, hipStream_t s);
    // Writes B (curl-curl/mu) and C (sigma mass) into device arrays (lower RSF + diagonal).
    void run(double* d_b_ggl, double* d_b_di, double* d_c_ggl, double* d_c_di);
    // Runs only the assembly kernel into [lower | diagonal] arrays; returns its GPU time.
    double run_timed(double* b_all, double* c_all);
    int64_t nnz() const { return nnz_; }
    int64_t contributions() const { return int64_t(codes_.n); }
    hipStream_t stream;

  private:
    int n_ = 0;
    int64_t nnz_ = 0;
    DevBuf<int64_t> cptr_;
    DevBuf<uint32_t> codes_;
    DevBuf<ElemData> el_;
    DevBuf<double> bscratch_, cscratch_;
};

}  // namespace emtd
