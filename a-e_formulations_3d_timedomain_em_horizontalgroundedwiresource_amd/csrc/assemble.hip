// Edge-element assembly on the GPU.
//
// Replaces T_Global_SLAE_Vec::Assembling (A2/T_Global_SLAE.cpp:76-215) for the curl-curl
// matrix B (Assembling(2), local matrix b/mu, A2/T_Brick.cpp:142-310,494-501) and the mass
// matrix C (Assembling(1), local matrix c*sigma, A2/T_Brick.cpp:311-474,481-491).
//
// Gather formulation: every lower RSF slot (row i, column j<i) and every diagonal entry owns
// the list of (element, local a, local b) contributions, built once on the host in element
// order. One thread per slot sums its contributions in that order (the reference's order)
// and writes the B and C values once: no atomics, no colouring, coalesced stores.
//
// Local entries of an axis-aligned brick with sizes h (volume V), edge a of direction d
// with transverse axes p=(d+1)%3, q=(d+2)%3 and bits (sp, sq):
//   C_ab = 4V/h_d^2 M(sp_a,sp_b) M(sq_a,sq_b)                      (same direction)
//   B_ab = 4V/h_d^2 [D(sp)M(sq)/h_p^2 + M(sp)D(sq)/h_q^2]         (same direction)
//   B_ab = -4V/(h_da^2 h_pa^2) L'(sp_a) M(sq_a, sp_b) L'(sq_b)     (db = pa; else swap a,b)
// with M(s,t) = 1/3 if s==t else 1/6, D(s,t) = 1 if s==t else -1, L'(s) = -1 for s=0, +1.
#include <hip/hip_runtime.h>

#include "assemble.hpp"

namespace emtd {
namespace {

__device__ __forceinline__ double m1(int s, int t) { return s == t ? (1.0 / 3.0) : (1.0 / 6.0); }
__device__ __forceinline__ double d1(int s, int t) { return s == t ? 1.0 : -1.0; }
__device__ __forceinline__ double lp(int s) { return s ? 1.0 : -1.0; }

// bit of local edge a along axis ax (ax transverse to the edge)
__device__ __forceinline__ int ebit(int a, int ax) {
    const int d = a >> 2, k = a & 3;
    return ax == (d + 1) % 3 ? (k & 1) : (k >> 1);
}

__device__ double local_c(int a, int b, const double h[3], double vol) {
    const int da = a >> 2, db = b >> 2;
    if (da != db) return 0.0;
    const int p = (da + 1) % 3, q = (da + 2) % 3;
    return 4.0 * vol / (h[da] * h[da]) * m1(ebit(a, p), ebit(b, p)) * m1(ebit(a, q), ebit(b, q));
}

__device__ double local_b(int a, int b, const double h[3], double vol) {
    const int da = a >> 2, db = b >> 2;
    if (da == db) {
        const int p = (da + 1) % 3, q = (da + 2) % 3;
        const int ap = ebit(a, p), bp = ebit(b, p), aq = ebit(a, q), bq = ebit(b, q);
        return 4.0 * vol / (h[da] * h[da]) *
               (d1(ap, bp) * m1(aq, bq) / (h[p] * h[p]) + m1(ap, bp) * d1(aq, bq) / (h[q] * h[q]));
    }
    if (db != (da + 1) % 3) {  // make b lie along a's first transverse axis
        const int t = a;
        a = b;
        b = t;
    }
    const int d = a >> 2, p = (d + 1) % 3, q = (d + 2) % 3;
    // b has transverse axes (q, d)
    return -4.0 * vol / (h[d] * h[d] * h[p] * h[p]) * lp(ebit(a, p)) * m1(ebit(a, q), ebit(b, q)) * lp(ebit(b, d));
}

__global__ void k_element_table(int kpar, const double* __restrict__ xyz, const int* __restrict__ nver,
                                const int* __restrict__ nvkat, const double* __restrict__ sigma,
                                const double* __restrict__ mu, ElemData* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kpar) return;
    const int n0 = nver[8 * e], n7 = nver[8 * e + 7];
    ElemData d;
    d.hx = xyz[3 * n7] - xyz[3 * n0];
    d.hy = xyz[3 * n7 + 1] - xyz[3 * n0 + 1];
    d.hz = xyz[3 * n7 + 2] - xyz[3 * n0 + 2];
    const int mat = nvkat[e];
    d.sigma = sigma[mat];
    d.mu = mu[mat];
    out[e] = d;
}

__global__ void k_assemble(int64_t nslots, const int64_t* __restrict__ cptr, const uint32_t* __restrict__ code,
                           const ElemData* __restrict__ el, double* __restrict__ bval, double* __restrict__ cval) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    double bs = 0.0, cs = 0.0;
    for (int64_t k = cptr[s]; k < cptr[s + 1]; ++k) {
        const uint32_t c = code[k];
        const uint32_t e = c / 144u, ab = c % 144u;
        const int a = int(ab / 12u), b = int(ab % 12u);
        const ElemData d = el[e];
        const double h[3] = {d.hx, d.hy, d.hz};
        const double vol = d.hx * d.hy * d.hz;
        bs += local_b(a, b, h, vol) / d.mu;
        cs += local_c(a, b, h, vol) * d.sigma;
    }
    bval[s] = bs;
    cval[s] = cs;
}

}  // namespace

void build_contributions(int kpar, const int* ed, int n, const int* ig, const int* jg, std::vector<int64_t>& cptr,
                         std::vector<uint32_t>& codes) {
    // slots: 0..nnz_low-1 lower entries, nnz_low..nnz_low+n-1 diagonal
    const int64_t nnz = ig[n];
    const int64_t ns = nnz + n;
    cptr.assign(ns + 1, 0);
    std::vector<int64_t> slot(size_t(kpar) * 78);
    int64_t w = 0;
    for (int e = 0; e < kpar; ++e) {
        const int* E = ed + 12 * e;
        for (int a = 0; a < 12; ++a) {
            const int ii = E[a];
            for (int b = 0; b < 12; ++b) {
                const int jj = E[b];
                int64_t s = -1;
                if (jj < ii) {
                    const int* lo = jg + ig[ii];
                    const int* hi = jg + ig[ii + 1];
                    const int* it = std::lower_bound(lo, hi, jj);
                    if (it == hi || *it != jj) throw Error(0x0004, "assembly: pattern lacks an element entry");
                    s = it - jg;
                } else if (jj == ii && a == b) {
                    s = nnz + ii;
                }
                if (s >= 0) cptr[s + 1]++;
            }
        }
        (void)w;
    }
    for (int64_t s = 0; s < ns; ++s) cptr[s + 1] += cptr[s];
    codes.resize(cptr[ns]);
    std::vector<int64_t> fill(cptr.begin(), cptr.end() - 1);
    for (int e = 0; e < kpar; ++e) {
        const int* E = ed + 12 * e;
        for (int a = 0; a < 12; ++a) {
            const int ii = E[a];
            for (int b = 0; b < 12; ++b) {
                const int jj = E[b];
                int64_t s = -1;
                if (jj < ii) {
                    const int* lo = jg + ig[ii];
                    const int* it = std::lower_bound(lo, jg + ig[ii + 1], jj);
                    s = it - jg;
                } else if (jj == ii && a == b) {
                    s = nnz + ii;
                }
                if (s >= 0) codes[fill[s]++] = uint32_t(e) * 144u + uint32_t(a * 12 + b);
            }
        }
    }
    (void)slot;
}

Assembler::Assembler(const emtd_mesh& m, int n, const int* ig, const int* jg, hipStream_t s) : stream(s) {
    if (int64_t(m.kpar) * 144 >= (int64_t(1) << 32)) throw Error(0x0004, "too many elements for the assembly code");
    n_ = n;
    nnz_ = ig[n];
    std::vector<int64_t> cptr;
    std::vector<uint32_t> codes;
    build_contributions(m.kpar, m.ed, n, ig, jg, cptr, codes);
    cptr_ = to_device(cptr, s);
    codes_ = to_device(codes, s);
    DevBuf<double> xyz(size_t(m.kuzlov) * 3), sig(m.n_materials), mu(m.n_materials);
    DevBuf<int> nver(size_t(m.kpar) * 8), nvkat(m.kpar);
    xyz.upload(m.xyz, size_t(m.kuzlov) * 3, s);
    sig.upload(m.sigma, m.n_materials, s);
    mu.upload(m.mu, m.n_materials, s);
    nver.upload(m.nver, size_t(m.kpar) * 8, s);
    nvkat.upload(m.nvkat, m.kpar, s);
    el_.alloc(m.kpar);
    hipLaunchKernelGGL(k_element_table, dim3((m.kpar + 255) / 256), dim3(256), 0, s, m.kpar, xyz.p, nver.p, nvkat.p,
                       sig.p, mu.p, el_.p);
    EMTD_HIP(hipGetLastError());
    EMTD_HIP(hipStreamSynchronize(s));
}

void Assembler::run(double* d_b_ggl, double* d_b_di, double* d_c_ggl, double* d_c_di) {
    // the slot space is [lower | diagonal]; kernel writes into one contiguous scratch then splits
    const int64_t ns = nnz_ + n_;
    if (bscratch_.n < size_t(ns)) { bscratch_.alloc(ns); cscratch_.alloc(ns); }
    hipLaunchKernelGGL(k_assemble, dim3(unsigned((ns + 255) / 256)), dim3(256), 0, stream, ns, cptr_.p, codes_.p,
                       el_.p, bscratch_.p, cscratch_.p);
    EMTD_HIP(hipGetLastError());
    EMTD_HIP(hipMemcpyAsync(d_b_ggl, bscratch_.p, nnz_ * sizeof(double), hipMemcpyDeviceToDevice, stream));
    EMTD_HIP(hipMemcpyAsync(d_c_ggl, cscratch_.p, nnz_ * sizeof(double), hipMemcpyDeviceToDevice, stream));
    EMTD_HIP(hipMemcpyAsync(d_b_di, bscratch_.p + nnz_, n_ * sizeof(double), hipMemcpyDeviceToDevice, stream));
    EMTD_HIP(hipMemcpyAsync(d_c_di, cscratch_.p + nnz_, n_ * sizeof(double), hipMemcpyDeviceToDevice, stream));
}

double Assembler::run_timed(double* b_all, double* c_all) {
    hipEvent_t e0, e1;
    EMTD_HIP(hipEventCreate(&e0));
    EMTD_HIP(hipEventCreate(&e1));
    const int64_t ns = nnz_ + n_;
    EMTD_HIP(hipEventRecord(e0, stream));
    hipLaunchKernelGGL(k_assemble, dim3(unsigned((ns + 255) / 256)), dim3(256), 0, stream, ns, cptr_.p, codes_.p,
                       el_.p, b_all, c_all);
    EMTD_HIP(hipEventRecord(e1, stream));
    EMTD_HIP(hipEventSynchronize(e1));
    float ms = 0;
    EMTD_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms * 1e-3;
}

}  // namespace emtd
