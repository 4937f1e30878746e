// GPU supernodal multifrontal Cholesky for CDNA4 (gfx950).
//
// Replaces MKL PARDISO's numeric factorization (phase 12) and solve (phase 33) called at
// A2/pardiso.cpp:104-119. Fronts are processed level by level (all fronts of a level are
// independent); within a level every kernel is batched over the level's fronts:
//
//   scatter    system entries  -> L storage            (one thread per matrix entry)
//   extend-add child update U_c -> parent front          (one wavefront per update column)
//   potrf      diagonal NB x NB block of each front      (one workgroup per front, LDS)
//   trsm       panel rows below the diagonal block       (one wavefront per 64 rows)
//   update     trailing lower matrix -= panel * panel^T  (64x64 tiles, MFMA 16x16x4)
//
// Storage per front f with p pivots, u update rows, m = p + u: L_f is m x p column-major
// (ld = m) and holds the factor columns; U_f is u x u column-major and receives the Schur
// complement passed to the parent.
//
// Triangular solves (per time step) are multifrontal too: the forward sweep gathers the
// children's update vectors into the front, solves with L11 and leaves w = v_U - L21 y for
// the parent; the backward sweep reads the already final ancestors' values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>

#include "chol.hpp"

namespace emtd {

namespace {

constexpr int NB = 32;    // panel width of the blocked factorization
constexpr int TILE = 64;  // trailing-update tile
constexpr int SB = 64;    // column block of the triangular sweeps

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// L[dest] = mult*C + B for the lower entries; diagonal gets the 1e-12 relative bump
// (A2/time_approx.cpp:987-994).
template <typename T>
__global__ void k_scatter_system(const int64_t* __restrict__ dest, const double* __restrict__ bg,
                                 const double* __restrict__ cg, int64_t nnz, double mult, T* __restrict__ L) {
    int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    const int64_t d = dest[s];
    if (d >= 0) L[d] = T(cg[s] * mult + bg[s]);
}

template <typename T>
__global__ void k_scatter_system_diag(const int64_t* __restrict__ dest, const double* __restrict__ bd,
                                      const double* __restrict__ cd, int n, double mult, T* __restrict__ L) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t d = dest[i];
    if (d < 0) return;
    double v = cd[i] * mult + bd[i];
    v += v * 1e-12;
    L[d] = T(v);
}

template <typename T>
__global__ void k_scatter_values(const int64_t* __restrict__ dest, const double* __restrict__ val, int64_t nnz,
                                 T* __restrict__ L) {
    int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    const int64_t d = dest[s];
    if (d >= 0) L[d] = T(val[s]);
}

template <typename T>
__device__ __forceinline__ T* front_elem(const DevFront& F, T* L, T* U, int r, int c) {
    // lower element (r >= c) of the frontal matrix
    if (c < F.p) return L + F.loff + int64_t(c) * F.m + r;
    const int u = F.m - F.p;
    return U + F.uoff + int64_t(c - F.p) * u + (r - F.p);
}

// Extend-add: one wavefront adds one column of a child's update matrix into its parent.
template <typename T>
__global__ void k_extend_add(const DevFront* __restrict__ F, const int2* __restrict__ items, int nitems,
                             const int* __restrict__ relmap, T* __restrict__ L, T* __restrict__ U) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int it = blockIdx.x * (blockDim.x >> 6) + wave;
    if (it >= nitems) return;
    const int c = items[it].x, b = items[it].y;
    const DevFront Fc = F[c];
    const DevFront Ff = F[Fc.parent];
    const int uc = Fc.m - Fc.p;
    const int* rm = relmap + Fc.relmap_off;
    const int tb = rm[b];
    const T* src = U + Fc.uoff + int64_t(b) * uc;
    if (tb < Ff.p) {
        T* dst = L + Ff.loff + int64_t(tb) * Ff.m;
        for (int a = b + lane; a < uc; a += 64) dst[rm[a]] += src[a];
    } else {
        const int uf = Ff.m - Ff.p;
        T* dst = U + Ff.uoff + int64_t(tb - Ff.p) * uf - Ff.p;
        for (int a = b + lane; a < uc; a += 64) dst[rm[a]] += src[a];
    }
}

// Unblocked Cholesky of the diagonal block [kb, kb+nb) of each active front.
template <typename T>
__global__ void k_potrf(const DevFront* __restrict__ F, const int* __restrict__ flist, int kb,
                        T* __restrict__ L, int* __restrict__ info) {
    const DevFront Fr = F[flist[blockIdx.x]];
    if (kb >= Fr.p) return;
    const int nb = min(NB, Fr.p - kb);
    __shared__ T A[NB][NB + 1];
    __shared__ int bad;
    T* base = L + Fr.loff + int64_t(kb) * Fr.m + kb;
    for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
        const int i = idx % nb, j = idx / nb;
        if (i >= j) A[i][j] = base[int64_t(j) * Fr.m + i];
    }
    if (threadIdx.x == 0) bad = -1;
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        if (threadIdx.x == 0) {
            T d = A[j][j];
            if (!(d > T(0))) {
                if (bad < 0) bad = j;
                d = T(1);
            }
            A[j][j] = sqrt(d);
        }
        __syncthreads();
        const T dj = A[j][j];
        for (int i = j + 1 + threadIdx.x; i < nb; i += blockDim.x) A[i][j] /= dj;
        __syncthreads();
        const int w = nb - j - 1;
        for (int idx = threadIdx.x; idx < w * w; idx += blockDim.x) {
            const int i = j + 1 + idx % w, l = j + 1 + idx / w;
            if (l <= i) A[i][l] -= A[i][j] * A[l][j];
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
        const int i = idx % nb, j = idx / nb;
        if (i >= j) base[int64_t(j) * Fr.m + i] = A[i][j];
    }
    if (threadIdx.x == 0 && bad >= 0) atomicMin(info, Fr.fbeg + kb + bad);
}

// Panel rows below the diagonal block: X * L_kk^T = B, 64 rows per wavefront.
template <typename T>
__global__ void k_trsm(const DevFront* __restrict__ F, const int* __restrict__ flist, int kb, T* __restrict__ L) {
    const DevFront Fr = F[flist[blockIdx.y]];
    if (kb >= Fr.p) return;
    const int nb = min(NB, Fr.p - kb);
    const int r0 = kb + nb + 64 * blockIdx.x;
    if (r0 >= Fr.m) return;
    __shared__ T Lkk[NB][NB + 1];
    __shared__ T X[64][NB + 1];
    T* col = L + Fr.loff + int64_t(kb) * Fr.m;
    for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
        const int i = idx % nb, j = idx / nb;
        if (i >= j) Lkk[i][j] = col[int64_t(j) * Fr.m + kb + i];
    }
    const int t = threadIdx.x;
    const int r = r0 + t;
    if (r < Fr.m)
        for (int j = 0; j < nb; ++j) X[t][j] = col[int64_t(j) * Fr.m + r];
    __syncthreads();
    if (r < Fr.m) {
        for (int j = 0; j < nb; ++j) {
            T s = X[t][j];
            for (int l = 0; l < j; ++l) s -= X[t][l] * Lkk[j][l];
            X[t][j] = s / Lkk[j][j];
        }
        for (int j = 0; j < nb; ++j) col[int64_t(j) * Fr.m + r] = X[t][j];
    }
}

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef v4d acc_t;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<float> {
    typedef v4f acc_t;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// Trailing update of the lower frontal matrix: F[r][c] -= sum_k P[r][k] P[c][k] for
// r >= c >= kb+nb, in 64x64 tiles of 4 wavefronts; each wavefront owns a 32x32 quarter as
// 2x2 MFMA 16x16x4 blocks (A lane l: row l%16, k l/16; C lane l: rows 4(l/16)..+3, col l%16).
template <typename T>
__global__ void __launch_bounds__(256) k_update(const DevFront* __restrict__ F, const int* __restrict__ flist,
                                                int kb, T* __restrict__ L, T* __restrict__ U) {
    const DevFront Fr = F[flist[blockIdx.y]];
    if (kb >= Fr.p) return;
    const int nb = min(NB, Fr.p - kb);
    const int base = kb + nb;
    const int rem = Fr.m - base;
    if (rem <= 0) return;
    const int nt = (rem + TILE - 1) / TILE;
    const int t = blockIdx.x;
    if (t >= nt * (nt + 1) / 2) return;
    int ti = int((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while (ti * (ti + 1) / 2 > t) --ti;
    while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
    const int tj = t - ti * (ti + 1) / 2;
    const int R0 = base + TILE * ti, C0 = base + TILE * tj;
    constexpr int KP = NB + 4;  // padded K extent (multiple of 4) + bank offset
    __shared__ T As[TILE][KP];
    __shared__ T Bs[TILE][KP];
    const T* pan = L + Fr.loff + int64_t(kb) * Fr.m;
    const int nb4 = (nb + 3) & ~3;
    for (int idx = threadIdx.x; idx < TILE * nb4; idx += blockDim.x) {
        const int i = idx % TILE, k = idx / TILE;
        const int ra = R0 + i, rb = C0 + i;
        As[i][k] = (k < nb && ra < Fr.m) ? pan[int64_t(k) * Fr.m + ra] : T(0);
        Bs[i][k] = (k < nb && rb < Fr.m) ? pan[int64_t(k) * Fr.m + rb] : T(0);
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    if (ti == tj && wr + 31 < wc) return;  // strictly upper quarter of a diagonal tile
    typename Mfma<T>::acc_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) acc[a][b][v] = T(0);
    const int li = lane & 15, lk = lane >> 4;
    for (int k0 = 0; k0 < nb4; k0 += 4) {
        T av[2], bv[2];
        for (int a = 0; a < 2; ++a) av[a] = As[wr + 16 * a + li][k0 + lk];
        for (int b = 0; b < 2; ++b) bv[b] = Bs[wc + 16 * b + li][k0 + lk];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) acc[a][b] = Mfma<T>::op(av[a], bv[b], acc[a][b]);
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                const int r = R0 + wr + 16 * a + 4 * lk + v;
                const int c = C0 + wc + 16 * b + li;
                if (r < Fr.m && c < Fr.m && r >= c) *front_elem(Fr, L, U, r, c) -= acc[a][b][v];
            }
}

// ------------------------------------------------------------------ triangular sweeps
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Forward sweep of one level: one workgroup per front and right-hand side.
template <typename T>
__global__ void __launch_bounds__(256) k_fwd(const DevFront* __restrict__ F, const int* __restrict__ flist,
                                             const int* __restrict__ children, const int* __restrict__ relmap,
                                             const T* __restrict__ L, double* __restrict__ W, int64_t wsize,
                                             const double* __restrict__ b, double* __restrict__ x, int64_t ld) {
    const DevFront Fr = F[flist[blockIdx.x]];
    const int rhs = blockIdx.y;
    double* v = W + rhs * wsize + Fr.woff;
    const double* bb = b + rhs * ld;
    double* xx = x + rhs * ld;
    const int p = Fr.p, m = Fr.m;
    for (int i = threadIdx.x; i < m; i += blockDim.x) v[i] = i < p ? bb[Fr.fbeg + i] : 0.0;
    __syncthreads();
    for (int ci = Fr.child_beg; ci < Fr.child_end; ++ci) {
        const DevFront Fc = F[children[ci]];
        const int uc = Fc.m - Fc.p;
        const int* rm = relmap + Fc.relmap_off;
        const double* wc = W + rhs * wsize + Fc.woff + Fc.p;
        for (int k = threadIdx.x; k < uc; k += blockDim.x) v[rm[k]] += wc[k];
        __syncthreads();
    }
    __shared__ double y[SB];
    const T* Lf = L + Fr.loff;
    const int lane = threadIdx.x & 63;
    for (int kb = 0; kb < p; kb += SB) {
        const int nb = min(SB, p - kb);
        if (threadIdx.x < 64) {
            double val = lane < nb ? v[kb + lane] : 0.0;
            for (int j = 0; j < nb; ++j) {
                const double yj = __shfl(val, j, 64) / double(Lf[int64_t(kb + j) * m + kb + j]);
                if (lane == j) y[j] = yj;
                if (lane > j && lane < nb) val -= double(Lf[int64_t(kb + j) * m + kb + lane]) * yj;
            }
            if (lane < nb) v[kb + lane] = y[lane];
        }
        __syncthreads();
        for (int r = kb + nb + threadIdx.x; r < m; r += blockDim.x) {
            double acc = v[r];
            const T* lr = Lf + int64_t(kb) * m + r;
            for (int j = 0; j < nb; ++j) acc -= double(lr[int64_t(j) * m]) * y[j];
            v[r] = acc;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < p; i += blockDim.x) xx[Fr.fbeg + i] = v[i];
}

// Backward sweep of one level (levels visited top-down).
template <typename T>
__global__ void __launch_bounds__(256) k_bwd(const DevFront* __restrict__ F, const int* __restrict__ flist,
                                             const int* __restrict__ rows, const T* __restrict__ L,
                                             double* __restrict__ W, int64_t wsize, double* __restrict__ x,
                                             int64_t ld) {
    const DevFront Fr = F[flist[blockIdx.x]];
    const int rhs = blockIdx.y;
    double* v = W + rhs * wsize + Fr.woff;
    double* xx = x + rhs * ld;
    const int p = Fr.p, m = Fr.m;
    const int* rw = rows + Fr.rows_off;
    for (int i = threadIdx.x; i < m; i += blockDim.x) v[i] = i < p ? xx[Fr.fbeg + i] : xx[rw[i]];
    __syncthreads();
    __shared__ double s[SB];
    const T* Lf = L + Fr.loff;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nwaves = blockDim.x >> 6;
    for (int kb = ((p - 1) / SB) * SB; kb >= 0; kb -= SB) {
        const int nb = min(SB, p - kb);
        for (int j = wave; j < nb; j += nwaves) {
            const T* lc = Lf + int64_t(kb + j) * m;
            double part = 0.0;
            for (int r = kb + nb + lane; r < m; r += 64) part += double(lc[r]) * v[r];
            part = wave_sum(part);
            if (lane == 0) s[j] = part;
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            double val = lane < nb ? v[kb + lane] - s[lane] : 0.0;
            for (int j = nb - 1; j >= 0; --j) {
                const double xj = __shfl(val, j, 64) / double(Lf[int64_t(kb + j) * m + kb + j]);
                if (lane == j) s[j] = xj;
                if (lane < j) val -= double(Lf[int64_t(kb + lane) * m + kb + j]) * xj;
            }
            if (lane < nb) v[kb + lane] = s[lane];
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < p; i += blockDim.x) xx[Fr.fbeg + i] = v[i];
}

inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    return unsigned(std::max<int64_t>(1, std::min<int64_t>(g, 1 << 20)));
}

}  // namespace

template <typename T>
GpuCholesky<T>::GpuCholesky(const Symbolic& S, hipStream_t stream) : S_(S), stream_(stream) {
    std::vector<DevFront> fr(S.nf);
    for (int f = 0; f < S.nf; ++f) {
        DevFront& d = fr[f];
        d.loff = S.loff[f];
        d.uoff = S.uoff[f];
        d.woff = S.woff[f];
        d.rows_off = S.rows_ptr[f];
        d.relmap_off = S.relmap_ptr[f];
        d.fbeg = S.fbeg[f];
        d.p = S.pcount[f];
        d.m = S.mcount[f];
        d.parent = S.parent[f];
        d.child_beg = S.child_ptr[f];
        d.child_end = S.child_ptr[f + 1];
    }
    fronts_ = to_device(fr, stream);
    rows_ = to_device(S.rows, stream);
    children_ = to_device(S.children, stream);
    relmap_ = to_device(S.relmap, stream);
    level_fronts_ = to_device(S.level_fronts, stream);
    L_.alloc(std::max<int64_t>(S.lsize, 1));
    U_.alloc(std::max<int64_t>(S.usize, 1));
    info_.alloc(1);
    // levels with fronts sorted by decreasing pivot count (active prefix per panel step)
    std::vector<int> sorted;
    level_sorted_off_.assign(S.nlevels + 1, 0);
    for (int l = 0; l < S.nlevels; ++l) {
        std::vector<int> fl(S.level_fronts.begin() + S.level_ptr[l], S.level_fronts.begin() + S.level_ptr[l + 1]);
        std::stable_sort(fl.begin(), fl.end(), [&](int a, int b) { return S.pcount[a] > S.pcount[b]; });
        sorted.insert(sorted.end(), fl.begin(), fl.end());
        level_sorted_off_[l + 1] = int(sorted.size());
        level_sorted_.push_back(fl);
    }
    level_sorted_dev_ = to_device(sorted, stream);
    // extend-add work items per (level, pass)
    const int P = std::max(1, S.max_children);
    std::vector<int2> items;
    ea_ptr_.assign(S.nlevels * P + 1, 0);
    for (int l = 0; l < S.nlevels; ++l)
        for (int k = 0; k < P; ++k) {
            for (int li = S.level_ptr[l]; li < S.level_ptr[l + 1]; ++li) {
                const int f = S.level_fronts[li];
                if (S.child_ptr[f] + k >= S.child_ptr[f + 1]) continue;
                const int c = S.children[S.child_ptr[f] + k];
                const int uc = S.mcount[c] - S.pcount[c];
                for (int b = 0; b < uc; ++b) items.push_back(make_int2(c, b));
            }
            ea_ptr_[l * P + k + 1] = int(items.size());
        }
    ea_items_ = to_device(items, stream);
    EMTD_HIP(hipStreamSynchronize(stream));
}

template <typename T>
void GpuCholesky<T>::numeric() {
    const Symbolic& S = S_;
    const int P = std::max(1, S.max_children);
    for (int l = 0; l < S.nlevels; ++l) {
        if (l > 0)
            for (int k = 0; k < P; ++k) {
                const int b = ea_ptr_[l * P + k], e = ea_ptr_[l * P + k + 1];
                if (e > b)
                    hipLaunchKernelGGL(k_extend_add<T>, dim3((e - b + 3) / 4), dim3(256), 0, stream_,
                                       fronts_.p, ea_items_.p + b, e - b, relmap_.p, L_.p, U_.p);
            }
        const std::vector<int>& fl = level_sorted_[l];
        const int* dfl = level_sorted_dev_.p + level_sorted_off_[l];
        const int maxp = S.pcount[fl[0]];
        int maxm = 0;
        for (int f : fl) maxm = std::max(maxm, S.mcount[f]);
        int nact = int(fl.size());
        for (int kb = 0; kb < maxp; kb += NB) {
            while (nact > 0 && S.pcount[fl[nact - 1]] <= kb) --nact;
            if (nact == 0) break;
            int mm = 0;
            for (int i = 0; i < nact; ++i) mm = std::max(mm, S.mcount[fl[i]]);
            hipLaunchKernelGGL(k_potrf<T>, dim3(nact), dim3(256), 0, stream_, fronts_.p, dfl, kb, L_.p, info_.p);
            const int rem = mm - kb - 1;
            if (rem > 0) {
                const int chunks = (rem + 63) / 64;
                hipLaunchKernelGGL(k_trsm<T>, dim3(chunks, nact), dim3(64), 0, stream_, fronts_.p, dfl, kb, L_.p);
                const int nt = (rem + TILE - 1) / TILE;
                hipLaunchKernelGGL(k_update<T>, dim3(nt * (nt + 1) / 2, nact), dim3(256), 0, stream_, fronts_.p,
                                   dfl, kb, L_.p, U_.p);
            }
        }
    }
    EMTD_HIP(hipGetLastError());
}

template <typename T>
void GpuCholesky<T>::factorize_system(const int64_t* d_dest_off, const int64_t* d_dest_diag, int64_t nnz_low,
                                      int n, const double* d_b_ggl, const double* d_b_di, const double* d_c_ggl,
                                      const double* d_c_di, double mult) {
    hipEvent_t e0, e1;
    EMTD_HIP(hipEventCreate(&e0));
    EMTD_HIP(hipEventCreate(&e1));
    EMTD_HIP(hipEventRecord(e0, stream_));
    const int init = INT_MAX;
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, stream_, info_.p, int64_t(1), init);
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(S_.lsize, 256)), dim3(256), 0, stream_, L_.p, S_.lsize, T(0));
    if (S_.usize)
        hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(S_.usize, 256)), dim3(256), 0, stream_, U_.p, S_.usize, T(0));
    if (nnz_low)
        hipLaunchKernelGGL(k_scatter_system<T>, dim3(unsigned((nnz_low + 255) / 256)), dim3(256), 0, stream_,
                           d_dest_off, d_b_ggl, d_c_ggl, nnz_low, mult, L_.p);
    hipLaunchKernelGGL(k_scatter_system_diag<T>, dim3((n + 255) / 256), dim3(256), 0, stream_, d_dest_diag, d_b_di,
                       d_c_di, n, mult, L_.p);
    numeric();
    EMTD_HIP(hipEventRecord(e1, stream_));
    EMTD_HIP(hipEventSynchronize(e1));
    float ms = 0;
    EMTD_HIP(hipEventElapsedTime(&ms, e0, e1));
    factor_seconds_ = ms * 1e-3;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

template <typename T>
void GpuCholesky<T>::factorize_values(const int64_t* d_dest_off, const int64_t* d_dest_diag, int64_t nnz_low,
                                      int n, const double* d_ggl, const double* d_di) {
    hipEvent_t e0, e1;
    EMTD_HIP(hipEventCreate(&e0));
    EMTD_HIP(hipEventCreate(&e1));
    EMTD_HIP(hipEventRecord(e0, stream_));
    const int init = INT_MAX;
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, stream_, info_.p, int64_t(1), init);
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(S_.lsize, 256)), dim3(256), 0, stream_, L_.p, S_.lsize, T(0));
    if (S_.usize)
        hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(S_.usize, 256)), dim3(256), 0, stream_, U_.p, S_.usize, T(0));
    if (nnz_low)
        hipLaunchKernelGGL(k_scatter_values<T>, dim3(unsigned((nnz_low + 255) / 256)), dim3(256), 0, stream_,
                           d_dest_off, d_ggl, nnz_low, L_.p);
    hipLaunchKernelGGL(k_scatter_values<T>, dim3((n + 255) / 256), dim3(256), 0, stream_, d_dest_diag, d_di,
                       int64_t(n), L_.p);
    numeric();
    EMTD_HIP(hipEventRecord(e1, stream_));
    EMTD_HIP(hipEventSynchronize(e1));
    float ms = 0;
    EMTD_HIP(hipEventElapsedTime(&ms, e0, e1));
    factor_seconds_ = ms * 1e-3;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

template <typename T>
int GpuCholesky<T>::check_info() {
    int h = INT_MAX;
    EMTD_HIP(hipMemcpyAsync(&h, info_.p, sizeof(int), hipMemcpyDeviceToHost, stream_));
    EMTD_HIP(hipStreamSynchronize(stream_));
    return h == INT_MAX ? -1 : h;
}

template <typename T>
void GpuCholesky<T>::solve(const double* d_b, double* d_x, int nrhs, int64_t ld) {
    const Symbolic& S = S_;
    if (nrhs > wrhs_) {
        W_.alloc(size_t(std::max<int64_t>(S.wsize, 1)) * nrhs);
        wrhs_ = nrhs;
    }
    for (int l = 0; l < S.nlevels; ++l) {
        const int nfl = S.level_ptr[l + 1] - S.level_ptr[l];
        hipLaunchKernelGGL(k_fwd<T>, dim3(nfl, nrhs), dim3(256), 0, stream_, fronts_.p,
                           level_fronts_.p + S.level_ptr[l], children_.p, relmap_.p, L_.p, W_.p, S.wsize, d_b, d_x,
                           ld);
    }
    for (int l = S.nlevels - 1; l >= 0; --l) {
        const int nfl = S.level_ptr[l + 1] - S.level_ptr[l];
        hipLaunchKernelGGL(k_bwd<T>, dim3(nfl, nrhs), dim3(256), 0, stream_, fronts_.p,
                           level_fronts_.p + S.level_ptr[l], rows_.p, L_.p, W_.p, S.wsize, d_x, ld);
    }
    EMTD_HIP(hipGetLastError());
}

template class GpuCholesky<double>;
template class GpuCholesky<float>;

}  // namespace emtd
