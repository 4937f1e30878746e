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
// Operands of both instructions: lane l carries A[l%16][l/16] and B[l/16][l%16]. The result
// layouts differ (checked on gfx950 with an in-bounds probe): accumulator element v of lane l
// is D[row(l/16, v)][l%16] with row = l/16 + 4v for f64 and 4(l/16) + v for f32.
template <>
struct Mfma<double> {
    typedef v4d acc_t;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct Mfma<float> {
    typedef v4f acc_t;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};

// Symmetric rank-k update of the lower frontal matrix with the factored columns
// [k0, k0+kd), kd = min(kmax, p-k0):  F[r][c] -= sum_k L[r][k] L[c][k]  for r >= c in the
// target columns [k0+kd, ce), where ce = min(kb+NBO, p) (mode 0: update inside the current
// 128-column block) or m (mode 1: the trailing matrix). 64x64 tiles (blockIdx.x = row tile,
// blockIdx.y = column tile, blockIdx.z = front); 4 wavefronts own a 32x32 quarter each as
// 2x2 MFMA 16x16x4 blocks (lane layouts in Mfma<T>);
// K is streamed through LDS in chunks of 32 and the tile is read and written once.
constexpr int NBO = 128;  // outer block of the factorization

template <typename T>
__global__ void __launch_bounds__(256) k_syrk(const DevFront* __restrict__ F, const int* __restrict__ flist, int kb,
                                              int k0, int kmax, int mode, T* __restrict__ L, T* __restrict__ U) {
    const DevFront Fr = F[flist[blockIdx.z]];
    const int kd = min(kmax, Fr.p - k0);
    if (kd <= 0) return;
    const int cs = k0 + kd;
    const int ce = mode == 0 ? min(kb + NBO, Fr.p) : Fr.m;
    if (cs >= ce) return;
    const int ti = blockIdx.x, tj = blockIdx.y;
    if (ti < tj) return;
    const int R0 = cs + TILE * ti, C0 = cs + TILE * tj;
    if (R0 >= Fr.m || C0 >= ce) return;
    constexpr int KC = 32, KP = KC + 4;
    __shared__ T As[TILE][KP];
    __shared__ T Bs[TILE][KP];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const bool skip = (ti == tj && wr + 31 < wc);  // strictly upper quarter of a diagonal tile
    typename Mfma<T>::acc_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) acc[a][b][v] = T(0);
    const int li = lane & 15, lk = lane >> 4;
    for (int kc = 0; kc < kd; kc += KC) {
        const int nk = min(KC, kd - kc);
        const T* pan = L + Fr.loff + int64_t(k0 + kc) * Fr.m;
        __syncthreads();
        for (int idx = threadIdx.x; idx < TILE * KC; idx += blockDim.x) {
            const int i = idx % TILE, k = idx / TILE;
            const int ra = R0 + i, rb = C0 + i;
            As[i][k] = (k < nk && ra < Fr.m) ? pan[int64_t(k) * Fr.m + ra] : T(0);
            Bs[i][k] = (k < nk && rb < ce) ? pan[int64_t(k) * Fr.m + rb] : T(0);
        }
        __syncthreads();
        if (!skip) {
            const int nk4 = (nk + 3) & ~3;
            for (int kk = 0; kk < nk4; kk += 4) {
                T av[2], bv[2];
                for (int a = 0; a < 2; ++a) av[a] = As[wr + 16 * a + li][kk + lk];
                for (int b = 0; b < 2; ++b) bv[b] = Bs[wc + 16 * b + li][kk + lk];
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) acc[a][b] = Mfma<T>::op(av[a], bv[b], acc[a][b]);
            }
        }
    }
    if (skip) return;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                const int r = R0 + wr + 16 * a + Mfma<T>::row(lk, v);
                const int c = C0 + wc + 16 * b + li;
                if (r < Fr.m && c < ce && r >= c) *front_elem(Fr, L, U, r, c) -= acc[a][b][v];
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

// ------------------------------------------------------------------ sweeps of large fronts
// Levels with few, large fronts split every front over many workgroups: 64-row blocks in
// the forward sweep, 64-column pivot blocks in the backward sweep. A workgroup takes a ticket
// (atomic counter) and processes items in ticket order, so it only ever waits for blocks held
// by workgroups that started earlier; results of a pivot block are published through a
// per-block buffer and flag (release/acquire at agent scope). The wait is bounded: after
// kSpinLimit polls the workgroup records an error and continues, so the kernel always ends.
constexpr int kSpinLimit = 1 << 22;

__device__ __forceinline__ void wait_flag(const int* flag, int* err) {
    int spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > kSpinLimit) {
            atomicOr(err, 1);
            break;
        }
    }
}

__device__ __forceinline__ void publish_flag(int* flag) {
    __threadfence();
    __hip_atomic_store(flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// v_f = [b_P ; 0] + children's update vectors (one workgroup per front)
__global__ void __launch_bounds__(256) k_fwd_gather(const DevFront* __restrict__ F, const int* __restrict__ flist,
                                                    const int* __restrict__ children, const int* __restrict__ relmap,
                                                    double* __restrict__ W, const double* __restrict__ b) {
    const DevFront Fr = F[flist[blockIdx.x]];
    double* v = W + Fr.woff;
    for (int i = threadIdx.x; i < Fr.m; i += blockDim.x) v[i] = i < Fr.p ? b[Fr.fbeg + i] : 0.0;
    __syncthreads();
    for (int ci = Fr.child_beg; ci < Fr.child_end; ++ci) {
        const DevFront Fc = F[children[ci]];
        const int uc = Fc.m - Fc.p;
        const int* rm = relmap + Fc.relmap_off;
        const double* wc = W + Fc.woff + Fc.p;
        for (int k = threadIdx.x; k < uc; k += blockDim.x) v[rm[k]] += wc[k];
        __syncthreads();
    }
}

// v_f = [y_P ; x_U] for the backward sweep (one workgroup per front)
__global__ void __launch_bounds__(256) k_bwd_gather(const DevFront* __restrict__ F, const int* __restrict__ flist,
                                                    const int* __restrict__ rows, double* __restrict__ W,
                                                    const double* __restrict__ x) {
    const DevFront Fr = F[flist[blockIdx.x]];
    double* v = W + Fr.woff;
    const int* rw = rows + Fr.rows_off;
    for (int i = threadIdx.x; i < Fr.m; i += blockDim.x) v[i] = i < Fr.p ? x[Fr.fbeg + i] : x[rw[i]];
}

// item.x = front, item.y = 64-row block; flag_off per front in foff[]
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_big(const DevFront* __restrict__ F, const int2* __restrict__ items,
                                                 const int* __restrict__ foff, int* __restrict__ flags,
                                                 double* __restrict__ ybuf, int* __restrict__ counter,
                                                 int* __restrict__ err, const T* __restrict__ L,
                                                 double* __restrict__ W, double* __restrict__ x) {
    __shared__ int s_ticket;
    __shared__ double red[4][64];
    __shared__ double ys[64];
    if (threadIdx.x == 0) s_ticket = atomicAdd(counter, 1);
    __syncthreads();
    const int2 it = items[s_ticket];
    const DevFront Fr = F[it.x];
    const int rb = it.y;
    const int p = Fr.p, m = Fr.m;
    const int np = (p + 63) / 64;
    const int r0 = 64 * rb;
    const int lr = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int r = r0 + lr;
    const T* Lf = L + Fr.loff;
    double* v = W + Fr.woff;
    int* fl = flags + foff[it.x];
    double* yb = ybuf + int64_t(foff[it.x]) * 64;
    const int ncb = (r0 < p) ? rb : np;
    double acc = 0.0;
    for (int cb = 0; cb < ncb; ++cb) {
        if (threadIdx.x == 0) wait_flag(fl + cb, err);
        __syncthreads();
        const int nbc = min(64, p - 64 * cb);
        double yv[16];
        for (int j = 0; j < 16; ++j) {
            const int jj = 16 * q + j;
            yv[j] = jj < nbc ? __hip_atomic_load(yb + 64 * cb + jj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
        }
        if (r < m) {
            const T* lc = Lf + int64_t(64 * cb + 16 * q) * m + r;
            for (int j = 0; j < 16; ++j)
                if (16 * q + j < nbc) acc += double(lc[int64_t(j) * m]) * yv[j];
        }
    }
    red[q][lr] = acc;
    __syncthreads();
    double rowv = 0.0;
    if (q == 0 && r < m) rowv = v[r] - (red[0][lr] + red[1][lr] + red[2][lr] + red[3][lr]);
    if (r0 < p) {
        const int nbp = min(64, p - r0);
        if (q == 0) {
            double val = rowv;
            for (int j = 0; j < nbp; ++j) {
                const double yj = __shfl(val, j, 64) / double(Lf[int64_t(r0 + j) * m + r0 + j]);
                if (lr == j) ys[j] = yj;
                if (lr > j && lr < nbp) val -= double(Lf[int64_t(r0 + j) * m + r0 + lr]) * yj;
            }
            if (lr < nbp) {
                const double y = ys[lr];
                __hip_atomic_store(yb + 64 * rb + lr, y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                v[r] = y;
                x[Fr.fbeg + r] = y;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) publish_flag(fl + rb);
        // update rows sharing this block: subtract the block's own contribution
        if (q == 0 && lr >= nbp && r < m) {
            double a2 = rowv;
            for (int j = 0; j < nbp; ++j) a2 -= double(Lf[int64_t(r0 + j) * m + r]) * ys[j];
            v[r] = a2;
        }
    } else if (q == 0 && r < m) {
        v[r] = rowv;
    }
}

// item.x = front, item.y = 64-column pivot block (items in decreasing block order per front)
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_big(const DevFront* __restrict__ F, const int2* __restrict__ items,
                                                 const int* __restrict__ foff, int* __restrict__ flags,
                                                 double* __restrict__ xbuf, int* __restrict__ counter,
                                                 int* __restrict__ err, const T* __restrict__ L,
                                                 double* __restrict__ W, double* __restrict__ x) {
    __shared__ int s_ticket;
    __shared__ double sred[64];
    __shared__ double xs[64];
    __shared__ double xr[64];
    if (threadIdx.x == 0) s_ticket = atomicAdd(counter, 1);
    __syncthreads();
    const int2 it = items[s_ticket];
    const DevFront Fr = F[it.x];
    const int cb = it.y;
    const int p = Fr.p, m = Fr.m;
    const int np = (p + 63) / 64;
    const int c0 = 64 * cb;
    const int nbc = min(64, p - c0);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const T* Lf = L + Fr.loff;
    double* v = W + Fr.woff;
    int* fl = flags + foff[it.x];
    double* xb = xbuf + int64_t(foff[it.x]) * 64;
    double part[16];
    for (int j = 0; j < 16; ++j) part[j] = 0.0;
    // update rows first (their values are final), in 64-row chunks
    for (int rr = p; rr < m; rr += 64) {
        const int r = rr + lane;
        const double xv = r < m ? v[r] : 0.0;
        if (r < m)
            for (int j = 0; j < 16; ++j) {
                const int jj = wave + 4 * j;
                if (jj < nbc) part[j] += double(Lf[int64_t(c0 + jj) * m + r]) * xv;
            }
    }
    // later pivot blocks, newest first
    for (int cb2 = np - 1; cb2 > cb; --cb2) {
        if (threadIdx.x == 0) wait_flag(fl + cb2, err);
        __syncthreads();
        const int nb2 = min(64, p - 64 * cb2);
        if (threadIdx.x < 64)
            xr[lane] = lane < nb2 ? __hip_atomic_load(xb + 64 * cb2 + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : 0.0;
        __syncthreads();
        const int r = 64 * cb2 + lane;
        if (lane < nb2)
            for (int j = 0; j < 16; ++j) {
                const int jj = wave + 4 * j;
                if (jj < nbc) part[j] += double(Lf[int64_t(c0 + jj) * m + r]) * xr[lane];
            }
        __syncthreads();
    }
    // rows below the diagonal inside the block itself are handled by the triangular solve
    for (int j = 0; j < 16; ++j) {
        double s = part[j];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        const int jj = wave + 4 * j;
        if (lane == 0 && jj < 64) sred[jj] = s;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        double val = lane < nbc ? v[c0 + lane] - sred[lane] : 0.0;
        for (int j = nbc - 1; j >= 0; --j) {
            const double xj = __shfl(val, j, 64) / double(Lf[int64_t(c0 + j) * m + c0 + j]);
            if (lane == j) xs[j] = xj;
            if (lane < j) val -= double(Lf[int64_t(c0 + lane) * m + c0 + j]) * xj;
        }
        if (lane < nbc) {
            const double xv = xs[lane];
            __hip_atomic_store(xb + 64 * cb + lane, xv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            v[c0 + lane] = xv;
            x[Fr.fbeg + c0 + lane] = xv;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) publish_flag(fl + cb);
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
        int nact = int(fl.size());
        for (int kb = 0; kb < maxp; kb += NBO) {
            while (nact > 0 && S.pcount[fl[nact - 1]] <= kb) --nact;
            if (nact == 0) break;
            int mm = 0;
            for (int i = 0; i < nact; ++i) mm = std::max(mm, S.mcount[fl[i]]);
            for (int nz = 0; nz < nact; nz += 65535) {
                const int cnt = std::min(65535, nact - nz);
                const int* fz = dfl + nz;
                for (int ks = kb; ks < std::min(kb + NBO, maxp); ks += NB) {
                    hipLaunchKernelGGL(k_potrf<T>, dim3(cnt), dim3(256), 0, stream_, fronts_.p, fz, ks, L_.p, info_.p);
                    const int rem = mm - ks - 1;
                    if (rem <= 0) continue;
                    hipLaunchKernelGGL(k_trsm<T>, dim3((rem + 63) / 64, cnt), dim3(64), 0, stream_, fronts_.p, fz, ks,
                                       L_.p);
                    // inside the outer block: columns [ks+nb, kb+NBO)
                    const int pc = (kb + NBO) - (ks + 1);
                    if (pc > 0)
                        hipLaunchKernelGGL(k_syrk<T>, dim3((rem + TILE - 1) / TILE, (pc + TILE - 1) / TILE, cnt),
                                           dim3(256), 0, stream_, fronts_.p, fz, kb, ks, NB, 0, L_.p, U_.p);
                }
                // trailing matrix: columns [kb+nbo, m) with K = nbo
                const int rem = mm - kb - 1;
                if (rem > 0) {
                    const int t = (rem + TILE - 1) / TILE;
                    hipLaunchKernelGGL(k_syrk<T>, dim3(t, t, cnt), dim3(256), 0, stream_, fronts_.p, fz, kb, kb, NBO, 1,
                                       L_.p, U_.p);
                }
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
void GpuCholesky<T>::prepare_big_levels() {
    const Symbolic& S = S_;
    big_.assign(S.nlevels, 0);
    std::vector<int2> fi, bi;
    std::vector<int> foff(S.nf, 0);
    big_fwd_ptr_.assign(S.nlevels + 1, 0);
    big_bwd_ptr_.assign(S.nlevels + 1, 0);
    int64_t nflags = 0;
    int maxflags = 0;
    for (int l = 0; l < S.nlevels; ++l) {
        const int nfl = S.level_ptr[l + 1] - S.level_ptr[l];
        int maxp = 0;
        for (int k = S.level_ptr[l]; k < S.level_ptr[l + 1]; ++k) maxp = std::max(maxp, S.pcount[S.level_fronts[k]]);
        big_[l] = (nfl < big_level_fronts_ && maxp > 64) ? 1 : 0;
        if (big_[l]) {
            int off = 0;
            for (int k = S.level_ptr[l]; k < S.level_ptr[l + 1]; ++k) {
                const int f = S.level_fronts[k];
                foff[f] = off;
                const int np = (S.pcount[f] + 63) / 64, nr = (S.mcount[f] + 63) / 64;
                for (int rb = 0; rb < nr; ++rb) fi.push_back(make_int2(f, rb));
                for (int cb = np - 1; cb >= 0; --cb) bi.push_back(make_int2(f, cb));
                off += np;
            }
            maxflags = std::max(maxflags, off);
        }
        big_fwd_ptr_[l + 1] = int(fi.size());
        big_bwd_ptr_[l + 1] = int(bi.size());
    }
    (void)nflags;
    big_fwd_items_ = to_device(fi.empty() ? std::vector<int2>{make_int2(0, 0)} : fi, stream_);
    big_bwd_items_ = to_device(bi.empty() ? std::vector<int2>{make_int2(0, 0)} : bi, stream_);
    foff_ = to_device(foff, stream_);
    max_flags_ = std::max(maxflags, 1);
    flags_.alloc(size_t(max_flags_) + 2);   // flags + ticket counter + error word
    xbuf_.alloc(size_t(max_flags_) * 64);
    std::vector<int> z(max_flags_ + 2, 0);
    flags_.upload(z.data(), z.size(), stream_);
    EMTD_HIP(hipStreamSynchronize(stream_));
    big_ready_ = true;
}

template <typename T>
int GpuCholesky<T>::sweep_error() {
    int h = 0;
    if (!big_ready_ || !flags_.p) return 0;
    EMTD_HIP(hipMemcpyAsync(&h, flags_.p + max_flags_ + 1, sizeof(int), hipMemcpyDeviceToHost, stream_));
    EMTD_HIP(hipStreamSynchronize(stream_));
    return h;
}

template <typename T>
void GpuCholesky<T>::solve(const double* d_b, double* d_x, int nrhs, int64_t ld) {
    const Symbolic& S = S_;
    if (!big_ready_) prepare_big_levels();
    if (nrhs > wrhs_) {
        W_.alloc(size_t(std::max<int64_t>(S.wsize, 1)) * nrhs);
        wrhs_ = nrhs;
    }
    int* counter = flags_.p + max_flags_;
    int* err = flags_.p + max_flags_ + 1;
    for (int l = 0; l < S.nlevels; ++l) {
        const int nfl = S.level_ptr[l + 1] - S.level_ptr[l];
        const int* lf = level_fronts_.p + S.level_ptr[l];
        if (!big_[l]) {
            hipLaunchKernelGGL(k_fwd<T>, dim3(nfl, nrhs), dim3(256), 0, stream_, fronts_.p, lf, children_.p,
                               relmap_.p, L_.p, W_.p, S.wsize, d_b, d_x, ld);
            continue;
        }
        const int b = big_fwd_ptr_[l], e = big_fwd_ptr_[l + 1];
        for (int r = 0; r < nrhs; ++r) {
            hipLaunchKernelGGL(k_fill<int>, dim3(grid_for(max_flags_ + 1, 256)), dim3(256), 0, stream_, flags_.p,
                               int64_t(max_flags_ + 1), 0);
            hipLaunchKernelGGL(k_fwd_gather, dim3(nfl), dim3(256), 0, stream_, fronts_.p, lf, children_.p, relmap_.p,
                               W_.p + r * S.wsize, d_b + r * ld);
            hipLaunchKernelGGL(k_fwd_big<T>, dim3(e - b), dim3(256), 0, stream_, fronts_.p, big_fwd_items_.p + b,
                               foff_.p, flags_.p, xbuf_.p, counter, err, L_.p, W_.p + r * S.wsize, d_x + r * ld);
        }
    }
    for (int l = S.nlevels - 1; l >= 0; --l) {
        const int nfl = S.level_ptr[l + 1] - S.level_ptr[l];
        const int* lf = level_fronts_.p + S.level_ptr[l];
        if (!big_[l]) {
            hipLaunchKernelGGL(k_bwd<T>, dim3(nfl, nrhs), dim3(256), 0, stream_, fronts_.p, lf, rows_.p, L_.p, W_.p,
                               S.wsize, d_x, ld);
            continue;
        }
        const int b = big_bwd_ptr_[l], e = big_bwd_ptr_[l + 1];
        for (int r = 0; r < nrhs; ++r) {
            hipLaunchKernelGGL(k_fill<int>, dim3(grid_for(max_flags_ + 1, 256)), dim3(256), 0, stream_, flags_.p,
                               int64_t(max_flags_ + 1), 0);
            hipLaunchKernelGGL(k_bwd_gather, dim3(nfl), dim3(256), 0, stream_, fronts_.p, lf, rows_.p,
                               W_.p + r * S.wsize, d_x + r * ld);
            hipLaunchKernelGGL(k_bwd_big<T>, dim3(e - b), dim3(256), 0, stream_, fronts_.p, big_bwd_items_.p + b,
                               foff_.p, flags_.p, xbuf_.p, counter, err, L_.p, W_.p + r * S.wsize, d_x + r * ld);
        }
    }
    EMTD_HIP(hipGetLastError());
}

template class GpuCholesky<double>;
template class GpuCholesky<float>;

}  // namespace emtd
