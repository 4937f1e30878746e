// extern "C" entry points of libemtd (declared in include/emtd.h).
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/emtd.h"
#include "assemble.hpp"
#include "chol.hpp"
#include "common.hpp"
#include "fem_host.hpp"
#include "symbolic.hpp"
#include "vec.hpp"

namespace emtd {
static thread_local std::string g_last_error;
void set_last_error(const std::string& s) { g_last_error = s; }
}  // namespace emtd

using namespace emtd;

#define EMTD_GUARD_BEGIN try {
#define EMTD_GUARD_END                           \
    }                                            \
    catch (const Error& e) {                     \
        set_last_error(e.what());                \
        return e.code;                           \
    }                                            \
    catch (const std::bad_alloc&) {              \
        set_last_error("host allocation failed"); \
        return EMTD_ERR_NOMEM;                   \
    }                                            \
    catch (const std::exception& e) {            \
        set_last_error(e.what());                \
        return EMTD_ERR_ERROR;                   \
    }

extern "C" const char* emtd_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------ portrait
extern "C" int emtd_portrait(const emtd_mesh* mesh, int* ig, int64_t* nnz, int* jg) {
    EMTD_GUARD_BEGIN
    if (!mesh || !ig || !nnz) throw Error(EMTD_ERR_RANGE, "null argument");
    std::vector<int> vig, vjg;
    build_portrait(mesh->kpar, mesh->ed, mesh->n_edges, vig, vjg);
    std::copy(vig.begin(), vig.end(), ig);
    *nnz = int64_t(vjg.size());
    if (jg) std::copy(vjg.begin(), vjg.end(), jg);
    return EMTD_OK;
    EMTD_GUARD_END
}

// ------------------------------------------------------------------ assembly
extern "C" int emtd_assemble(const emtd_mesh* mesh, const int* ig, const int* jg, int device, double* b_ggl,
                             double* b_di, double* c_ggl, double* c_di) {
    EMTD_GUARD_BEGIN
    if (!mesh || !ig || !jg || !b_ggl || !b_di || !c_ggl || !c_di) throw Error(EMTD_ERR_RANGE, "null argument");
    select_device(device);
    hipStream_t s;
    EMTD_HIP(hipStreamCreate(&s));
    {
        const int n = mesh->n_edges;
        Assembler A(*mesh, n, ig, jg, s);
        const int64_t nnz = ig[n];
        DevBuf<double> bg(std::max<int64_t>(nnz, 1)), bd(n), cg(std::max<int64_t>(nnz, 1)), cd(n);
        A.run(bg.p, bd.p, cg.p, cd.p);
        bg.download(b_ggl, nnz, s);
        bd.download(b_di, n, s);
        cg.download(c_ggl, nnz, s);
        cd.download(c_di, n, s);
        EMTD_HIP(hipStreamSynchronize(s));
    }
    (void)hipStreamDestroy(s);
    return EMTD_OK;
    EMTD_GUARD_END
}

// ------------------------------------------------------------------ pardiso seam
struct emtd_solver {
    int device = -1;
    int precision = EMTD_FP64;
    hipStream_t stream = nullptr;
    int n = 0;
    std::vector<int> ig, jg;
    std::vector<char> active;
    std::unique_ptr<Symbolic> S;
    std::unique_ptr<GpuCholesky<double>> c64;
    std::unique_ptr<GpuCholesky<float>> c32;
    DevBuf<int64_t> dest_off, dest_diag;
    DevBuf<double> ggl, di;
    DevBuf<int> perm, singles;
    DevBuf<double> single_inv;
    DevBuf<double> b, x, bp, xp;
    int nb_rhs = 0;
    bool factored = false;
    emtd_solver_stats st{};
};

extern "C" int emtd_solver_create(int device, int precision, emtd_solver** out) {
    EMTD_GUARD_BEGIN
    if (!out) throw Error(EMTD_ERR_RANGE, "null argument");
    if (precision != EMTD_FP64 && precision != EMTD_FP32_IR) throw Error(EMTD_ERR_RANGE, "unknown precision");
    select_device(device);
    auto s = std::make_unique<emtd_solver>();
    s->device = device;
    s->precision = precision;
    EMTD_HIP(hipStreamCreate(&s->stream));
    *out = s.release();
    return EMTD_OK;
    EMTD_GUARD_END
}

extern "C" int emtd_solver_factorize(emtd_solver* s, int n, const int* ig, const int* jg, const double* ggl,
                                     const double* di) {
    EMTD_GUARD_BEGIN
    if (!s || !ig || !jg || !ggl || !di || n <= 0) throw Error(EMTD_ERR_RANGE, "bad argument");
    select_device(s->device);
    const int64_t nnz = ig[n];
    // singleton DOFs: no nonzero off-diagonal entry in row or column
    std::vector<char> active(n, 0);
    for (int i = 0; i < n; ++i)
        for (int k = ig[i]; k < ig[i + 1]; ++k)
            if (ggl[k] != 0.0) { active[i] = 1; active[jg[k]] = 1; }
    const bool same = s->S && s->n == n && (int64_t)s->jg.size() == nnz && s->active == active &&
                      std::equal(ig, ig + n + 1, s->ig.begin()) && std::equal(jg, jg + nnz, s->jg.begin());
    if (!same) {
        s->c64.reset();
        s->c32.reset();
        s->n = n;
        s->ig.assign(ig, ig + n + 1);
        s->jg.assign(jg, jg + nnz);
        s->active = active;
        OrderingOptions opt;
        s->S = std::make_unique<Symbolic>(analyze(n, ig, jg, active, {}, opt));
        std::vector<int64_t> doff, ddiag;
        assembly_map(*s->S, n, ig, jg, doff, ddiag);
        s->dest_off = to_device(doff, s->stream);
        s->dest_diag = to_device(ddiag, s->stream);
        s->perm = to_device(std::vector<int>(s->S->perm.begin(), s->S->perm.begin() + s->S->nn), s->stream);
        std::vector<int> sing(s->S->perm.begin() + s->S->nn, s->S->perm.end());
        s->singles = to_device(sing, s->stream);
        s->ggl.alloc(std::max<int64_t>(nnz, 1));
        s->di.alloc(n);
        if (s->precision == EMTD_FP64)
            s->c64 = std::make_unique<GpuCholesky<double>>(*s->S, s->stream);
        else
            s->c32 = std::make_unique<GpuCholesky<float>>(*s->S, s->stream);
        s->nb_rhs = 0;
    }
    // singleton inverses
    {
        std::vector<double> inv;
        for (int k = s->S->nn; k < n; ++k) inv.push_back(1.0 / di[s->S->perm[k]]);
        s->single_inv = to_device(inv, s->stream);
    }
    s->ggl.upload(ggl, nnz, s->stream);
    s->di.upload(di, n, s->stream);
    int bad;
    if (s->c64) {
        s->c64->factorize_values(s->dest_off.p, s->dest_diag.p, nnz, n, s->ggl.p, s->di.p);
        bad = s->c64->check_info();
        s->st.factor_seconds = s->c64->last_factor_seconds();
    } else {
        s->c32->factorize_values(s->dest_off.p, s->dest_diag.p, nnz, n, s->ggl.p, s->di.p);
        bad = s->c32->check_info();
        s->st.factor_seconds = s->c32->last_factor_seconds();
    }
    s->st.n = n;
    s->st.n_factored = s->S->nn;
    s->st.n_fronts = s->S->nf;
    s->st.n_levels = s->S->nlevels;
    s->st.nnz_l = s->S->nnz_l;
    s->st.factor_flops = s->S->flops;
    s->factored = bad < 0;
    if (bad >= 0)
        throw Error(EMTD_ERR_NOTSPD, "non-positive pivot at permuted index " + std::to_string(bad));
    return EMTD_OK;
    EMTD_GUARD_END
}

extern "C" int emtd_solver_solve_nrhs(emtd_solver* s, int nrhs, const double* pr, double* x) {
    EMTD_GUARD_BEGIN
    if (!s || !pr || !x || nrhs <= 0) throw Error(EMTD_ERR_RANGE, "bad argument");
    if (!s->factored) throw Error(EMTD_ERR_NOTINIT, "solve before a successful factorize");
    select_device(s->device);
    const int n = s->n, nn = s->S->nn;
    if (nrhs > s->nb_rhs) {
        s->b.alloc(size_t(n) * nrhs);
        s->x.alloc(size_t(n) * nrhs);
        s->bp.alloc(size_t(std::max(nn, 1)) * nrhs);
        s->xp.alloc(size_t(std::max(nn, 1)) * nrhs);
        s->nb_rhs = nrhs;
    }
    hipEvent_t e0, e1;
    EMTD_HIP(hipEventCreate(&e0));
    EMTD_HIP(hipEventCreate(&e1));
    s->b.upload(pr, size_t(n) * nrhs, s->stream);
    EMTD_HIP(hipEventRecord(e0, s->stream));
    gather(s->stream, s->bp.p, s->b.p, s->perm.p, nn, nrhs, nn, n);
    if (s->c64)
        s->c64->solve(s->bp.p, s->xp.p, nrhs, nn);
    else
        s->c32->solve(s->bp.p, s->xp.p, nrhs, nn);
    scatter(s->stream, s->x.p, s->xp.p, s->perm.p, nn, nrhs, n, nn);
    singletons(s->stream, s->x.p, s->b.p, s->singles.p, s->single_inv.p, n - nn, nrhs, n);
    EMTD_HIP(hipEventRecord(e1, s->stream));
    s->x.download(x, size_t(n) * nrhs, s->stream);
    EMTD_HIP(hipStreamSynchronize(s->stream));
    const std::string serr = s->c64 ? s->c64->sweep_error() : s->c32->sweep_error();
    if (!serr.empty()) throw Error(EMTD_ERR_ERROR, serr);
    float ms = 0;
    EMTD_HIP(hipEventElapsedTime(&ms, e0, e1));
    s->st.solve_seconds = ms * 1e-3;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return EMTD_OK;
    EMTD_GUARD_END
}

extern "C" int emtd_solver_stop(emtd_solver* s) {
    EMTD_GUARD_BEGIN
    if (!s) throw Error(EMTD_ERR_RANGE, "null argument");
    s->c64.reset();
    s->c32.reset();
    s->S.reset();
    s->factored = false;
    return EMTD_OK;
    EMTD_GUARD_END
}

extern "C" void emtd_solver_destroy(emtd_solver* s) {
    if (!s) return;
    try {
        select_device(s->device);
        s->c64.reset();
        s->c32.reset();
        if (s->stream) (void)hipStreamDestroy(s->stream);
    } catch (...) {
    }
    delete s;
}

extern "C" int emtd_solver_get_stats(const emtd_solver* s, emtd_solver_stats* st) {
    if (!s || !st) return EMTD_ERR_RANGE;
    *st = s->st;
    return EMTD_OK;
}
