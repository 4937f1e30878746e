// The anomalous-edge export before a primary/secondary run (host code, no GPU).
//
// emtd_unload_anomal_nonstat replaces UnloadAnomalNonstat (M/UnloadAnomalNonstat/TaskSP.cpp:104-127
// and Time_approx_for_vfem::Read_data, Unload_anomal_nodes, Unload_anomal_nodesA0 at
// M/UnloadAnomalNonstat/time_approx.cpp:118-363). An edge is anomalous when it belongs to an
// element whose conductivity differs from the normal model's (|sigma0 - sigma| > 1e-6) and
// neither of its nodes lies above the top of the 2D layered model plus 1e-3 (the last depth of
// z_sig_2d). The module writes, per element, whether it is anomalous or touches an anomalous
// edge (isElemAnomal), per edge the flag (mtrsizeforE0, which the calculation module reads),
// and for the anomalous edges their midpoints (xyzVectorE0, preceded by their count) and unit
// tangents (TgCompE0), where the 2D modules evaluate the normal field. When some material's
// permeability differs from the normal model's, the same is written for the elements with
// |mu0 - mu| > 1e-6 (absolute, H/m) without the depth limit: isElemAnomal_A, mtrsizeforA0,
// xyzVectorA0, TgCompA0, the last two with the stream's default six significant digits.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/emtd.h"
#include "common.hpp"
#include "model_io.hpp"

using namespace emtd;

namespace {

constexpr double kMu0 = 1.25663706143591729e-6;  // M/UnloadAnomalNonstat/vfem_const.h:21

void write_text(const std::string& path, const std::string& s) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw Error(EMTD_ERR_NOFILE, "cannot write " + path);
    f << s;
}

// z_sig_2d: the layer count, then one line per layer whose first number is its depth
double top_of_layers(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw Error(EMTD_ERR_NOFILE, "cannot open " + path);
    int nzl = 0;
    in >> nzl;
    if (!in || nzl <= 0) throw Error(EMTD_ERR_BADFILE, "bad z_sig_2d");
    double z = 0;
    for (int i = 0; i < nzl; ++i) {
        in >> z;
        if (!in) throw Error(EMTD_ERR_BADFILE, "truncated z_sig_2d");
        in.ignore(1000, '\n');
    }
    return z;
}

// The four files of one export: element flags, edge flags, midpoints and tangents of the
// flagged edges. fmt is the number format of the last two ("%.14e" or "%g").
void write_export(const std::string& dir, const ModelFiles& M, const std::vector<char>& elem,
                  const std::vector<char>& edge, const char* elem_name, const char* edge_name,
                  const char* mid_name, const char* tg_name, const char* fmt) {
    auto P = [&](const char* s) { return join_path(dir, s); };
    std::string s;
    for (char f : elem) s += f ? "1\n" : "0\n";
    write_text(P(elem_name), s);
    s.clear();
    for (char f : edge) s += f ? "1\n" : "0\n";
    write_text(P(edge_name), s);
    int count = 0;
    for (char f : edge) count += f != 0;
    char b[128];
    auto num3 = [&](double x, double y, double z) {
        char f3[32];
        std::snprintf(f3, sizeof f3, "%s %s %s\n", fmt, fmt, fmt);
        std::snprintf(b, sizeof b, f3, x, y, z);
        return std::string(b);
    };
    std::string mid = std::to_string(count) + "\n", tg;
    for (int i = 0; i < M.n; ++i) {
        if (!edge[i]) continue;
        const double* a = &M.xyz[3 * size_t(M.edges[2 * i])];
        const double* c = &M.xyz[3 * size_t(M.edges[2 * i + 1])];
        mid += num3(0.5 * (a[0] + c[0]), 0.5 * (a[1] + c[1]), 0.5 * (a[2] + c[2]));
        const double px = c[0] - a[0], py = c[1] - a[1], pz = c[2] - a[2];
        const double len = std::sqrt(px * px + py * py + pz * pz);
        tg += num3(px / len, py / len, pz / len);
    }
    write_text(P(mid_name), mid);
    write_text(P(tg_name), tg);
}

}  // namespace

extern "C" int emtd_unload_anomal_nonstat(const char* workdir) {
    try {
        if (!workdir) throw Error(EMTD_ERR_RANGE, "null argument");
        const std::string d = workdir;
        const double z0 = top_of_layers(join_path(d, "z_sig_2d")) + 1e-3;
        const ModelFiles M = read_mesh_files(d);
        for (int e = 0; e < M.kpar; ++e)
            if (M.nvkat[e] < 0 || M.nvkat[e] >= M.nmat) throw Error(EMTD_ERR_BADFILE, "material number out of range");
        // conductivity anomaly (Unload_anomal_nodes)
        std::vector<char> elem(M.kpar, 0), edge(M.n, 0);
        for (int e = 0; e < M.kpar; ++e) {
            const int k = M.nvkat[e];
            if (std::fabs(M.sigma0[k] - M.sigma[k]) > 1e-6) {
                elem[e] = 1;
                for (int j = 0; j < 12; ++j) edge[M.ed[12 * size_t(e) + j]] = 1;
            }
        }
        for (int i = 0; i < M.n; ++i)
            if (edge[i] && (M.xyz[3 * size_t(M.edges[2 * i]) + 2] > z0 || M.xyz[3 * size_t(M.edges[2 * i + 1]) + 2] > z0))
                edge[i] = 0;
        for (int e = 0; e < M.kpar; ++e)
            for (int j = 0; j < 12 && !elem[e]; ++j)
                if (edge[M.ed[12 * size_t(e) + j]]) elem[e] = 1;
        write_export(d, M, elem, edge, "isElemAnomal", "mtrsizeforE0", "xyzVectorE0", "TgCompE0", "%.14e");
        // permeability anomaly (Read_data, Unload_anomal_nodesA0)
        int n_anom_mu = 0;
        for (int k = 0; k < M.nmat; ++k) n_anom_mu += std::fabs(M.mu[k] - M.mu0[k]) / kMu0 > 1e-6;
        if (n_anom_mu) {
            std::vector<char> elem_a(M.kpar, 0), edge_a(M.n, 0);
            for (int e = 0; e < M.kpar; ++e) {
                const int k = M.nvkat[e];
                if (std::fabs(M.mu0[k] - M.mu[k]) > 1e-6) {
                    elem_a[e] = 1;
                    for (int j = 0; j < 12; ++j) edge_a[M.ed[12 * size_t(e) + j]] = 1;
                }
            }
            write_export(d, M, elem_a, edge_a, "isElemAnomal_A", "mtrsizeforA0", "xyzVectorA0", "TgCompA0", "%g");
        }
        return EMTD_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return EMTD_ERR_ERROR;
    }
}
