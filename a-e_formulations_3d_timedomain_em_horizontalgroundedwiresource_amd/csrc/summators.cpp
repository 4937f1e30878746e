// The output chain after the calculation module: per-component receiver files and the E-field
// integration along MN lines (host code, no GPU).
//
// emtd_sum_nonstat_3d_only replaces SumNonstat3DOnly (M/SumNonstat3DOnly/main.cpp:171-384):
// the module's anomalous tables ball_anom.N, edsall_anom.N, eall_anom.N become one file per
// component (Bxall.N ... Ezall.N) whose rows are (t ms, normal = 0, anomalous, total), with
// the times taken from infite.0 inside timeintervalforprint (picket format of
// M/SumNonstat3DOnly/picket.cpp ReadPickets/WritePickets).
//
// emtd_summator_edmn replaces SummatorEDMN (M/SummatorEDMN/main.cpp:40-295): for MN line i of
// xyzmn<place+1>_<profile> with points p_k of xyzVectorElin<place+1>_<profile>, the voltage
// dV = sum_k E(p_k).d_k with d_0 = 2 (p_0 - A) and d_k = p_k - p_{k-1} (the midpoint rule when
// the points are segment midpoints), from exall/eyall/ezall<profile>_<place>; with ddV the
// points of the second half of the line enter with a minus sign (edmoonall...).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/emtd.h"
#include "common.hpp"
#include "model_io.hpp"

using namespace emtd;

namespace {

struct Picket {
    int ix = 0, iy = 0;
    double p[3] = {0, 0, 0};
    std::vector<double> t;
    std::vector<double> v[3];
};

const char* kLegB[3] = {"t (ms)        Bx (mTl)     Bx (mTl)     Bx (mTl)", "t (ms)        By (mTl)     By (mTl)     By (mTl)",
                        "t (ms)        Bz (mTl)     Bz (mTl)     Bz (mTl)"};
const char* kLegF[3] = {"t (ms)      Emfx (mV)   Emfx (mV)   Emfx (mV)", "t (ms)      Emfy (mV)   Emfy (mV)   Emfy (mV)",
                        "t (ms)      Emfz (mV)   Emfz (mV)   Emfz (mV)"};
const char* kLegE[3] = {"t (ms)        Ex (V/m)     Ex (V/m)     Ex (V/m)", "t (ms)        Ey (V/m)     Ey (V/m)     Ey (V/m)",
                        "t (ms)        Ez (V/m)     Ez (V/m)     Ez (V/m)"};
const char* kNameB[3] = {"Bxall", "Byall", "Bzall"};
const char* kNameF[3] = {"Edsxall", "Edsyall", "Edsall"};
const char* kNameE[3] = {"Exall", "Eyall", "Ezall"};
const char* kLegDV =
    "         t, ms                         dVp, mV                         dVs, mV                         dV, mV";
const char* kLegDDV =
    "         t, ms                         ddVp, mV                         ddVs, mV                         ddV, mV";

// ReadPickets (M/SumNonstat3DOnly/picket.cpp): nr pickets of nt rows each.
std::vector<Picket> read_pickets(const std::string& path, int nr, int nt) {
    std::ifstream in(path);
    if (!in) throw Error(EMTD_ERR_NOFILE, "cannot open " + path);
    std::vector<Picket> out(nr);
    for (Picket& k : out) {
        int tmp;
        in >> k.ix >> k.iy >> k.p[0] >> k.p[1] >> k.p[2] >> tmp >> tmp >> tmp >> tmp >> tmp >> tmp;
        std::string leg;
        std::getline(in, leg);  // rest of the line
        std::getline(in, leg);  // legend
        k.t.resize(nt);
        for (auto& c : k.v) c.resize(nt);
        for (int i = 0; i < nt; ++i) in >> k.t[i] >> k.v[0][i] >> k.v[1][i] >> k.v[2][i];
        if (!in) throw Error(EMTD_ERR_BADFILE, "truncated table " + path);
    }
    return out;
}

void write_component(const std::string& path, const std::vector<Picket>& pk, const char* leg,
                     const std::vector<double>& t, int comp) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw Error(EMTD_ERR_NOFILE, "cannot write " + path);
    for (const Picket& k : pk) {
        std::fprintf(f, "%d %d\n%.16e %.16e %.16e\n1\n1 1 1 1\n1\n%s\n", k.ix, k.iy, k.p[0], k.p[1], k.p[2], leg);
        for (size_t i = 0; i < t.size(); ++i) {
            const double v = k.v[comp][i];
            std::fprintf(f, "%.16e %.16e %.16e %.16e\n", t[i], 0.0, v, v);
        }
    }
    std::fclose(f);
}

// Picket::ReadPicket repeated (M/SummatorEDMN/Picket.h): header lines, then rows of four
// numbers up to the first line that is not one.
struct Series {
    std::vector<double> t;
    std::vector<double> v;  // rows x 3 (normal, anomalous, total)
};

std::vector<Series> read_series(const std::string& path, int count) {
    std::ifstream in(path);
    if (!in) throw Error(EMTD_ERR_NOFILE, "cannot open " + path);
    std::vector<std::string> lines;
    for (std::string s; std::getline(in, s);) lines.push_back(s);
    std::vector<Series> out(count);
    size_t k = 0;
    for (Series& s : out) {
        k += 6;
        while (k < lines.size()) {
            std::istringstream is(lines[k]);
            double r[4];
            if (!(is >> r[0] >> r[1] >> r[2] >> r[3])) break;
            s.t.push_back(r[0]);
            s.v.insert(s.v.end(), r + 1, r + 4);
            ++k;
        }
    }
    return out;
}

std::vector<double> read_numbers(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw Error(EMTD_ERR_NOFILE, "cannot open " + path);
    std::vector<double> v;
    for (double x; in >> x;) v.push_back(x);
    return v;
}

}  // namespace

extern "C" int emtd_sum_nonstat_3d_only(const char* workdir) {
    try {
        if (!workdir) throw Error(EMTD_ERR_RANGE, "null argument");
        const std::string d = workdir;
        auto P = [&](const std::string& s) { return join_path(d, s); };
        std::vector<double> g = read_numbers(P("group"));
        if (g.empty()) throw Error(EMTD_ERR_BADFILE, "bad group file");
        int npls = 0;
        for (int k = 0; k < int(g[0]); ++k) npls += int(g[3 + 3 * k]) - int(g[2 + 3 * k]) + 1;
        std::vector<double> win = read_numbers(P("timeintervalforprint"));
        if (win.size() < 2) throw Error(EMTD_ERR_BADFILE, "bad timeintervalforprint");
        const std::vector<double> times = read_time_grid(P("infite.0"));
        std::vector<double> tw;
        for (double tj : times)
            if (tj >= win[0] && tj <= win[1]) tw.push_back(tj * 1000.0);
        const int nt = int(tw.size());
        for (int kind = 0; kind < 2; ++kind) {
            std::vector<double> counts = read_numbers(P(kind == 0 ? "recvsb" : "recvse"));
            if (int(counts.size()) < npls) throw Error(EMTD_ERR_BADFILE, "bad receiver count file");
            for (int ip = 0; ip < npls; ++ip) {
                const int nall = int(counts[ip]);
                const std::string sfx = "." + std::to_string(ip + 1);
                if (kind == 0) {
                    auto pk = read_pickets(P("ball_anom" + sfx), nall, nt);
                    for (int c = 0; c < 3; ++c) write_component(P(kNameB[c] + sfx), pk, kLegB[c], tw, c);
                    pk = read_pickets(P("edsall_anom" + sfx), nall, nt);
                    for (int c = 0; c < 3; ++c) write_component(P(kNameF[c] + sfx), pk, kLegF[c], tw, c);
                } else {
                    auto pk = read_pickets(P("eall_anom" + sfx), nall, nt);
                    for (int c = 0; c < 3; ++c) write_component(P(kNameE[c] + sfx), pk, kLegE[c], tw, c);
                }
            }
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

extern "C" int emtd_summator_edmn(const char* workdir, int with_ddv, int profile, int place) {
    try {
        if (!workdir) throw Error(EMTD_ERR_RANGE, "null argument");
        const std::string d = workdir;
        auto P = [&](const std::string& s) { return join_path(d, s); };
        std::vector<double> nl = read_numbers(P("xyzmn"));
        if (nl.empty() || int(nl[0]) == 0) return EMTD_OK;
        const std::string pp = std::to_string(place + 1) + "_" + std::to_string(profile);
        const std::string tag = std::to_string(profile) + "_" + std::to_string(place);
        std::vector<double> ab = read_numbers(P("xyzmn" + pp));
        std::vector<double> pe = read_numbers(P("xyzVectorElin" + pp));
        if (ab.empty() || pe.empty()) throw Error(EMTD_ERR_BADFILE, "bad MN line files");
        const int q = int(ab[0]), ne = int(pe[0]);
        if (q <= 0 || int(ab.size()) < 1 + 6 * q || int(pe.size()) < 1 + 3 * ne)
            throw Error(EMTD_ERR_BADFILE, "bad MN line files");
        const int per = ne / q;
        std::vector<Series> E[3] = {read_series(P("exall" + tag), q * per), read_series(P("eyall" + tag), q * per),
                                    read_series(P("ezall" + tag), q * per)};
        FILE* fdv = std::fopen(P("edmnall" + tag).c_str(), "w");
        if (!fdv) throw Error(EMTD_ERR_NOFILE, "cannot write edmnall" + tag);
        FILE* fddv = with_ddv ? std::fopen(P("edmoonall" + tag).c_str(), "w") : nullptr;
        if (with_ddv && !fddv) {
            std::fclose(fdv);
            throw Error(EMTD_ERR_NOFILE, "cannot write edmoonall" + tag);
        }
        int m = 0;
        for (int i = 0; i < q; ++i) {
            const double* A = &ab[1 + 6 * i];
            const double* B = A + 3;
            std::vector<double> acc, acc2, times;
            double prev[3] = {0, 0, 0};
            for (int k = 0; k < per; ++k, ++m) {
                const double* p = &pe[1 + 3 * m];
                double v[3];
                for (int a = 0; a < 3; ++a) v[a] = k == 0 ? (p[a] - A[a]) * 2.0 : p[a] - prev[a];
                for (int a = 0; a < 3; ++a) prev[a] = p[a];
                const size_t nt = E[0][m].t.size();
                if (k == 0) {
                    times = E[0][m].t;
                    acc.assign(3 * nt, 0.0);
                    acc2.assign(3 * nt, 0.0);
                }
                const double sign = k > 0 && k >= per / 2 ? -1.0 : 1.0;
                for (size_t l = 0; l < nt && l < times.size(); ++l)
                    for (int c = 0; c < 3; ++c) {
                        const double e = E[0][m].v[3 * l + c] * v[0] + E[1][m].v[3 * l + c] * v[1] +
                                         E[2][m].v[3 * l + c] * v[2];
                        acc[3 * l + c] += e;
                        acc2[3 * l + c] += sign * e;
                    }
            }
            const double mid[3] = {0.5 * (A[0] + B[0]), 0.5 * (A[1] + B[1]), 0.5 * (A[2] + B[2])};
            for (int w = 0; w < (with_ddv ? 2 : 1); ++w) {
                FILE* f = w ? fddv : fdv;
                const std::vector<double>& a = w ? acc2 : acc;
                std::fprintf(f, "%d %d\n%.16e %.16e %.16e\n1\n1 1 1 1\n1\n%s\n", int(mid[0]), int(mid[1]), mid[0], mid[1],
                             mid[2], w ? kLegDDV : kLegDV);
                for (size_t l = 0; l < times.size(); ++l)
                    std::fprintf(f, "%.16e \t%.16e \t%.16e \t%.16e\n", times[l], a[3 * l], a[3 * l + 1], a[3 * l + 2]);
            }
        }
        std::fclose(fdv);
        if (fddv) std::fclose(fddv);
        return EMTD_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return EMTD_ERR_ERROR;
    }
}
