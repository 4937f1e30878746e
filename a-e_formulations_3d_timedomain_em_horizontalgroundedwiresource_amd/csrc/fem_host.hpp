// Host-side pieces of the calculation module that run once per model: portrait, boundary,
// line-source placement, receiver functionals, decades and time differentiation.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "model_io.hpp"

namespace emtd {

// Strictly lower, sorted, unique pattern of the edge matrix (A2/T_Portrait.cpp:45-122).
void build_portrait(int kpar, const int* ed, int n, std::vector<int>& ig, std::vector<int>& jg);

// Edges whose two nodes are both boundary nodes (A2/T_Bound_Cond.cpp:58-89).
std::vector<char> boundary_edge_mask(const ModelFiles& M);

// Line-source RHS shapes, one dense vector of n per source (original numbering), or empty
// when the current function is zero after t0 (A2/vec_prep_data.cpp:982-1270).
struct SourceInfo {
    std::vector<double> shape;  // npls*n
    std::vector<int> elem;      // per dipole
    std::vector<double> loc;    // per dipole: 3 local coordinates
    std::vector<double> pos;    // per dipole: 3 global coordinates
};
SourceInfo place_line_sources(const ModelFiles& M);

// Receiver functionals: row (receiver, component) -> sparse weights over edges (original
// numbering). kind 0 = curl of the field at receivers (one subdomain), kind 1 = the field
// itself (one subdomain per material) (A2/OutputResultant3d.cpp:343-515, A2/Subdomain.cpp).
struct ReceiverOps {
    int nrec = 0;
    std::vector<int64_t> rp;  // 3*nrec + 1
    std::vector<int> idx;
    std::vector<double> w;
};
ReceiverOps build_receiver_ops(const ModelFiles& M, const std::vector<double>& pts, int kind);

// Decades (A2/TaskSP.cpp:326-398, 628-660).
struct Decade {
    int tbeg, tend, npre1, npre2, nlst;
};
std::vector<Decade> decade_plan(const std::vector<double>& t);
void print_window(const ModelFiles& M, int& first, int& last);

// Time differentiation of receiver series (A2/give_out_vec_loop.cpp:240-470; 3-point:
// A3/give_out_vec_loop.cpp:313-367). F and out: (last-first+1) x stride values.
void diff_series(const std::vector<double>& t, const double* F, double* out, int first, int last, int stride,
                 const std::vector<int>& vpre1, const std::vector<int>& vpre2, int scheme);

// Output tables eall_anom.N / edsall_anom.N / ball_anom.N (A2/give_out_vec_loop.cpp:532-609).
void write_tables(const std::string& dir, const ModelFiles& M, int first, int last, const double* eds,
                  const double* bfield, const double* efield);

}  // namespace emtd
