// Host-side ordering and multifrontal symbolic analysis of the sparse SPD system.
#pragma once

#include <cstdint>
#include <vector>

namespace emtd {

// Integer position of a DOF on the doubled tensor grid (edge midpoint * 2 in node-index units).
struct GridCoord {
    int x, y, z;
};

// Result of the symbolic phase. Permuted ("new") indices: 0..nn-1 are the factored DOFs in
// elimination order, nn..n-1 the singleton DOFs (no off-diagonal coupling, solved as b/d).
struct Symbolic {
    int n = 0;
    int nn = 0;
    std::vector<int> perm;   // new -> original, size n
    std::vector<int> iperm;  // original -> new, size n
    int nf = 0;              // fronts (supernodes), numbered so that children precede parents
    std::vector<int> fbeg;   // pivot range [fbeg[f], fbeg[f+1]) in new numbering, size nf+1
    std::vector<int> parent; // -1 for roots
    std::vector<int> level;  // 0 for leaves, 1 + max(children) otherwise
    std::vector<int> pcount, mcount;          // pivots and rows per front
    std::vector<int64_t> rows_ptr;            // rows of front f: pivots then update rows
    std::vector<int> rows;
    std::vector<int> child_ptr, children;     // children of each front, ascending
    std::vector<int64_t> relmap_ptr;          // per front: u_f entries
    std::vector<int> relmap;                  // update row k of f -> position in parent's rows
    std::vector<int64_t> loff, uoff, woff;    // offsets of L_f (m*p), U_f (u*u), workspace (m)
    int64_t lsize = 0, usize = 0, wsize = 0;
    std::vector<int> level_ptr, level_fronts; // fronts grouped by level
    int nlevels = 0;
    int max_children = 0;
    int max_p = 0, max_m = 0;
    double flops = 0, nnz_l = 0;
};

// Ordering options.
struct OrderingOptions {
    int leaf_size = 64;   // stop dissecting below this many DOFs
};

// Builds the ordering and symbolic structure of the matrix with lower RSF pattern (ig, jg).
// active[i] == 0 marks singleton DOFs (their rows/columns are ignored). When coords is
// non-empty, geometric nested dissection on the grid is used; otherwise nested dissection on
// breadth-first level structures of the matrix graph.
Symbolic analyze(int n, const int* ig, const int* jg, const std::vector<char>& active,
                 const std::vector<GridCoord>& coords, const OrderingOptions& opt);

// Assigns U_f offsets by lifetime (level of f .. level of its parent) so that update matrices
// that are never live at the same time share storage; sets S.uoff and S.usize.
void place_update_matrices(Symbolic& S);

// For each lower RSF entry (and the diagonal) of the original pattern, the position of the
// entry in the front storage L (or -1 when it belongs to a singleton row/column).
void assembly_map(const Symbolic& S, int n, const int* ig, const int* jg,
                  std::vector<int64_t>& dest_off, std::vector<int64_t>& dest_diag);

// Grid coordinates of the edge DOFs of a tensor-product brick mesh (nodes numbered x fastest,
// then y, then z). Returns an empty vector when the nodes do not form a tensor grid.
std::vector<GridCoord> edge_grid_coords(int kuzlov, const double* xyz, int n_edges, const int* edges);

}  // namespace emtd
