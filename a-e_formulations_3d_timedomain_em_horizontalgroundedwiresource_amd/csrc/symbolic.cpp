// Ordering (nested dissection) and multifrontal symbolic analysis.
//
// The reference hands its lower RSF matrix to MKL PARDISO with default parameters
// (A2/pardiso.cpp:69-111), i.e. a nested-dissection ordering followed by a supernodal
// symbolic factorization, redone every decade. Here the analysis runs once per pattern and
// is reused by every refactorization:
//
//  * ordering: geometric nested dissection on the doubled tensor grid when the mesh is a
//    tensor grid (separators are node planes: the edges lying in the plane x = x_c split the
//    edges with smaller and larger x, because no brick spans both sides), otherwise nested
//    dissection on breadth-first level structures of the matrix graph;
//  * fronts: every dissection leaf and every separator is one dense front (supernode);
//  * structure: update rows U_f = (lower rows of the pivot columns  U  children's U) minus the
//    pivots; the parent is the front holding min(U_f) (the elimination-tree parent).
#include "symbolic.hpp"

#include <algorithm>
#include <atomic>
#include <thread>
#include <map>
#include <climits>
#include <cmath>
#include <numeric>

#include "common.hpp"

namespace emtd {
namespace {

struct Graph {
    std::vector<int64_t> ptr;
    std::vector<int> adj;
};

Graph build_graph(int n, const int* ig, const int* jg, const std::vector<char>& active) {
    Graph g;
    g.ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i)
        for (int k = ig[i]; k < ig[i + 1]; ++k) {
            int j = jg[k];
            if (active[i] && active[j]) { g.ptr[i + 1]++; g.ptr[j + 1]++; }
        }
    for (int i = 0; i < n; ++i) g.ptr[i + 1] += g.ptr[i];
    g.adj.resize(g.ptr[n]);
    std::vector<int64_t> fill(g.ptr.begin(), g.ptr.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int k = ig[i]; k < ig[i + 1]; ++k) {
            int j = jg[k];
            if (active[i] && active[j]) { g.adj[fill[i]++] = j; g.adj[fill[j]++] = i; }
        }
    return g;
}

// ------------------------------------------------------------------ geometric dissection
void nd_geom(int* b, int* e, const std::vector<GridCoord>& C, int leaf, std::vector<int>& order,
             std::vector<int>& fsize) {
    const int cnt = int(e - b);
    if (cnt == 0) return;
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (int* it = b; it != e; ++it) {
        const GridCoord& c = C[*it];
        const int v[3] = {c.x, c.y, c.z};
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], v[a]); hi[a] = std::max(hi[a], v[a]); }
    }
    int axis = -1, cut = 0, ext_best = -1;
    if (cnt > leaf) {
        for (int a = 0; a < 3; ++a) {
            int mid = (lo[a] + hi[a]) / 2;
            int c = mid & ~1;
            if (c <= lo[a]) c += 2;
            if (c >= hi[a]) continue;
            if (hi[a] - lo[a] > ext_best) { ext_best = hi[a] - lo[a]; axis = a; cut = c; }
        }
    }
    if (axis < 0) {
        std::sort(b, e);
        order.insert(order.end(), b, e);
        fsize.push_back(cnt);
        return;
    }
    auto coord = [&](int d) { const GridCoord& c = C[d]; return axis == 0 ? c.x : axis == 1 ? c.y : c.z; };
    int* m1 = std::partition(b, e, [&](int d) { return coord(d) < cut; });
    int* m2 = std::partition(m1, e, [&](int d) { return coord(d) > cut; });
    nd_geom(b, m1, C, leaf, order, fsize);
    nd_geom(m1, m2, C, leaf, order, fsize);
    if (m2 < e) {
        std::sort(m2, e);
        order.insert(order.end(), m2, e);
        fsize.push_back(int(e - m2));
    }
}

// ------------------------------------------------------------------ graph dissection
struct GraphND {
    const Graph& g;
    int leaf;
    std::vector<int>& order;
    std::vector<int>& fsize;
    std::vector<int> mark, dist;
    int stamp = 0;

    GraphND(const Graph& g_, int n, int leaf_, std::vector<int>& o, std::vector<int>& f)
        : g(g_), leaf(leaf_), order(o), fsize(f), mark(n, 0), dist(n, -1) {}

    // BFS inside the set marked `in`; returns visit order, fills dist.
    std::vector<int> bfs(int src, int in) {
        std::vector<int> q{src};
        dist[src] = 0;
        int vis = ++stamp;
        mark[src] = vis;
        for (size_t h = 0; h < q.size(); ++h) {
            int v = q[h];
            for (int64_t k = g.ptr[v]; k < g.ptr[v + 1]; ++k) {
                int w = g.adj[k];
                if (mark[w] == in) {
                    mark[w] = vis;
                    dist[w] = dist[v] + 1;
                    q.push_back(w);
                }
            }
        }
        // restore membership of the visited nodes
        for (int v : q) mark[v] = in;
        return q;
    }

    void emit(std::vector<int>& s) {
        std::sort(s.begin(), s.end());
        order.insert(order.end(), s.begin(), s.end());
        fsize.push_back(int(s.size()));
    }

    void run(std::vector<int> set) {
        if (set.empty()) return;
        if ((int)set.size() <= leaf) { emit(set); return; }
        int in = ++stamp;
        for (int v : set) mark[v] = in;
        // connected components
        std::vector<std::vector<int>> comps;
        int done = ++stamp;
        for (int v : set) {
            if (mark[v] != in) continue;
            std::vector<int> comp = bfs(v, in);
            for (int w : comp) mark[w] = done;
            comps.push_back(std::move(comp));
        }
        for (int v : set) mark[v] = 0;
        if (comps.size() > 1) {
            std::vector<int> small;
            for (auto& c : comps) {
                if ((int)c.size() <= leaf) {
                    if ((int)(small.size() + c.size()) > leaf && !small.empty()) { emit(small); small.clear(); }
                    small.insert(small.end(), c.begin(), c.end());
                } else {
                    run(std::move(c));
                }
            }
            if (!small.empty()) emit(small);
            return;
        }
        // single component: level structure from a pseudo-peripheral node
        in = ++stamp;
        for (int v : set) mark[v] = in;
        std::vector<int> q = bfs(set[0], in);
        int far = q.back();
        q = bfs(far, in);
        far = q.back();
        q = bfs(far, in);
        int depth = dist[q.back()];
        for (int v : set) mark[v] = 0;
        if (depth < 2) { emit(set); return; }
        std::vector<int> cnt(depth + 1, 0);
        for (int v : q) cnt[dist[v]]++;
        int half = int(q.size()) / 2, acc = 0, sl = 1;
        for (int l = 0; l <= depth; ++l) {
            if (acc + cnt[l] > half) { sl = l; break; }
            acc += cnt[l];
        }
        sl = std::max(1, std::min(sl, depth - 1));
        std::vector<int> left, right, sep;
        for (int v : q) {
            int d = dist[v];
            (d < sl ? left : d > sl ? right : sep).push_back(v);
        }
        run(std::move(left));
        run(std::move(right));
        emit(sep);
    }
};

}  // namespace

std::vector<GridCoord> edge_grid_coords(int kuzlov, const double* xyz, int n_edges, const int* edges) {
    std::vector<double> ux, uy, uz;
    for (int i = 0; i < kuzlov; ++i) { ux.push_back(xyz[3 * i]); uy.push_back(xyz[3 * i + 1]); uz.push_back(xyz[3 * i + 2]); }
    auto uniq = [](std::vector<double>& v) { std::sort(v.begin(), v.end()); v.erase(std::unique(v.begin(), v.end()), v.end()); };
    uniq(ux); uniq(uy); uniq(uz);
    const int nx = int(ux.size()), ny = int(uy.size()), nz = int(uz.size());
    if ((int64_t)nx * ny * nz != kuzlov) return {};
    // tensor grid with x fastest: node i has indices (i % nx, (i / nx) % ny, i / (nx*ny))
    for (int i = 0; i < kuzlov; ++i) {
        int ix = i % nx, iy = (i / nx) % ny, iz = i / (nx * ny);
        if (xyz[3 * i] != ux[ix] || xyz[3 * i + 1] != uy[iy] || xyz[3 * i + 2] != uz[iz]) return {};
    }
    std::vector<GridCoord> c(n_edges);
    for (int e = 0; e < n_edges; ++e) {
        int a = edges[2 * e], b = edges[2 * e + 1];
        c[e].x = a % nx + b % nx;
        c[e].y = (a / nx) % ny + (b / nx) % ny;
        c[e].z = a / (nx * ny) + b / (nx * ny);
    }
    return c;
}

Symbolic analyze(int n, const int* ig, const int* jg, const std::vector<char>& active,
                 const std::vector<GridCoord>& coords, const OrderingOptions& opt) {
    Symbolic S;
    S.n = n;
    std::vector<int> act;
    act.reserve(n);
    for (int i = 0; i < n; ++i)
        if (active[i]) act.push_back(i);
    std::vector<int> order, fsize;
    order.reserve(n);
    if (!coords.empty()) {
        nd_geom(act.data(), act.data() + act.size(), coords, opt.leaf_size, order, fsize);
    } else {
        Graph g = build_graph(n, ig, jg, active);
        GraphND nd(g, n, opt.leaf_size, order, fsize);
        nd.run(act);
    }
    S.nn = int(order.size());
    if (S.nn != (int)act.size()) throw Error(0x0080, "ordering lost DOFs");
    S.perm = order;
    for (int i = 0; i < n; ++i)
        if (!active[i]) S.perm.push_back(i);
    S.iperm.assign(n, -1);
    for (int k = 0; k < n; ++k) S.iperm[S.perm[k]] = k;

    const int nf = int(fsize.size());
    S.nf = nf;
    S.fbeg.resize(nf + 1);
    S.fbeg[0] = 0;
    for (int f = 0; f < nf; ++f) S.fbeg[f + 1] = S.fbeg[f] + fsize[f];
    std::vector<int> front_of(S.nn);
    for (int f = 0; f < nf; ++f)
        for (int k = S.fbeg[f]; k < S.fbeg[f + 1]; ++k) front_of[k] = f;

    // lower structure by column in the new numbering
    std::vector<int64_t> cptr(S.nn + 1, 0);
    for (int i = 0; i < n; ++i)
        for (int k = ig[i]; k < ig[i + 1]; ++k) {
            int a = S.iperm[i], b = S.iperm[jg[k]];
            if (a >= S.nn || b >= S.nn) continue;
            cptr[std::min(a, b) + 1]++;
        }
    for (int j = 0; j < S.nn; ++j) cptr[j + 1] += cptr[j];
    std::vector<int> crow(cptr[S.nn]);
    {
        std::vector<int64_t> fill(cptr.begin(), cptr.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int k = ig[i]; k < ig[i + 1]; ++k) {
                int a = S.iperm[i], b = S.iperm[jg[k]];
                if (a >= S.nn || b >= S.nn) continue;
                crow[fill[std::min(a, b)]++] = std::max(a, b);
            }
    }

    // multifrontal symbolic factorization
    std::vector<std::vector<int>> U(nf), kids(nf);
    std::vector<int> mark(S.nn, -1);
    S.parent.assign(nf, -1);
    S.level.assign(nf, 0);
    for (int f = 0; f < nf; ++f) {
        const int fb = S.fbeg[f], fe = S.fbeg[f + 1];
        std::vector<int>& u = U[f];
        for (int j = fb; j < fe; ++j)
            for (int64_t k = cptr[j]; k < cptr[j + 1]; ++k) {
                int r = crow[k];
                if (r >= fe && mark[r] != f) { mark[r] = f; u.push_back(r); }
            }
        for (int c : kids[f]) {
            for (int r : U[c])
                if (r >= fe && mark[r] != f) { mark[r] = f; u.push_back(r); }
            S.level[f] = std::max(S.level[f], S.level[c] + 1);
        }
        std::sort(u.begin(), u.end());
        if (!u.empty()) {
            int pf = front_of[u[0]];
            S.parent[f] = pf;
            kids[pf].push_back(f);
        }
    }

    // flatten
    S.pcount.resize(nf);
    S.mcount.resize(nf);
    S.rows_ptr.assign(nf + 1, 0);
    S.child_ptr.assign(nf + 1, 0);
    S.relmap_ptr.assign(nf + 1, 0);
    S.loff.resize(nf);
    S.uoff.resize(nf);
    S.woff.resize(nf);
    int64_t lo = 0, uo = 0, wo = 0;
    for (int f = 0; f < nf; ++f) {
        const int p = S.fbeg[f + 1] - S.fbeg[f];
        const int u = int(U[f].size());
        const int m = p + u;
        S.pcount[f] = p;
        S.mcount[f] = m;
        S.rows_ptr[f + 1] = S.rows_ptr[f] + m;
        S.child_ptr[f + 1] = S.child_ptr[f] + int(kids[f].size());
        S.relmap_ptr[f + 1] = S.relmap_ptr[f] + u;
        S.loff[f] = lo;
        S.uoff[f] = uo;
        S.woff[f] = wo;
        lo += int64_t(m) * p;
        uo += int64_t(u) * u;
        wo += m;
        S.max_p = std::max(S.max_p, p);
        S.max_m = std::max(S.max_m, m);
        S.max_children = std::max(S.max_children, int(kids[f].size()));
        const double P = p, Uu = u;
        S.flops += P * P * P / 3.0 + P * P * Uu + P * Uu * Uu;
        S.nnz_l += P * (P + 1) / 2.0 + P * Uu;
    }
    S.lsize = lo;
    S.usize = uo;
    S.wsize = wo;
    S.rows.resize(S.rows_ptr[nf]);
    S.children.resize(S.child_ptr[nf]);
    S.relmap.resize(S.relmap_ptr[nf]);
    for (int f = 0; f < nf; ++f) {
        int64_t o = S.rows_ptr[f];
        for (int k = S.fbeg[f]; k < S.fbeg[f + 1]; ++k) S.rows[o++] = k;
        for (int r : U[f]) S.rows[o++] = r;
        std::copy(kids[f].begin(), kids[f].end(), S.children.begin() + S.child_ptr[f]);
    }
    // relative maps: child's update rows inside the parent's rows
    for (int c = 0; c < nf; ++c) {
        const int f = S.parent[c];
        if (f < 0) continue;
        const int fb = S.fbeg[f], fe = S.fbeg[f + 1], p = fe - fb;
        const std::vector<int>& uf = U[f];
        int64_t o = S.relmap_ptr[c];
        for (int r : U[c]) {
            int pos;
            if (r < fe) {
                pos = r - fb;
            } else {
                auto it = std::lower_bound(uf.begin(), uf.end(), r);
                pos = p + int(it - uf.begin());
            }
            S.relmap[o++] = pos;
        }
    }
    // group by level
    S.nlevels = 0;
    for (int f = 0; f < nf; ++f) S.nlevels = std::max(S.nlevels, S.level[f] + 1);
    S.level_ptr.assign(S.nlevels + 1, 0);
    for (int f = 0; f < nf; ++f) S.level_ptr[S.level[f] + 1]++;
    for (int l = 0; l < S.nlevels; ++l) S.level_ptr[l + 1] += S.level_ptr[l];
    S.level_fronts.resize(nf);
    {
        std::vector<int> fill(S.level_ptr.begin(), S.level_ptr.end() - 1);
        for (int f = 0; f < nf; ++f) S.level_fronts[fill[S.level[f]]++] = f;
    }
    place_update_matrices(S);
    return S;
}

// U_f is written when front f is factored (its level) and read by the extend-add into its
// parent, which finishes at the end of the parent's level; outside [level(f), level(parent)]
// its storage is free. Offsets are assigned level by level, first fit over a free list of
// intervals (largest matrices of a level first, 64-element alignment), which reduces the U
// pool from the sum of all update matrices to the peak of the live ones.
void place_update_matrices(Symbolic& S) {
    const int nf = int(S.pcount.size());
    std::map<int64_t, int64_t> free_list;  // offset -> length
    int64_t top = 0;
    std::vector<std::vector<int>> dies(S.nlevels + 1);
    auto release = [&](int64_t off, int64_t len) {
        auto it = free_list.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_list.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_list.erase(nx);
        }
        if (it != free_list.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_list.erase(it);
            }
        }
    };
    auto size_of = [&](int f) {
        const int64_t u = S.mcount[f] - S.pcount[f];
        return (u * u + 63) / 64 * 64;
    };
    for (int l = 0; l < S.nlevels; ++l) {
        for (int f : dies[l]) release(S.uoff[f], size_of(f));
        std::vector<int> fl(S.level_fronts.begin() + S.level_ptr[l], S.level_fronts.begin() + S.level_ptr[l + 1]);
        std::stable_sort(fl.begin(), fl.end(), [&](int a, int b) { return size_of(a) > size_of(b); });
        for (int f : fl) {
            const int64_t need = size_of(f);
            if (need == 0 || S.parent[f] < 0) {
                S.uoff[f] = 0;
                continue;
            }
            int64_t off = -1;
            for (auto it = free_list.begin(); it != free_list.end(); ++it)
                if (it->second >= need) {
                    off = it->first;
                    const int64_t rest = it->second - need;
                    free_list.erase(it);
                    if (rest > 0) free_list.emplace(off + need, rest);
                    break;
                }
            if (off < 0) {
                // extend the pool; a free interval ending at the top is merged into the new block
                if (!free_list.empty()) {
                    auto last = std::prev(free_list.end());
                    if (last->first + last->second == top) {
                        off = last->first;
                        free_list.erase(last);
                    }
                }
                if (off < 0) off = top;
                top = off + need;
            }
            S.uoff[f] = off;
            // freed after the parent's level (the extend-add into the parent runs at that level)
            const int dl = S.level[S.parent[f]] + 1;
            if (dl <= S.nlevels) dies[dl].push_back(f);
        }
    }
    S.usize = top;
}

void assembly_map(const Symbolic& S, int n, const int* ig, const int* jg,
                  std::vector<int64_t>& dest_off, std::vector<int64_t>& dest_diag) {
    std::vector<int> front_of(S.nn);
    for (int f = 0; f < S.nf; ++f)
        for (int k = S.fbeg[f]; k < S.fbeg[f + 1]; ++k) front_of[k] = f;
    std::atomic<bool> bad{false};
    auto pos_in_front = [&](int f, int r) -> int {
        const int fb = S.fbeg[f], fe = S.fbeg[f + 1];
        if (r < fe) return r - fb;
        const int* b = S.rows.data() + S.rows_ptr[f] + (fe - fb);
        const int* e = S.rows.data() + S.rows_ptr[f + 1];
        const int* it = std::lower_bound(b, e, r);
        if (it == e || *it != r) {
            bad.store(true, std::memory_order_relaxed);
            return 0;
        }
        return (fe - fb) + int(it - b);
    };
    dest_off.assign(ig[n], -1);
    dest_diag.assign(n, -1);
    // rows are independent: each writes its own entries
    auto work = [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i) {
            int a = S.iperm[i];
            if (a < S.nn) {
                int f = front_of[a];
                int lc = a - S.fbeg[f];
                dest_diag[i] = S.loff[f] + int64_t(lc) * S.mcount[f] + lc;
            }
            for (int k = ig[i]; k < ig[i + 1]; ++k) {
                int b = S.iperm[jg[k]];
                if (a >= S.nn || b >= S.nn) continue;
                int col = std::min(a, b), row = std::max(a, b);
                int f = front_of[col];
                int lc = col - S.fbeg[f];
                int lr = pos_in_front(f, row);
                dest_off[k] = S.loff[f] + int64_t(lc) * S.mcount[f] + lr;
            }
        }
    };
    const int nt = std::max(1, std::min<int>(8, int(std::thread::hardware_concurrency())));
    if (nt > 1 && n > 16384) {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) th.emplace_back(work, int(int64_t(n) * t / nt), int(int64_t(n) * (t + 1) / nt));
        for (auto& x : th) x.join();
    } else {
        work(0, n);
    }
    if (bad) throw Error(0x0080, "assembly map: row outside front structure");
}

}  // namespace emtd
