// Vertex partition of a tet mesh into subdomains of at most `bank` vertices
// for the MAS preconditioner, and the part-contiguous renumbering.
//
// Reference: metis_sort (MeshProcess/metis_partition/src/metis_sort.cpp:
// 25-383, tet branch :217-290): the vertex graph with edge weight = number of
// tets holding the edge (TetMesh::_build_adj, mesh.cpp:486-549), k-way
// partitioned by METIS_PartGraphKway (NodeEdgeModel::k_way_partition,
// node_edge_model.cpp:5-28) into nParts = ceil(n / (bank - off)) parts,
// retried with off + 1 until no part exceeds `bank`; then the vertices sorted
// by part id (TetMesh::sort_index, mesh.cpp:65-88).
//
// The k-way step here is this library's own partitioner (METIS is not used):
// recursive bisection, each bisection multilevel --
//   coarsen by heavy-edge matching (vertex weights = fine vertices merged),
//   bisect the coarsest graph by greedy graph growing from several seeds
//   (each step adds the frontier vertex with the largest gain),
//   project back level by level with Fiduccia-Mattheyses boundary refinement
//   (single moves in gain order inside the allowed side weights, rolled back
//   to the best cut of the pass).
// then a greedy k-way pass moves boundary vertices to the neighbouring part
// they are most connected to where that lowers the cut (parts stay within
// [n/k - 1, bank]).  A bisection of a vertex set due to become k parts gives
// its first side k1 = k / 2 parts and a weight within k1 / 2 of the
// proportional split and between n - k2 bank and k1 bank, so every part ends
// with at most `bank` vertices.  The offset loop starts at off = 1: one
// vertex of slack per part, which the refinement uses to lower the cut (METIS
// needs it on the reference's meshes too: high_mat stops at offset 1); with
// parts that always fit, the loop stops at its first try.  Seeds and visit
// orders come from a fixed-seed generator: the result is deterministic.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <queue>
#include <random>
#include <stdexcept>
#include <vector>

namespace sgipc {
namespace {

struct Graph {
    int n = 0;
    std::vector<int> xadj, adj, w;  // CSR with edge weights
    std::vector<int> vw;            // vertex weights
    int total_w() const { return std::accumulate(vw.begin(), vw.end(), 0); }
};

// induced subgraph of `verts` (global ids) of the fine graph G
Graph induced(const Graph& G, const std::vector<int>& verts, std::vector<int>& local) {
    Graph S;
    S.n = (int)verts.size();
    for (int i = 0; i < S.n; ++i) local[verts[i]] = i;
    S.xadj.assign(S.n + 1, 0);
    for (int i = 0; i < S.n; ++i) {
        const int v = verts[i];
        for (int e = G.xadj[v]; e < G.xadj[v + 1]; ++e)
            if (local[G.adj[e]] >= 0) {
                S.adj.push_back(local[G.adj[e]]);
                S.w.push_back(G.w[e]);
            }
        S.xadj[i + 1] = (int)S.adj.size();
    }
    S.vw.assign(S.n, 1);
    for (int i = 0; i < S.n; ++i) local[verts[i]] = -1;
    return S;
}

// heavy-edge matching; cmap[v] = coarse vertex of v.  Returns the coarse
// graph (parallel edges merged, weights summed).
Graph coarsen(const Graph& G, std::vector<int>& cmap, int max_vw, std::mt19937& rng) {
    const int n = G.n;
    std::vector<int> order(n), match(n, -1);
    std::iota(order.begin(), order.end(), 0);
    std::shuffle(order.begin(), order.end(), rng);
    for (int v : order) {
        if (match[v] >= 0) continue;
        int best = -1, bw = -1;
        for (int e = G.xadj[v]; e < G.xadj[v + 1]; ++e) {
            const int u = G.adj[e];
            if (match[u] >= 0 || u == v || G.vw[u] + G.vw[v] > max_vw) continue;
            if (G.w[e] > bw || (G.w[e] == bw && G.vw[u] < G.vw[best])) { bw = G.w[e]; best = u; }
        }
        if (best >= 0) { match[v] = best; match[best] = v; }
        else match[v] = v;
    }
    cmap.assign(n, -1);
    int nc = 0;
    for (int v = 0; v < n; ++v)
        if (cmap[v] < 0) {
            cmap[v] = nc;
            cmap[match[v]] = nc;
            ++nc;
        }
    Graph C;
    C.n = nc;
    C.vw.assign(nc, 0);
    for (int v = 0; v < n; ++v) C.vw[cmap[v]] += G.vw[v];
    // edges of each coarse vertex: both fine members' edges, merged
    std::vector<int> first(nc, -1);
    for (int v = 0; v < n; ++v)
        if (first[cmap[v]] < 0) first[cmap[v]] = v;
    std::vector<int> slot(nc, -1);
    C.xadj.assign(nc + 1, 0);
    for (int c = 0; c < nc; ++c) {
        const int a = first[c], b = match[a];
        const size_t beg = C.adj.size();
        for (int m = 0; m < (a == b ? 1 : 2); ++m) {
            const int v = m == 0 ? a : b;
            for (int e = G.xadj[v]; e < G.xadj[v + 1]; ++e) {
                const int cu = cmap[G.adj[e]];
                if (cu == c) continue;
                if (slot[cu] < 0) {
                    slot[cu] = (int)C.adj.size();
                    C.adj.push_back(cu);
                    C.w.push_back(G.w[e]);
                } else {
                    C.w[slot[cu]] += G.w[e];
                }
            }
        }
        for (size_t e = beg; e < C.adj.size(); ++e) slot[C.adj[e]] = -1;
        C.xadj[c + 1] = (int)C.adj.size();
    }
    return C;
}

long long cut_of(const Graph& G, const std::vector<char>& side) {
    long long cut = 0;
    for (int v = 0; v < G.n; ++v)
        for (int e = G.xadj[v]; e < G.xadj[v + 1]; ++e)
            if (side[v] != side[G.adj[e]]) cut += G.w[e];
    return cut / 2;
}

// gain of moving v to the other side: edge weight to the other side minus
// edge weight to its own side
int gain_of(const Graph& G, const std::vector<char>& side, int v) {
    int g = 0;
    for (int e = G.xadj[v]; e < G.xadj[v + 1]; ++e) g += side[G.adj[e]] != side[v] ? G.w[e] : -G.w[e];
    return g;
}

// Greedy graph growing: side 1 grows from `seed` by the frontier vertex of
// largest gain until its weight reaches `target` (restarting from an
// unvisited vertex when the frontier empties).
void grow(const Graph& G, int seed, int target, std::vector<char>& side, std::mt19937& rng) {
    const int n = G.n;
    side.assign(n, 0);
    std::vector<int> conn(n, 0), degw(n, 0);  // edge weight into side 1; total edge weight
    for (int v = 0; v < n; ++v)
        for (int e = G.xadj[v]; e < G.xadj[v + 1]; ++e) degw[v] += G.w[e];
    auto key = [&](int u) { return 2 * conn[u] - degw[u]; };  // cut decrease if u joins side 1
    typedef std::pair<int, int> Item;                        // (gain, -vertex)
    std::priority_queue<Item> heap;
    int wt = 0;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::shuffle(order.begin(), order.end(), rng);
    size_t next_start = 0;
    auto add = [&](int v) {
        side[v] = 1;
        wt += G.vw[v];
        for (int e = G.xadj[v]; e < G.xadj[v + 1]; ++e) {
            const int u = G.adj[e];
            if (side[u]) continue;
            conn[u] += G.w[e];
            heap.push({key(u), -u});
        }
    };
    auto pop_valid = [&](int& v) {
        while (!heap.empty()) {
            const Item it = heap.top();
            heap.pop();
            const int u = -it.second;
            if (side[u] || it.first != key(u)) continue;  // stale entry
            v = u;
            return true;
        }
        return false;
    };
    add(seed);
    while (wt < target) {
        int v = -1;
        if (!pop_valid(v)) {  // frontier empty: start a new component
            while (next_start < order.size() && side[order[next_start]]) ++next_start;
            if (next_start == order.size()) break;
            v = order[next_start];
        }
        if (wt + G.vw[v] > target) {
            // too heavy: take the best frontier vertex that still fits
            std::vector<Item> keep{{key(v), -v}};
            int u = -1;
            bool found = false;
            while (pop_valid(u)) {
                if (wt + G.vw[u] <= target) { found = true; break; }
                keep.push_back({key(u), -u});
            }
            for (auto& k : keep) heap.push(k);
            if (!found) break;
            v = u;
        }
        add(v);
    }
}

// FM refinement of a bisection: side-1 weight kept in [lo, hi].
// A pass moves boundary vertices one at a time, largest gain first (each at
// most once), and keeps the prefix of moves with the smallest cut.  Moves
// that repair a weight outside [lo, hi] are taken first.
void refine(const Graph& G, std::vector<char>& side, int lo, int hi, int passes) {
    const int n = G.n;
    int w1 = 0;
    for (int v = 0; v < n; ++v) w1 += side[v] ? G.vw[v] : 0;
    std::vector<int> gain(n);
    std::vector<char> moved(n);
    std::vector<int> moves;
    for (int pass = 0; pass < passes; ++pass) {
        for (int v = 0; v < n; ++v) gain[v] = gain_of(G, side, v);
        std::fill(moved.begin(), moved.end(), 0);
        typedef std::pair<int, int> Item;
        std::priority_queue<Item> heap[2];  // by source side
        for (int v = 0; v < n; ++v) {
            bool boundary = false;
            for (int e = G.xadj[v]; e < G.xadj[v + 1] && !boundary; ++e) boundary = side[G.adj[e]] != side[v];
            if (boundary || w1 < lo || w1 > hi) heap[(int)side[v]].push({gain[v], -v});
        }
        long long cur = 0, best = 0;
        int best_len = 0;
        bool best_balanced = w1 >= lo && w1 <= hi;
        moves.clear();
        int since_best = 0;
        while (since_best < 64 + n / 100) {
            // pick the source side: forced by the balance, else the better top
            int src = -1;
            auto top = [&](int s, Item& out) {
                while (!heap[s].empty()) {
                    const Item it = heap[s].top();
                    const int v = -it.second;
                    if (moved[v] || side[v] != s || it.first != gain[v]) { heap[s].pop(); continue; }
                    out = it;
                    return true;
                }
                return false;
            };
            Item t0, t1;
            const bool h0 = top(0, t0), h1 = top(1, t1);
            // moving from side 0 raises w1, from side 1 lowers it
            const bool ok0 = h0 && w1 + G.vw[-t0.second] <= hi;
            const bool ok1 = h1 && w1 - G.vw[-t1.second] >= lo;
            if (w1 < lo) src = h0 ? 0 : -1;
            else if (w1 > hi) src = h1 ? 1 : -1;
            else if (ok0 && ok1) src = t0.first >= t1.first ? 0 : 1;
            else if (ok0) src = 0;
            else if (ok1) src = 1;
            if (src < 0) break;
            const Item it = src == 0 ? t0 : t1;
            heap[src].pop();
            const int v = -it.second;
            moved[v] = 1;
            cur -= gain[v];
            side[v] ^= 1;
            w1 += side[v] ? G.vw[v] : -G.vw[v];
            moves.push_back(v);
            gain[v] = -gain[v];
            for (int e = G.xadj[v]; e < G.xadj[v + 1]; ++e) {
                const int u = G.adj[e];
                gain[u] += side[u] == side[v] ? -2 * G.w[e] : 2 * G.w[e];
                if (!moved[u]) heap[(int)side[u]].push({gain[u], -u});
            }
            const bool bal = w1 >= lo && w1 <= hi;
            if ((bal && !best_balanced) || (bal && cur < best)) {
                best = cur;
                best_len = (int)moves.size();
                best_balanced = true;
                since_best = 0;
            } else {
                ++since_best;
            }
        }
        // roll back past the best prefix
        for (int i = (int)moves.size() - 1; i >= best_len; --i) {
            const int v = moves[i];
            side[v] ^= 1;
            w1 += side[v] ? G.vw[v] : -G.vw[v];
        }
        if (best_len == 0) break;
    }
}

// multilevel bisection of G: side-1 weight in [lo, hi], aiming at target
void bisect(const Graph& G, int lo, int hi, int target, std::vector<char>& side, std::mt19937& rng) {
    std::vector<Graph> levels;
    std::vector<std::vector<int>> maps;
    levels.push_back(G);
    const int total = G.total_w();
    const int max_vw = std::max(2, total / 32);  // coarse vertices stay light against the side weights
    while (levels.back().n > 120) {
        std::vector<int> cmap;
        Graph C = coarsen(levels.back(), cmap, max_vw, rng);
        if (C.n > 0.9 * levels.back().n) break;
        maps.push_back(std::move(cmap));
        levels.push_back(std::move(C));
    }
    const Graph& Cg = levels.back();
    // initial bisection: best of several grown seeds
    std::vector<char> best, trial;
    long long best_cut = -1;
    bool best_ok = false;
    const int tries = std::min(Cg.n, 8);
    std::uniform_int_distribution<int> pick(0, Cg.n - 1);
    for (int t = 0; t < tries; ++t) {
        grow(Cg, pick(rng), target, trial, rng);
        refine(Cg, trial, lo, hi, 4);
        int w1 = 0;
        for (int v = 0; v < Cg.n; ++v) w1 += trial[v] ? Cg.vw[v] : 0;
        const bool ok = w1 >= lo && w1 <= hi;
        const long long c = cut_of(Cg, trial);
        if (best_cut < 0 || (ok && !best_ok) || (ok == best_ok && c < best_cut)) {
            best = trial;
            best_cut = c;
            best_ok = ok;
        }
    }
    side = best;
    for (int L = (int)levels.size() - 2; L >= 0; --L) {
        std::vector<char> fine(levels[L].n);
        for (int v = 0; v < levels[L].n; ++v) fine[v] = side[maps[L][v]];
        side.swap(fine);
        refine(levels[L], side, lo, hi, L == 0 ? 6 : 3);
    }
}

// recursive bisection of `verts` into k parts numbered from part0
void split(const Graph& G, std::vector<int>& verts, int k, int cap, int part0, std::vector<int>& part,
           std::vector<int>& local, std::mt19937& rng) {
    const int n = (int)verts.size();
    if (k <= 1 || n <= 1) {
        for (int v : verts) part[v] = part0;
        return;
    }
    const int k1 = k / 2, k2 = k - k1;
    const int flo = std::max(1, n - k2 * cap), fhi = std::min(n - 1, k1 * cap);  // feasible side weights
    if (flo > fhi) throw std::logic_error("partition: infeasible bisection");
    int target = (int)(((long long)n * k1 + k / 2) / k);
    target = std::min(fhi, std::max(flo, target));
    // balance: within half a vertex per part of the proportional split
    const int tol = std::max(1, k1 / 2);
    const int lo = std::max(flo, target - tol), hi = std::min(fhi, target + tol);
    Graph S = induced(G, verts, local);
    std::vector<char> side;
    bisect(S, lo, hi, target, side, rng);
    int w1 = 0;
    for (char c : side) w1 += c;
    if (w1 < lo || w1 > hi) throw std::logic_error("partition: bisection outside its bounds");
    std::vector<int> a, b;
    a.reserve(w1);
    b.reserve(n - w1);
    for (int i = 0; i < n; ++i) (side[i] ? a : b).push_back(verts[i]);
    std::vector<int>().swap(verts);
    split(G, a, k1, cap, part0, part, local, rng);
    split(G, b, k2, cap, part0 + k1, part, local, rng);
}

// Greedy k-way refinement: a boundary vertex moves to the neighbouring part
// it is most connected to when that lowers the cut and the part has room
// (at most cap vertices; the source keeps at least one).  Passes repeat
// until one moves nothing (at most `passes`).
void kway_refine(const Graph& G, std::vector<int>& part, int k, int cap, int min_size, int passes,
                 std::mt19937& rng) {
    const int n = G.n;
    std::vector<int> size(k, 0);
    for (int v = 0; v < n; ++v) size[part[v]]++;
    std::vector<int> order(n), conn(k, 0), touched;
    std::iota(order.begin(), order.end(), 0);
    for (int pass = 0; pass < passes; ++pass) {
        std::shuffle(order.begin(), order.end(), rng);
        int moved = 0;
        for (int v : order) {
            const int own = part[v];
            touched.clear();
            for (int e = G.xadj[v]; e < G.xadj[v + 1]; ++e) {
                const int q = part[G.adj[e]];
                if (conn[q] == 0) touched.push_back(q);
                conn[q] += G.w[e];
            }
            int best = own, bw = conn[own];
            for (int q : touched)
                if (q != own && size[q] < cap && (conn[q] > bw || (conn[q] == bw && best != own && q < best))) {
                    best = q;
                    bw = conn[q];
                }
            const int own_w = conn[own];
            for (int q : touched) conn[q] = 0;
            if (best != own && bw > own_w && size[own] > min_size) {
                part[v] = best;
                size[own]--;
                size[best]++;
                ++moved;
            }
        }
        if (!moved) break;
    }
}

}  // namespace

// Vertex graph of the tets with co-occurrence edge weights (each tet adds 1
// to each of its 6 edges, both directions), neighbours ascending.
void tet_vertex_graph(int n, int T, const uint32_t* tets, std::vector<int>& xadj, std::vector<int>& adj,
                      std::vector<int>& w) {
    std::vector<std::pair<uint64_t, int>> e;
    e.reserve((size_t)12 * T);
    for (int t = 0; t < T; ++t)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                if (a == b) continue;
                const uint32_t i = tets[4 * (size_t)t + a], j = tets[4 * (size_t)t + b];
                if (i >= (uint32_t)n || j >= (uint32_t)n) throw std::invalid_argument("tet vertex index out of range");
                if (i == j) continue;
                e.push_back({((uint64_t)i << 32) | j, 1});
            }
    std::sort(e.begin(), e.end());
    xadj.assign(n + 1, 0);
    adj.clear();
    w.clear();
    for (size_t k = 0; k < e.size();) {
        size_t m = k;
        int c = 0;
        while (m < e.size() && e[m].first == e[k].first) c += e[m++].second;
        adj.push_back((int)(e[k].first & 0xffffffffu));
        w.push_back(c);
        xadj[(e[k].first >> 32) + 1]++;
        k = m;
    }
    for (int i = 0; i < n; ++i) xadj[i + 1] += xadj[i];
}

// Parts of at most `bank` vertices: part[v] in [0, n_parts), then order =
// vertex ids sorted by part (ascending id within a part).  Returns n_parts;
// offset = the reference's metis_offset at which the loop stopped.
int partition_mesh(int n, int T, const uint32_t* tets, int bank, int* part, int* order, int* offset) {
    if (n <= 0) return 0;
    if (bank < 2) throw std::invalid_argument("partition: bank must be at least 2");
    Graph G;
    G.n = n;
    tet_vertex_graph(n, T, tets, G.xadj, G.adj, G.w);
    G.vw.assign(n, 1);
    std::vector<int> p(n, 0), local(n, -1);
    int nparts = 1, off = 0;
    for (off = 1; off < bank; ++off) {
        const int cap = bank - off;
        nparts = (n + cap - 1) / cap;
        if (nparts <= 1) {
            std::fill(p.begin(), p.end(), 0);
            nparts = 1;
        } else {
            std::vector<int> all(n);
            std::iota(all.begin(), all.end(), 0);
            std::mt19937 rng(12345u);
            split(G, all, nparts, bank, 0, p, local, rng);
            kway_refine(G, p, nparts, bank, std::max(1, n / nparts - 1), 20, rng);
        }
        std::vector<int> cnt(nparts, 0);
        for (int v = 0; v < n; ++v) cnt[p[v]]++;
        if (*std::max_element(cnt.begin(), cnt.end()) <= bank) break;
    }
    if (off == bank) throw std::logic_error("partition: no offset gave parts within the bank size");
    for (int v = 0; v < n; ++v) part[v] = p[v];
    if (order) {
        std::vector<int> ord(n);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return p[a] < p[b]; });
        std::copy(ord.begin(), ord.end(), order);
    }
    if (offset) *offset = off;
    return nparts;
}

long long partition_edge_cut(int n, int T, const uint32_t* tets, const int* part) {
    std::vector<int> xadj, adj, w;
    tet_vertex_graph(n, T, tets, xadj, adj, w);
    long long cut = 0;
    for (int v = 0; v < n; ++v)
        for (int e = xadj[v]; e < xadj[v + 1]; ++e)
            if (part[v] != part[adj[e]]) cut += w[e];
    return cut / 2;
}

}  // namespace sgipc
