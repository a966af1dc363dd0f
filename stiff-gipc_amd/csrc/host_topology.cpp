// Host-side topology preprocessing for the MAS preconditioner.
//   - vertex neighbour CSR from tets: StiffGIPC/load_mesh.cpp:934-1040
//   - bank slot maps from ascending part ids: StiffGIPC/gl_main.cu:1066-1093
//   - level count: StiffGIPC/MASPreconditioner.cu:1724-1743
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace sgipc {

// Edges are deduplicated regardless of orientation, keeping the orientation
// first met (tet order, then the pair order (0,1),(0,2),(0,3),(1,2),(1,3),
// (2,3)); the edge set is walked in lexicographic (a,b) order and each edge
// appends b to a's list and a to b's list.
int build_vertex_neighbors(int n, int T, const uint32_t* tets, uint32_t* start, uint32_t* num, uint32_t* list,
                           long long cap) {
    struct Tag {
        uint32_t lo, hi, a, b;
        uint64_t pos;
    };
    static const int pr[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    std::vector<Tag> tg;
    tg.reserve((size_t)6 * T);
    for (int t = 0; t < T; ++t)
        for (int p = 0; p < 6; ++p) {
            uint32_t a = tets[4 * (size_t)t + pr[p][0]], b = tets[4 * (size_t)t + pr[p][1]];
            if (a >= (uint32_t)n || b >= (uint32_t)n) throw std::invalid_argument("tet vertex index out of range");
            tg.push_back({std::min(a, b), std::max(a, b), a, b, (uint64_t)tg.size()});
        }
    std::sort(tg.begin(), tg.end(), [](const Tag& x, const Tag& y) {
        if (x.lo != y.lo) return x.lo < y.lo;
        if (x.hi != y.hi) return x.hi < y.hi;
        return x.pos < y.pos;
    });
    std::vector<std::pair<uint32_t, uint32_t>> e;
    e.reserve(tg.size());
    for (size_t i = 0; i < tg.size(); ++i) {
        if (i > 0 && tg[i].lo == tg[i - 1].lo && tg[i].hi == tg[i - 1].hi) continue;
        e.emplace_back(tg[i].a, tg[i].b);
    }
    std::sort(e.begin(), e.end());
    std::fill(num, num + n, 0u);
    for (auto& ed : e) { num[ed.first]++; num[ed.second]++; }
    uint32_t off = 0;
    for (int i = 0; i < n; ++i) { start[i] = off; off += num[i]; }
    if ((long long)off > cap) throw std::length_error("neighbour list capacity too small");
    std::vector<uint32_t> cur(n, 0);
    for (auto& ed : e) {
        list[start[ed.first] + cur[ed.first]++] = ed.second;
        list[start[ed.second] + cur[ed.second]++] = ed.first;
    }
    return (int)off;
}

int build_bank_map(int n, const int* part, int bank, int* part_map, long long cap, int* real_map) {
    if (n <= 0) return 0;
    for (int i = 1; i < n; ++i)
        if (part[i] < part[i - 1]) throw std::invalid_argument("part ids must be ascending");
    if (part[0] < 0) throw std::invalid_argument("negative part id");
    long long M = (long long)bank * ((long long)part[n - 1] + 1);
    if (M > cap) throw std::length_error("part map capacity too small");
    std::fill(part_map, part_map + M, -1);
    int rank = 0;
    for (int i = 0; i < n; ++i) {
        if (rank >= bank) throw std::invalid_argument("a part holds more vertices than the bank size");
        part_map[(size_t)bank * part[i] + rank] = i;
        rank++;
        if (i + 1 < n && part[i + 1] != part[i]) rank = 0;
    }
    int index = 0;
    for (long long i = 0; i < M; ++i)
        if (part_map[i] == index) { real_map[index] = (int)i; index++; }
    if (index != n) throw std::invalid_argument("bank map is not monotone");
    return (int)M;
}

int mas_num_levels(int nn, int bank) {
    int nLevel = 1;
    int levelSz = (nn + bank - 1) / bank * bank;
    while (levelSz > bank) {
        levelSz /= bank;
        nLevel++;
        levelSz = (levelSz + bank - 1) / bank * bank;
    }
    nLevel = nLevel + 1;
    return nLevel > 6 ? 6 : nLevel;
}

}  // namespace sgipc
