// Contact-pair broad phase: a linear BVH over the surface triangles and one
// over the surface edges, queried per surface vertex (point-triangle) and
// per edge (edge-edge) for boxes within sqrt(dhat) of each other.  The
// output is the candidate quadruples that the narrow phase
// (sgipc_classify_contacts) classifies and encodes.  With a search
// direction the boxes are swept from x to x - alpha dir (the reference's
// _calcLeafBvs_ccd / _selfQuery_vf_ccd / _selfQuery_ee_ccd, mlbvh.cu
// :1172-1204, :1408-1507, :1634-1760): the candidates of the CCD step bound.
//
// Reference: S/mlbvh.cu -- leaf boxes _calcLeafBvs (:1147-1170), scene box
// _reduct_max_box (:1062-1145), Morton keys _calcMChash (:1206-1225, 10 bits
// per axis of the box centre in the scene box, original index in the low
// 32 bits), Karras internal nodes _calcInternalNodes with determine_range /
// find_split (:126-227, :1247-1267), bottom-up boxes _calcInternalAABB
// (:1269-1296), self queries _selfQuery_vf (:1306-1406) and _selfQuery_ee
// (:1509-1632) with the reference's filters (point not in the face, edges
// sharing no vertex and obj > self, not all four vertices of btype >= 2,
// different bodies or body id -1) and box test `overlap` (:33-42).
// Differences: the reference appends pairs with an atomic counter while it
// traverses; here each query counts its hits, an exclusive scan places
// them, and a second traversal writes them, so the candidate order is
// (query index, traversal order) -- deterministic.  The key sort is
// rocPRIM's radix sort of the 30-bit codes with the leaf index as value
// (stable), which orders exactly as the reference's 64-bit keys.
#include "context.hpp"
#include <rocprim/device/device_radix_sort.hpp>

namespace sgipc {
namespace {

constexpr unsigned kNone = 0xFFFFFFFFu;
constexpr int kStack = 64;

struct Box {
    double lx, ly, lz, ux, uy, uz;
};

__device__ inline Box empty_box() { return Box{1e32, 1e32, 1e32, -1e32, -1e32, -1e32}; }
__device__ inline void grow(Box& b, const double* __restrict__ x, unsigned v) {
    double px = x[3 * (size_t)v], py = x[3 * (size_t)v + 1], pz = x[3 * (size_t)v + 2];
    b.lx = fmin(b.lx, px); b.ly = fmin(b.ly, py); b.lz = fmin(b.lz, pz);
    b.ux = fmax(b.ux, px); b.uy = fmax(b.uy, py); b.uz = fmax(b.uz, pz);
}
// the vertex at x and, when swept, at x - alpha dir
__device__ inline void grow_swept(Box& b, const double* __restrict__ x, const double* __restrict__ dir, double alpha,
                                  unsigned v) {
    grow(b, x, v);
    if (dir) {
        double px = x[3 * (size_t)v] - alpha * dir[3 * (size_t)v];
        double py = x[3 * (size_t)v + 1] - alpha * dir[3 * (size_t)v + 1];
        double pz = x[3 * (size_t)v + 2] - alpha * dir[3 * (size_t)v + 2];
        b.lx = fmin(b.lx, px); b.ly = fmin(b.ly, py); b.lz = fmin(b.lz, pz);
        b.ux = fmax(b.ux, px); b.uy = fmax(b.uy, py); b.uz = fmax(b.uz, pz);
    }
}
__device__ inline Box merge(const Box& a, const Box& b) {
    return Box{fmin(a.lx, b.lx), fmin(a.ly, b.ly), fmin(a.lz, b.lz), fmax(a.ux, b.ux), fmax(a.uy, b.uy),
               fmax(a.uz, b.uz)};
}
// boxes closer than gap on every axis (the reference's overlap with gapL)
__device__ inline bool near(const Box& a, const Box& b, double gap) {
    if (b.lx - a.ux >= gap || a.lx - b.ux >= gap) return false;
    if (b.ly - a.uy >= gap || a.ly - b.uy >= gap) return false;
    if (b.lz - a.uz >= gap || a.lz - b.uz >= gap) return false;
    return true;
}

struct Node {
    unsigned left, right, parent, elem;
};

template <int K>
__global__ void k_leaf_boxes(int m, const unsigned* __restrict__ prim, const double* __restrict__ x,
                             const double* __restrict__ dir, double alpha, Box* __restrict__ box) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    Box b = empty_box();
#pragma unroll
    for (int k = 0; k < K; ++k) grow_swept(b, x, dir, alpha, prim[K * (size_t)i + k]);
    box[i] = b;
}

// scene box: per-block partial boxes, then one block merges them
__global__ void __launch_bounds__(256) k_box_reduce(int m, const Box* __restrict__ in, Box* __restrict__ out) {
    __shared__ Box sh[256];
    Box b = empty_box();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) b = merge(b, in[i]);
    sh[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = merge(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__device__ inline unsigned spread10(unsigned v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
__device__ inline unsigned cell10(double t) {
    t = fmin(fmax(t * 1024.0, 0.0), 1023.0);
    return (unsigned)t;
}

__global__ void k_morton(int m, const Box* __restrict__ box, const Box* __restrict__ scene, unsigned* __restrict__ code,
                         unsigned* __restrict__ idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const Box S = *scene, b = box[i];
    // a flat scene axis maps every centre to cell 0 on it
    double sx = S.ux - S.lx, sy = S.uy - S.ly, sz = S.uz - S.lz;
    double cx = 0.5 * (b.lx + b.ux) - S.lx, cy = 0.5 * (b.ly + b.uy) - S.ly, cz = 0.5 * (b.lz + b.uz) - S.lz;
    unsigned xx = spread10(sx > 0.0 ? cell10(cx / sx) : 0u);
    unsigned yy = spread10(sy > 0.0 ? cell10(cy / sy) : 0u);
    unsigned zz = spread10(sz > 0.0 ? cell10(cz / sz) : 0u);
    code[i] = (xx << 2) + (yy << 1) + zz;
    idx[i] = (unsigned)i;
}

// leaves at m-1 .. 2m-2 in sorted order; internal nodes 0 .. m-2
__global__ void k_leaves(int m, const unsigned* __restrict__ sidx, const unsigned* __restrict__ scode,
                         const Box* __restrict__ box, Node* __restrict__ nodes, Box* __restrict__ nbox,
                         unsigned long long* __restrict__ key, unsigned* __restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const unsigned e = sidx[i];
    nodes[m - 1 + i] = Node{kNone, kNone, kNone, e};
    nbox[m - 1 + i] = box[e];
    key[i] = ((unsigned long long)scode[i] << 32) | e;
    if (i < m - 1) {
        nodes[i] = Node{kNone, kNone, kNone, kNone};
        flags[i] = 0u;
    }
}

__device__ inline int prefix(const unsigned long long* __restrict__ key, int m, int i, int j) {
    if (j < 0 || j >= m) return -1;
    return __clzll(key[i] ^ key[j]);
}

// Karras (2012) radix-tree construction over the unique sorted keys
__global__ void k_internal(int m, const unsigned long long* __restrict__ key, Node* __restrict__ nodes) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m - 1) return;
    const int d = (prefix(key, m, i, i + 1) - prefix(key, m, i, i - 1)) > 0 ? 1 : -1;
    const int dmin = prefix(key, m, i, i - d);
    int lmax = 2;
    while (prefix(key, m, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t > 0; t >>= 1)
        if (prefix(key, m, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = prefix(key, m, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (prefix(key, m, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    const unsigned left = (lo == gamma) ? (unsigned)(gamma + m - 1) : (unsigned)gamma;
    const unsigned right = (hi == gamma + 1) ? (unsigned)(gamma + 1 + m - 1) : (unsigned)(gamma + 1);
    nodes[i].left = left;
    nodes[i].right = right;
    nodes[left].parent = (unsigned)i;
    nodes[right].parent = (unsigned)i;
}

// bottom-up boxes: the second child to arrive at a parent merges it
__global__ void k_internal_boxes(int m, const Node* __restrict__ nodes, Box* __restrict__ nbox,
                                 unsigned* __restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    unsigned p = nodes[m - 1 + i].parent;
    while (p != kNone) {
        __threadfence();
        if (atomicAdd(flags + p, 1u) == 0u) return;  // the sibling's thread finishes this parent
        __threadfence();
        const Node n = nodes[p];
        nbox[p] = merge(nbox[n.left], nbox[n.right]);
        p = n.parent;
    }
}

struct QueryCtx {
    const Node* nodes;
    const Box* box;
    int m;
    double gap;
};

// visit every leaf whose box is near q (root 0; a single leaf when m == 1)
template <typename F>
__device__ inline bool traverse(const QueryCtx& Q, const Box& q, F&& on_leaf) {
    if (Q.m == 1) {
        if (near(q, Q.box[0], Q.gap)) on_leaf(Q.nodes[0].elem);
        return true;
    }
    unsigned stack[kStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const Node n = Q.nodes[stack[--sp]];
        const unsigned ch[2] = {n.left, n.right};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (!near(q, Q.box[ch[c]], Q.gap)) continue;
            const unsigned e = Q.nodes[ch[c]].elem;
            if (e != kNone) {
                on_leaf(e);
            } else {
                if (sp >= kStack) return false;
                stack[sp++] = ch[c];
            }
        }
    }
    return true;
}

__device__ inline bool all_static(const int* __restrict__ bt, unsigned a, unsigned b, unsigned c, unsigned d) {
    return bt && bt[a] >= 2 && bt[b] >= 2 && bt[c] >= 2 && bt[d] >= 2;
}
__device__ inline bool other_body(const int* __restrict__ body, unsigned a, unsigned b) {
    return !body || body[a] == -1 || body[a] != body[b];
}

// pass 0 counts hits per query, pass 1 writes them at the scanned offsets
template <int PASS>
__global__ void k_query_vf(int nq, const unsigned* __restrict__ sv, const unsigned* __restrict__ faces,
                           const double* __restrict__ x, const double* __restrict__ dir, double alpha, QueryCtx Q,
                           const int* __restrict__ bt,
                           const int* __restrict__ body, int* __restrict__ cnt, const int* __restrict__ off,
                           int4* __restrict__ out, int cap, int* __restrict__ err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const unsigned p = sv[i];
    Box q = empty_box();
    grow_swept(q, x, dir, alpha, p);
    int k = 0;
    const int base = PASS ? off[i] : 0;
    bool ok = traverse(Q, q, [&](unsigned f) {
        const unsigned a = faces[3 * (size_t)f], b = faces[3 * (size_t)f + 1], c = faces[3 * (size_t)f + 2];
        if (p == a || p == b || p == c) return;
        if (!other_body(body, p, a) || all_static(bt, p, a, b, c)) return;
        if (PASS && base + k < cap) out[base + k] = make_int4((int)p, (int)a, (int)b, (int)c);
        ++k;
    });
    if (!ok) atomicOr(err, 1);
    if (!PASS) cnt[i] = k;
}

template <int PASS>
__global__ void k_query_ee(int ne, const unsigned* __restrict__ edges, const Box* __restrict__ ebox, QueryCtx Q,
                           const int* __restrict__ bt, const int* __restrict__ body, int* __restrict__ cnt,
                           const int* __restrict__ off, int4* __restrict__ out, int cap, int* __restrict__ err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    const unsigned a0 = edges[2 * (size_t)i], a1 = edges[2 * (size_t)i + 1];
    const Box q = ebox[i];  // edge i's own box (leaf boxes are kept in edge order)
    int k = 0;
    const int base = PASS ? off[i] : 0;
    bool ok = traverse(Q, q, [&](unsigned o) {
        if (o <= (unsigned)i) return;
        const unsigned b0 = edges[2 * (size_t)o], b1 = edges[2 * (size_t)o + 1];
        if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
        if (!other_body(body, a0, b0) || all_static(bt, a0, a1, b0, b1)) return;
        if (PASS && base + k < cap) out[base + k] = make_int4((int)a0, (int)a1, (int)b0, (int)b1);
        ++k;
    });
    if (!ok) atomicOr(err, 1);
    if (!PASS) cnt[i] = k;
}

}  // namespace

// Builds the tree over m primitives of K vertices into t (boxes, nodes).
template <int K>
static void lbvh_build(Ctx& c, LbvhTree& t, int m, const unsigned* prim, const double* x, const double* dir,
                       double alpha) {
    hipStream_t s = c.stream;
    t.m = m;
    if (m <= 0) return;
    Box* leaf = (Box*)t.leaf.reserve((size_t)m * sizeof(Box));
    hipLaunchKernelGGL(k_leaf_boxes<K>, dim3(div_up(m, 256)), dim3(256), 0, s, m, prim, x, dir, alpha, leaf);
    SGIPC_LAUNCH_CHECK();
    const int gb = std::max(1, std::min(256, (int)div_up(m, 256)));
    Box* part = (Box*)t.scene.reserve((size_t)(gb + 1) * sizeof(Box));
    hipLaunchKernelGGL(k_box_reduce, dim3(gb), dim3(256), 0, s, m, (const Box*)leaf, part);
    SGIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_box_reduce, dim3(1), dim3(256), 0, s, gb, (const Box*)part, part + gb);
    SGIPC_LAUNCH_CHECK();
    unsigned* code = t.code.reserve(2 * (size_t)m);
    unsigned* idx = t.idx.reserve(2 * (size_t)m);
    hipLaunchKernelGGL(k_morton, dim3(div_up(m, 256)), dim3(256), 0, s, m, (const Box*)leaf,
                       (const Box*)(part + gb), code, idx);
    SGIPC_LAUNCH_CHECK();
    if (c.sort_impl == SORT_OWN) {
        radix_sort_pairs(c, code, code + m, (const int*)idx, 0, (int*)(idx + m), m, 30, s);
    } else {  // rocPRIM as the cross-check
        size_t bytes = 0;
        SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, code, code + m, idx, idx + m, (unsigned)m, 0, 30, s));
        void* tmp = t.sort_tmp.reserve(bytes + 1);
        SGIPC_HIP(rocprim::radix_sort_pairs(tmp, bytes, code, code + m, idx, idx + m, (unsigned)m, 0, 30, s));
    }
    Node* nodes = (Node*)t.nodes.reserve((size_t)(2 * m - 1) * sizeof(Node));
    Box* nbox = (Box*)t.nbox.reserve((size_t)(2 * m - 1) * sizeof(Box));
    unsigned long long* key = t.key.reserve((size_t)m);
    unsigned* flags = t.flags.reserve((size_t)m);
    hipLaunchKernelGGL(k_leaves, dim3(div_up(m, 256)), dim3(256), 0, s, m, (const unsigned*)(idx + m),
                       (const unsigned*)(code + m), (const Box*)leaf, nodes, nbox, key, flags);
    SGIPC_LAUNCH_CHECK();
    if (m > 1) {
        hipLaunchKernelGGL(k_internal, dim3(div_up(m - 1, 256)), dim3(256), 0, s, m,
                           (const unsigned long long*)key, nodes);
        SGIPC_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_internal_boxes, dim3(div_up(m, 256)), dim3(256), 0, s, m, (const Node*)nodes, nbox,
                           flags);
        SGIPC_LAUNCH_CHECK();
    }
}

void broad_phase(Ctx& c, int nsv, const unsigned* sv, int nf, const unsigned* faces, int ne, const unsigned* edges,
                 const double* x, const int* btype, const int* body, double dhat, int* pt_out, int pt_cap, int& n_pt,
                 int* ee_out, int ee_cap, int& n_ee, const double* dir, double alpha) {
    if (!(dhat > 0.0)) throw arg_error("broad_phase: dhat must be positive");
    hipStream_t s = c.stream;
    const double gap = sqrt(dhat);
    Lbvh& B = c.lbvh;
    int* err = B.err.reserve(1);
    fill<int>(err, 0, 1, s);
    n_pt = n_ee = 0;
    int hcount[2] = {0, 0};
    // point - triangle
    if (nsv > 0 && nf > 0) {
        lbvh_build<3>(c, B.tri, nf, faces, x, dir, alpha);
        QueryCtx Q{(const Node*)B.tri.nodes.p, (const Box*)B.tri.nbox.p, nf, gap};
        int* cnt = B.cnt.reserve((size_t)nsv + 1);
        int* off = B.off.reserve((size_t)nsv + 1);
        hipLaunchKernelGGL(k_query_vf<0>, dim3(div_up(nsv, 128)), dim3(128), 0, s, nsv, sv, faces, x, dir, alpha, Q,
                           btype, body, cnt, (const int*)nullptr, (int4*)nullptr, 0, err);
        SGIPC_LAUNCH_CHECK();
        exclusive_scan(c, cnt, off, nsv, nullptr, s);
        SGIPC_HIP(hipMemcpyAsync(&hcount[0], off + nsv, sizeof(int), hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
        if (hcount[0] <= pt_cap && hcount[0] > 0) {
            hipLaunchKernelGGL(k_query_vf<1>, dim3(div_up(nsv, 128)), dim3(128), 0, s, nsv, sv, faces, x, dir, alpha,
                               Q, btype, body, cnt, (const int*)off, (int4*)pt_out, pt_cap, err);
            SGIPC_LAUNCH_CHECK();
        }
    }
    // edge - edge
    if (ne > 1) {
        lbvh_build<2>(c, B.edge, ne, edges, x, dir, alpha);
        const Box* ebox = (const Box*)B.edge.leaf.p;  // leaf boxes in edge order
        QueryCtx Q{(const Node*)B.edge.nodes.p, (const Box*)B.edge.nbox.p, ne, gap};
        int* cnt = B.cnt.reserve((size_t)ne + 1);
        int* off = B.off.reserve((size_t)ne + 1);
        hipLaunchKernelGGL(k_query_ee<0>, dim3(div_up(ne, 128)), dim3(128), 0, s, ne, edges, ebox, Q, btype, body,
                           cnt, (const int*)nullptr, (int4*)nullptr, 0, err);
        SGIPC_LAUNCH_CHECK();
        exclusive_scan(c, cnt, off, ne, nullptr, s);
        SGIPC_HIP(hipMemcpyAsync(&hcount[1], off + ne, sizeof(int), hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
        if (hcount[1] <= ee_cap && hcount[1] > 0) {
            hipLaunchKernelGGL(k_query_ee<1>, dim3(div_up(ne, 128)), dim3(128), 0, s, ne, edges, ebox, Q, btype,
                               body, cnt, (const int*)off, (int4*)ee_out, ee_cap, err);
            SGIPC_LAUNCH_CHECK();
        }
    }
    int herr = 0;
    SGIPC_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    if (herr) throw state_error("broad_phase: BVH traversal stack exceeded");
    n_pt = hcount[0];
    n_ee = hcount[1];
}

}  // namespace sgipc
