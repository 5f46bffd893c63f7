// The leave-one-out connected set of a (worker, firm) panel, for gfx950 and
// the host.
//
// Reference (hdfe_stream/): _find_components(leave_one_out=True)
// core/connected.py:378-493, _leave_one_out :503-562, and the kernels
// _nb_pairs_degree/_fill, _nb_bfs_components, _nb_articulation_points,
// _nb_pair_rows, _nb_level_rows (core/kernels_graph.py:22-352).
//
// Everything runs on the device: level codes of both dimensions, the distinct
// (worker, firm) pairs with their row counts (one radix sort of the packed
// keys and a run-length pass), the adjacency in both directions as a CSR, and
// the pruning loop itself -- the largest component (lock-free union-find),
// workers with one row, and rounds of cut workers.  Cut vertices come from
// Tarjan and Vishkin's biconnectivity over a breadth-first spanning tree
// (R. E. Tarjan and U. Vishkin, "An efficient parallel biconnectivity
// algorithm", SIAM J. Comput. 14(4), 1985): preorder numbers and subtree
// sizes of the tree, the lowest and highest preorder number reachable from
// each subtree by a non-tree edge, then the connected components of the
// auxiliary graph whose vertices are the tree edges; a vertex is a cut
// vertex when the tree edges at it fall in more than one component.  Every
// result is a set (components, cut vertices), so nothing depends on search or
// scheduling order: the output is bit-exact against the reference's rule,
// which finds the same sets by depth-first search.
#include <algorithm>
#include <memory>
#include <vector>

#include "devbuf.h"

namespace hdfe {

__global__ void k_pair_keys(int64_t n, const int *__restrict__ a, const int *__restrict__ b,
                            int64_t n1, uint64_t *__restrict__ key) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        key[i] = (uint64_t)a[i] * (uint64_t)n1 + (uint64_t)b[i];
}

__global__ void k_run_heads(int64_t n, const uint64_t *__restrict__ key, int *__restrict__ head) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// Distinct keys and their run lengths (the rows of each pair).
__global__ void k_run_unique(int64_t n, const uint64_t *__restrict__ key,
                             const int *__restrict__ head, const int *__restrict__ rank,
                             uint64_t *__restrict__ ukey, int *__restrict__ ustart) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (head[i]) {
            ukey[rank[i] - 1] = key[i];
            ustart[rank[i] - 1] = (int)i;
        }
}

// Directed edges, both ways: (src << 32 | dst), workers 0..n0-1, firms n0...
__global__ void k_edges(int64_t np, const uint64_t *__restrict__ ukey, int64_t n0, int64_t n1,
                        uint64_t *__restrict__ edge) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < np;
         e += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t a = ukey[e] / (uint64_t)n1, b = (uint64_t)n0 + ukey[e] % (uint64_t)n1;
        edge[2 * e] = (a << 32) | b;
        edge[2 * e + 1] = (b << 32) | a;
    }
}

// CSR from the sorted directed edges: indptr by binary search per vertex.
__global__ void k_csr_ptr(int64_t nv, int64_t ne, const uint64_t *__restrict__ edge,
                          int64_t *__restrict__ indptr) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v <= nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = ne;
        const uint64_t target = (uint64_t)v << 32;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (edge[mid] < target) lo = mid + 1;
            else hi = mid;
        }
        indptr[v] = lo;
    }
}

__global__ void k_csr_dst(int64_t ne, const uint64_t *__restrict__ edge, int *__restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne;
         e += (int64_t)gridDim.x * blockDim.x)
        dst[e] = (int)(edge[e] & 0xFFFFFFFFull);
}

// Rows kept: both of the row's levels survive.
__global__ void k_keep_rows(int64_t n, const int *__restrict__ a, const int *__restrict__ b,
                            int64_t n0, const uint8_t *__restrict__ alive,
                            uint8_t *__restrict__ keep) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        keep[i] = alive[a[i]] && alive[n0 + b[i]];
}

// The distinct pairs' endpoints as vertices (workers 0..n0-1, firms n0...)
// and their rows.
__global__ void k_pair_ends(int64_t np, int64_t n, const uint64_t *__restrict__ ukey,
                            const int *__restrict__ ustart, int64_t n0, int64_t n1,
                            int *__restrict__ pa, int *__restrict__ pb, int *__restrict__ cnt) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < np;
         e += (int64_t)gridDim.x * blockDim.x) {
        pa[e] = (int)(ukey[e] / (uint64_t)n1);
        pb[e] = (int)(n0 + (int64_t)(ukey[e] % (uint64_t)n1));
        cnt[e] = (e + 1 < np ? ustart[e + 1] : (int)n) - ustart[e];
    }
}

// ------------------------------------------------- pruning on the device
//
// alive (nv, uint8) marks the vertices still in; dist doubles as the alive
// test inside a search (-2 dead, -1 not reached, else the depth).

// Kernels that sum into one counter run at most this many blocks (grid-stride),
// so that the counter takes few atomics.
constexpr int kSumBlocks = 256;

// One atomic per wave for a per-thread count (every lane calls it once,
// after its loop).
__device__ __forceinline__ void wave_count(int64_t c, unsigned long long *ctr) {
    c = wave_sum(c);
    if ((threadIdx.x & (kWave - 1)) == 0 && c) atomicAdd(ctr, (unsigned long long)c);
}

__global__ void k_vertex_ids(int64_t n, int *__restrict__ x) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (int)i;
}

__global__ void k_fill_int(int64_t n, int val, int *__restrict__ x) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = val;
}

// Union of the two ends of every pair whose ends are both alive.
__global__ void k_cc_pairs(int64_t np, const int *__restrict__ pa, const int *__restrict__ pb,
                           const uint8_t *__restrict__ alive, int *root) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < np;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int a = pa[e], b = pb[e];
        if (alive[a] && alive[b]) uf_join(root, a, b);
    }
}

__global__ void k_cc_flatten(int64_t nv, const uint8_t *__restrict__ alive, int *root) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x)
        if (alive[v]) root[v] = uf_root(root, (int)v);
}

// Per-thread sums keyed by component root: added into out[key] when the key
// changes, and at the end across the wave's runs of equal keys, so that a
// giant component costs one atomic per wave rather than one per element.
struct KeyedSum {
    int key = -1;
    unsigned long long v = 0;
    __device__ void add(int k, unsigned long long x, unsigned long long *out) {
        if (k != key) {
            if (key >= 0 && v) atomicAdd(out + key, v);
            key = k;
            v = 0;
        }
        v += x;
    }
    // every lane of the wave calls it once, after its loop
    __device__ void flush(unsigned long long *out) {
        bool last;
        const unsigned long long t = run_total(key, v, last);
        if (last && key >= 0 && t) atomicAdd(out + key, t);
    }
};

// A component's size: its rows (the pairs' counts) ...
__global__ void k_cc_size_rows(int64_t np, const int *__restrict__ pa, const int *__restrict__ pb,
                               const int *__restrict__ cnt, const uint8_t *__restrict__ alive,
                               const int *__restrict__ root, unsigned long long *size) {
    KeyedSum acc;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < np;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int a = pa[e], b = pb[e];
        if (alive[a] && alive[b]) acc.add(root[a], (unsigned long long)cnt[e], size);
    }
    acc.flush(size);
}

// ... or its levels of one dimension, the vertices [v0, v1).
__global__ void k_cc_size_levels(int64_t v0, int64_t v1, const uint8_t *__restrict__ alive,
                                 const int *__restrict__ root, unsigned long long *size) {
    KeyedSum acc;
    for (int64_t v = v0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < v1;
         v += (int64_t)gridDim.x * blockDim.x)
        if (alive[v]) acc.add(root[v], 1ull, size);
    acc.flush(size);
}

// The largest component, ties to the smallest root (the components are
// labelled in the order of their smallest vertex): the maximum of
// size << 32 | (2^32 - 1 - root) over the roots; and the number of roots.
__global__ void k_cc_best(int64_t nv, const uint8_t *__restrict__ alive,
                          const int *__restrict__ root, const unsigned long long *__restrict__ size,
                          unsigned long long *best, unsigned long long *ncomp) {
    unsigned long long m = 0;
    int64_t c = 0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x)
        if (alive[v] && root[v] == (int)v) {
            const unsigned long long key = (size[v] << 32) | (0xFFFFFFFFull - (unsigned long long)v);
            m = key > m ? key : m;
            ++c;
        }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(m, o, kWave);
        m = t > m ? t : m;
    }
    if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicMax(best, m);
    wave_count(c, ncomp);
}

__global__ void k_cc_keep(int64_t nv, const int *__restrict__ root,
                          const unsigned long long *__restrict__ best, uint8_t *alive) {
    const int br = (int)(0xFFFFFFFFull - (*best & 0xFFFFFFFFull));
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x)
        if (alive[v] && root[v] != br) alive[v] = 0;
}

// Rows of every alive worker over its alive pairs; workers with one row go.
__global__ void k_rows_of_workers(int64_t np, const int *__restrict__ pa,
                                  const int *__restrict__ pb, const int *__restrict__ cnt,
                                  const uint8_t *__restrict__ alive, int *rows0) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < np;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int a = pa[e], b = pb[e];
        if (alive[a] && alive[b]) atomicAdd(rows0 + a, cnt[e]);
    }
}

__global__ void k_drop_marked(int64_t n0, const int *__restrict__ rows0,
                              const uint8_t *__restrict__ cut, uint8_t *alive,
                              unsigned long long *ctr) {
    int64_t c = 0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n0;
         v += (int64_t)gridDim.x * blockDim.x)
        if (alive[v] && (rows0 ? rows0[v] == 1 : cut[v] != 0)) {
            alive[v] = 0;
            ++c;
        }
    wave_count(c, ctr);
}

// Breadth-first search from the root of every component (root[v] == v).
__global__ void k_bfs_init(int64_t nv, const uint8_t *__restrict__ alive,
                           const int *__restrict__ root, int *__restrict__ dist,
                           int *__restrict__ more) {
    bool any = false;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int d = !alive[v] ? -2 : (root[v] == (int)v ? 0 : -1);
        dist[v] = d;
        any |= d == 0;
    }
    if (any) more[0] = 1;
}

// One level: every vertex at depth d marks its unreached neighbours d + 1
// (concurrent writers write the same value) and flags level d + 1.
__global__ void k_bfs_level(int64_t nv, int d, const int64_t *__restrict__ indptr,
                            const int *__restrict__ dst, int *dist, int *__restrict__ more) {
    bool any = false;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        if (dist[v] != d) continue;
        for (int64_t q = indptr[v]; q < indptr[v + 1]; ++q) {
            const int w = dst[q];
            if (dist[w] == -1) {
                dist[w] = d + 1;
                any = true;
            }
        }
    }
    if (any) more[d + 1] = 1;
}

// The tree: each reached vertex's parent is its first neighbour (adjacency
// order) one level up; sort key (depth, parent + 1) over vb bits of parent,
// so that the sorted vertices run level by level, a parent's children
// together, in vertex order within (a stable sort of vertex order).
__global__ void k_tree_parent(int64_t nv, int vb, uint64_t dead,
                              const int64_t *__restrict__ indptr, const int *__restrict__ dst,
                              const int *__restrict__ dist, int *__restrict__ parent,
                              uint64_t *__restrict__ key) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int d = dist[v];
        int p = -1;
        if (d > 0)
            for (int64_t q = indptr[v]; q < indptr[v + 1]; ++q)
                if (dist[dst[q]] == d - 1) {
                    p = dst[q];
                    break;
                }
        parent[v] = p;
        key[v] = d < 0 ? dead : ((uint64_t)d << vb) | (uint64_t)(p + 1);
    }
}

// lstart[d] = first sorted position of depth d; lstart[nlev] = the reached
// vertices' count (preset to nv).
__global__ void k_level_starts(int64_t nv, int vb, int nlev, const uint64_t *__restrict__ skey,
                               int *__restrict__ lstart) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = (int64_t)(skey[i] >> vb);
        const int64_t dp = i ? (int64_t)(skey[i - 1] >> vb) : -1;
        if (d != dp) {
            if (d < nlev) lstart[d] = (int)i;
            else if (dp < nlev) lstart[nlev] = (int)i;
        }
    }
}

// Subtree sizes, one level at a time from the deepest (children done first).
__global__ void k_up_size(int d, const int *__restrict__ lstart, const int *__restrict__ ord,
                          const int *__restrict__ parent, int *nd) {
    const int b = lstart[d], e = lstart[d + 1];
    for (int i = b + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
        const int v = ord[i];
        atomicAdd(nd + parent[v], nd[v]);
    }
}

__global__ void k_sorted_sizes(int64_t nv, int nlev, const int *__restrict__ lstart,
                               const int *__restrict__ ord, const int *__restrict__ nd,
                               int *__restrict__ out) {
    const int na = lstart[nlev];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = i < na ? nd[ord[i]] : 0;
}

// The sorted position of every parent's first child.
__global__ void k_first_child(int nlev, const int *__restrict__ lstart,
                              const uint64_t *__restrict__ skey, const int *__restrict__ ord,
                              const int *__restrict__ parent, int *__restrict__ sfirst) {
    const int na = lstart[nlev];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x)
        if (i == 0 || skey[i] != skey[i - 1]) {
            const int p = parent[ord[i]];
            if (p >= 0) sfirst[p] = i;
        }
}

// Preorder numbers, one level at a time from the roots: a child follows its
// parent and the subtrees of its earlier siblings (scan = exclusive sums of
// the sorted subtree sizes); the trees follow each other.
__global__ void k_down_pre(int d, const int *__restrict__ lstart, const int *__restrict__ ord,
                           const int *__restrict__ parent, const int *__restrict__ scan,
                           const int *__restrict__ sfirst, int *pre) {
    const int b = lstart[d], e = lstart[d + 1];
    for (int i = b + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
        const int v = ord[i], p = parent[v];
        pre[v] = p < 0 ? scan[i] : pre[p] + 1 + scan[i] - scan[sfirst[p]];
    }
}

// A non-tree edge of the search (neither end the other's parent).
__device__ __forceinline__ bool nontree(int v, int w, const int *parent, const int *dist) {
    return dist[w] >= 0 && w != parent[v] && parent[w] != v;
}

// The lowest and highest preorder numbers at v and across its non-tree edges.
__global__ void k_low_high(int64_t nv, const int64_t *__restrict__ indptr,
                           const int *__restrict__ dst, const int *__restrict__ dist,
                           const int *__restrict__ parent, const int *__restrict__ pre,
                           int *__restrict__ lo, int *__restrict__ hi) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        if (dist[v] < 0) continue;
        int l = pre[v], h = pre[v];
        for (int64_t q = indptr[v]; q < indptr[v + 1]; ++q) {
            const int w = dst[q];
            if (!nontree((int)v, w, parent, dist)) continue;
            l = min(l, pre[w]);
            h = max(h, pre[w]);
        }
        lo[v] = l;
        hi[v] = h;
    }
}

// ... then over whole subtrees, one level at a time from the deepest.
__global__ void k_up_low_high(int d, const int *__restrict__ lstart, const int *__restrict__ ord,
                              const int *__restrict__ parent, int *lo, int *hi) {
    const int b = lstart[d], e = lstart[d + 1];
    for (int i = b + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
        const int v = ord[i], p = parent[v];
        atomicMin(lo + p, lo[v]);
        atomicMax(hi + p, hi[v]);
    }
}

// The auxiliary graph (a tree edge is named by its child).  A non-tree edge
// (v, w) with w after v's subtree in preorder joins the tree edges of v and
// w ...
__global__ void k_tv_cross(int64_t nv, const int64_t *__restrict__ indptr,
                           const int *__restrict__ dst, const int *__restrict__ dist,
                           const int *__restrict__ parent, const int *__restrict__ pre,
                           const int *__restrict__ nd, int *aux) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        if (dist[v] < 0) continue;
        const int end = pre[v] + nd[v];
        for (int64_t q = indptr[v]; q < indptr[v + 1]; ++q) {
            const int w = dst[q];
            if (nontree((int)v, w, parent, dist) && pre[w] >= end) uf_join(aux, (int)v, w);
        }
    }
}

// ... and the tree edge of w joins its parent v's own (v not a root) when
// w's subtree reaches outside v's subtree by a non-tree edge.
__global__ void k_tv_tree(int64_t nv, const int *__restrict__ dist,
                          const int *__restrict__ parent, const int *__restrict__ pre,
                          const int *__restrict__ nd, const int *__restrict__ lo,
                          const int *__restrict__ hi, int *aux) {
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < nv;
         w += (int64_t)gridDim.x * blockDim.x) {
        if (dist[w] <= 0) continue;
        const int v = parent[w];
        if (parent[v] < 0) continue;
        if (lo[w] < pre[v] || hi[w] >= pre[v] + nd[v]) uf_join(aux, (int)w, v);
    }
}

// v is a cut vertex when a child's tree edge is in another component than
// v's own (a root: than its first child's).
__global__ void k_tv_cut(int64_t nv, const int *__restrict__ dist, const int *__restrict__ parent,
                         const int *__restrict__ sfirst, const int *__restrict__ ord,
                         const int *__restrict__ aux, uint8_t *__restrict__ cut) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nv;
         c += (int64_t)gridDim.x * blockDim.x) {
        if (dist[c] <= 0) continue;
        const int p = parent[c];
        const int ref = parent[p] >= 0 ? p : ord[sfirst[p]];
        if (aux[c] != aux[ref]) cut[p] = 1;
    }
}

// Alive vertices of [v0, v1), counted.
__global__ void k_count_alive(int64_t v0, int64_t v1, const uint8_t *__restrict__ alive,
                              unsigned long long *ctr) {
    int64_t c = 0;
    for (int64_t v = v0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < v1;
         v += (int64_t)gridDim.x * blockDim.x)
        c += alive[v] != 0;
    wave_count(c, ctr);
}

// ------------------------------------------------------------- host search

struct Graph {
    int64_t n0 = 0, n1 = 0;
    std::vector<int64_t> indptr;
    std::vector<int> indices;
    std::vector<uint64_t> keys;     // distinct pairs a * n1 + b
    std::vector<int64_t> counts;    // their rows
};

// Components of the alive subgraph, labelled in order of their smallest
// vertex (core/kernels_graph.py:89-116); returns the count.
static int64_t components(const Graph &G, const std::vector<uint8_t> &alive,
                          std::vector<int64_t> &label) {
    const int64_t nv = G.n0 + G.n1;
    label.assign(nv, -1);
    std::vector<int64_t> queue(nv);
    int64_t comp = 0;
    for (int64_t s = 0; s < nv; ++s) {
        if (!alive[s] || label[s] != -1) continue;
        label[s] = comp;
        int64_t head = 0, tail = 0;
        queue[tail++] = s;
        while (head < tail) {
            const int64_t u = queue[head++];
            for (int64_t q = G.indptr[u]; q < G.indptr[u + 1]; ++q) {
                const int64_t v = G.indices[q];
                if (alive[v] && label[v] == -1) {
                    label[v] = comp;
                    queue[tail++] = v;
                }
            }
        }
        ++comp;
    }
    return comp;
}

// Rows of the pairs whose two levels are alive.
static int64_t rows_alive(const Graph &G, const std::vector<uint8_t> &alive) {
    int64_t r = 0;
    for (size_t e = 0; e < G.keys.size(); ++e) {
        const int64_t a = (int64_t)(G.keys[e] / (uint64_t)G.n1);
        const int64_t b = G.n0 + (int64_t)(G.keys[e] % (uint64_t)G.n1);
        if (alive[a] && alive[b]) r += G.counts[e];
    }
    return r;
}

// The (k0, k1)-core of the distinct pairs (core/kernels_graph.py:187-221):
// every level with fewer alive partners than its dimension's k goes, all at
// once per round, until a round removes nothing.  Returns the rounds that
// removed something.
static int64_t peel(const Graph &G, std::vector<uint8_t> &alive, int k0, int k1) {
    const int64_t n0 = G.n0, nv = G.n0 + G.n1;
    std::vector<int64_t> deg(nv);
    int64_t rounds = 0;
    while (true) {
        std::fill(deg.begin(), deg.end(), 0);
        for (size_t e = 0; e < G.keys.size(); ++e) {
            const int64_t a = (int64_t)(G.keys[e] / (uint64_t)G.n1);
            const int64_t b = n0 + (int64_t)(G.keys[e] % (uint64_t)G.n1);
            if (alive[a] && alive[b]) {
                ++deg[a];
                ++deg[b];
            }
        }
        bool changed = false;
        for (int64_t v = 0; v < nv; ++v)
            if (alive[v] && deg[v] < (v < n0 ? k0 : k1)) {
                alive[v] = 0;
                changed = true;
            }
        if (!changed) return rounds;
        ++rounds;
    }
}

// The distinct pairs and their adjacency on the device.
struct DevGraph {
    int64_t n0 = 0, n1 = 0, nv = 0, np = 0, ne = 0;
    DevBuf<int64_t> indptr;   // (nv + 1) CSR over the directed edges
    DevBuf<int> dst;          // (ne)
    DevBuf<int> pa, pb, cnt;  // (np) the pairs' ends as vertices, their rows
};

// Level codes of both columns (sorted level order), the distinct pairs with
// their rows, and the adjacency of the pairs in both directions, on the
// device, and with `host` copied to the host as well.
struct PairTables {
    Graph G;
    DevGraph D;
    DevBuf<int> ca, cb;       // row codes (device)
    DevBuf<int64_t> la, lb;   // the levels' ids, sorted (device)
};

static int pair_tables(hipStream_t s, Temp &tmp, int64_t n, const int64_t *a_ids,
                       const int64_t *b_ids, PairTables &T, const char *who, bool host) {
    int64_t n0 = 0, n1 = 0;
    int st = factorize_col(s, tmp, a_ids, n, T.ca, T.la, n0);
    if (st) return st;
    st = factorize_col(s, tmp, b_ids, n, T.cb, T.lb, n1);
    if (st) return st;
    if (n0 + n1 >= INT32_MAX) {
        set_error(std::string(who) + ": too many levels");
        return HDFE_STATUS_ARG;
    }
    // distinct pairs and their rows
    const int kbits = std::max(1, bits_for((uint64_t)n0 * (uint64_t)n1));
    DevBuf<uint64_t> k1, k2;
    DevBuf<int> head, rank;
    DEV_ALLOC(k1, n, s);
    DEV_ALLOC(k2, n, s);
    DEV_ALLOC(head, n, s);
    DEV_ALLOC(rank, n, s);
    hipLaunchKernelGGL(k_pair_keys, dim3(grid1(n)), dim3(256), 0, s, n, T.ca.p, T.cb.p, n1, k1.p);
    {
        size_t bytes = 0;
        HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, k1.p, k2.p, (int)n, 0, kbits, s));
        HIP_TRY(tmp.ensure(bytes, s));
        bytes = (size_t)tmp.buf.n;
        HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp.buf.p, bytes, k1.p, k2.p, (int)n, 0, kbits, s));
    }
    hipLaunchKernelGGL(k_run_heads, dim3(grid1(n)), dim3(256), 0, s, n, k2.p, head.p);
    HIP_TRY(inclusive_sum(tmp, head.p, rank.p, n, s));
    int np = 0;
    HIP_TRY(scalar_to_host(&np, rank.p + n - 1, s));
    DevBuf<uint64_t> ukey;
    DevBuf<int> ustart;
    DEV_ALLOC(ukey, np, s);
    DEV_ALLOC(ustart, np + 1, s);
    hipLaunchKernelGGL(k_run_unique, dim3(grid1(n)), dim3(256), 0, s, n, k2.p, head.p, rank.p,
                       ukey.p, ustart.p);
    // directed edges, sorted: the CSR
    const int64_t ne = 2 * (int64_t)np;
    DevBuf<uint64_t> e1, e2;
    DEV_ALLOC(e1, ne, s);
    DEV_ALLOC(e2, ne, s);
    hipLaunchKernelGGL(k_edges, dim3(grid1(np)), dim3(256), 0, s, (int64_t)np, ukey.p, n0, n1, e1.p);
    {
        const int vbits = std::max(1, bits_for((uint64_t)(n0 + n1)));
        size_t bytes = 0;
        HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, e1.p, e2.p, (int)ne, 0, 32 + vbits, s));
        HIP_TRY(tmp.ensure(bytes, s));
        bytes = (size_t)tmp.buf.n;
        HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp.buf.p, bytes, e1.p, e2.p, (int)ne, 0,
                                                  32 + vbits, s));
    }
    const int64_t nv = n0 + n1;
    DevGraph &D = T.D;
    D.n0 = n0;
    D.n1 = n1;
    D.nv = nv;
    D.np = np;
    D.ne = ne;
    DevBuf<int64_t> &indptr = D.indptr;
    DevBuf<int> &dst = D.dst;
    DEV_ALLOC(indptr, nv + 1, s);
    DEV_ALLOC(dst, ne, s);
    DEV_ALLOC(D.pa, np, s);
    DEV_ALLOC(D.pb, np, s);
    DEV_ALLOC(D.cnt, np, s);
    hipLaunchKernelGGL(k_csr_ptr, dim3(grid1(nv + 1)), dim3(256), 0, s, nv, ne, e2.p, indptr.p);
    hipLaunchKernelGGL(k_csr_dst, dim3(grid1(ne)), dim3(256), 0, s, ne, e2.p, dst.p);
    hipLaunchKernelGGL(k_pair_ends, dim3(grid1(np)), dim3(256), 0, s, (int64_t)np, n, ukey.p,
                       ustart.p, n0, n1, D.pa.p, D.pb.p, D.cnt.p);
    HIP_TRY(hipGetLastError());
    if (!host) return HDFE_STATUS_OK;
    Graph &G = T.G;
    G.n0 = n0;
    G.n1 = n1;
    HIP_TRY(to_host(G.indptr, indptr.p, nv + 1, s));
    HIP_TRY(to_host(G.indices, dst.p, ne, s));
    HIP_TRY(to_host(G.keys, ukey.p, np, s));
    std::vector<int> us;
    HIP_TRY(to_host(us, ustart.p, np, s));
    G.counts.resize(np);
    for (int q = 0; q < np; ++q) G.counts[q] = (q + 1 < np ? us[q + 1] : (int)n) - us[q];
    return HDFE_STATUS_OK;
}

// Work arrays of the pruning loop, over the vertices.
struct LooWork {
    DevBuf<int> root, dist, parent, idx, ord, nd, sizes, scan, sfirst, pre, lo, hi, aux, more,
        lstart, rows0;
    DevBuf<uint64_t> key, skey;
    DevBuf<unsigned long long> size, ctr;   // ctr: see loo_device
    DevBuf<uint8_t> cut;
    hipError_t alloc(int64_t nv, hipStream_t s) {
        DevBuf<int> *ints[] = {&root, &dist, &parent, &idx, &ord, &nd, &sizes, &scan, &sfirst,
                               &pre, &lo, &hi, &aux, &rows0};
        for (DevBuf<int> *b : ints) {
            hipError_t e = b->alloc(nv + 1, s);
            if (e != hipSuccess) return e;
        }
        hipError_t e;
        if ((e = more.alloc(nv + kBfsBatch + 2, s)) != hipSuccess) return e;
        if ((e = lstart.alloc(nv + 2, s)) != hipSuccess) return e;
        if ((e = key.alloc(nv, s)) != hipSuccess) return e;
        if ((e = skey.alloc(nv, s)) != hipSuccess) return e;
        if ((e = size.alloc(nv + 1, s)) != hipSuccess) return e;
        if ((e = ctr.alloc(8, s)) != hipSuccess) return e;
        return cut.alloc(nv, s);
    }
    static constexpr int kBfsBatch = 4;
};

// Keeps the largest alive component, by rows (rank_by < 0) or by the levels
// of dimension rank_by, ties to the one with the smallest vertex
// (core/connected.py:525-536); adds the number of components before to
// *ncomp (device).  Leaves W.root = each alive vertex's component root.
static int keep_largest_dev(hipStream_t s, const DevGraph &D, uint8_t *alive, int rank_by,
                            LooWork &W, unsigned long long *ncomp) {
    const int64_t nv = D.nv, np = D.np;
    unsigned long long *best = W.ctr.p + 7;
    hipLaunchKernelGGL(k_vertex_ids, dim3(grid1(nv)), dim3(256), 0, s, nv, W.root.p);
    hipLaunchKernelGGL(k_cc_pairs, dim3(grid1(np)), dim3(256), 0, s, np, D.pa.p, D.pb.p, alive,
                       W.root.p);
    hipLaunchKernelGGL(k_cc_flatten, dim3(grid1(nv)), dim3(256), 0, s, nv, alive, W.root.p);
    HIP_TRY(hipMemsetAsync(W.size.p, 0, sizeof(unsigned long long) * nv, s));
    HIP_TRY(hipMemsetAsync(best, 0, sizeof(unsigned long long), s));
    if (rank_by < 0)
        hipLaunchKernelGGL(k_cc_size_rows, dim3(grid1(np, 256, kSumBlocks)), dim3(256), 0, s, np,
                           D.pa.p, D.pb.p, D.cnt.p, alive, W.root.p, W.size.p);
    else {
        const int64_t v0 = rank_by == 0 ? 0 : D.n0, v1 = rank_by == 0 ? D.n0 : nv;
        hipLaunchKernelGGL(k_cc_size_levels, dim3(grid1(v1 - v0, 256, kSumBlocks)), dim3(256), 0,
                           s, v0, v1, alive, W.root.p, W.size.p);
    }
    hipLaunchKernelGGL(k_cc_best, dim3(grid1(nv, 256, kSumBlocks)), dim3(256), 0, s, nv, alive,
                       W.root.p, W.size.p, best, ncomp);
    hipLaunchKernelGGL(k_cc_keep, dim3(grid1(nv)), dim3(256), 0, s, nv, W.root.p, best, alive);
    HIP_TRY(hipGetLastError());
    return HDFE_STATUS_OK;
}

// Marks in W.cut the cut vertices of the alive subgraph (every component
// searched from its root in W.root), by Tarjan and Vishkin's rule over a
// breadth-first tree (the header).  Levels are searched kBfsBatch at a time
// between host reads of the level flags.
static int cut_vertices_dev(hipStream_t s, Temp &tmp, const DevGraph &D, const uint8_t *alive,
                            LooWork &W) {
    const int64_t nv = D.nv;
    const int B = LooWork::kBfsBatch;
    HIP_TRY(hipMemsetAsync(W.cut.p, 0, nv, s));
    HIP_TRY(hipMemsetAsync(W.more.p, 0, sizeof(int) * (nv + B + 2), s));
    hipLaunchKernelGGL(k_bfs_init, dim3(grid1(nv)), dim3(256), 0, s, nv, alive, W.root.p,
                       W.dist.p, W.more.p);
    int nlev = 0;
    std::vector<int> h(B + 1);
    for (int d = 0;; d += B) {
        for (int j = 0; j < B; ++j)
            hipLaunchKernelGGL(k_bfs_level, dim3(grid1(nv)), dim3(256), 0, s, nv, d + j,
                               D.indptr.p, D.dst.p, W.dist.p, W.more.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(h.data(), W.more.p + d, sizeof(int) * (B + 1),
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        int c = 0;
        while (c <= B && h[c]) ++c;   // the levels are contiguous
        if (c <= B) {
            nlev = d + c;
            break;
        }
    }
    if (nlev == 0) return HDFE_STATUS_OK;
    const int vb = std::max(1, bits_for((uint64_t)nv)), db = bits_for((uint64_t)nlev);
    const int end_bit = vb + db;
    const uint64_t dead = end_bit >= 64 ? ~0ull : ((1ull << end_bit) - 1);
    hipLaunchKernelGGL(k_tree_parent, dim3(grid1(nv)), dim3(256), 0, s, nv, vb, dead,
                       D.indptr.p, D.dst.p, W.dist.p, W.parent.p, W.key.p);
    hipLaunchKernelGGL(k_vertex_ids, dim3(grid1(nv)), dim3(256), 0, s, nv, W.idx.p);
    HIP_TRY(sort_pairs(tmp, W.key.p, W.skey.p, W.idx.p, W.ord.p, nv, 0, end_bit, s));
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, (int64_t)1, (int)nv, W.lstart.p + nlev);
    hipLaunchKernelGGL(k_level_starts, dim3(grid1(nv)), dim3(256), 0, s, nv, vb, nlev, W.skey.p,
                       W.lstart.p);
    // subtree sizes, one level at a time from the deepest
    hipLaunchKernelGGL(k_fill_int, dim3(grid1(nv)), dim3(256), 0, s, nv, 1, W.nd.p);
    const int gl = grid1(nv);
    for (int d = nlev - 1; d >= 1; --d)
        hipLaunchKernelGGL(k_up_size, dim3(gl), dim3(256), 0, s, d, W.lstart.p, W.ord.p,
                           W.parent.p, W.nd.p);
    // preorder numbers
    hipLaunchKernelGGL(k_sorted_sizes, dim3(grid1(nv)), dim3(256), 0, s, nv, nlev, W.lstart.p,
                       W.ord.p, W.nd.p, W.sizes.p);
    HIP_TRY(exclusive_sum(tmp, W.sizes.p, W.scan.p, nv, s));
    hipLaunchKernelGGL(k_first_child, dim3(grid1(nv)), dim3(256), 0, s, nlev, W.lstart.p,
                       W.skey.p, W.ord.p, W.parent.p, W.sfirst.p);
    for (int d = 0; d < nlev; ++d)
        hipLaunchKernelGGL(k_down_pre, dim3(gl), dim3(256), 0, s, d, W.lstart.p, W.ord.p,
                           W.parent.p, W.scan.p, W.sfirst.p, W.pre.p);
    // low and high over subtrees
    hipLaunchKernelGGL(k_low_high, dim3(grid1(nv)), dim3(256), 0, s, nv, D.indptr.p, D.dst.p,
                       W.dist.p, W.parent.p, W.pre.p, W.lo.p, W.hi.p);
    for (int d = nlev - 1; d >= 1; --d)
        hipLaunchKernelGGL(k_up_low_high, dim3(gl), dim3(256), 0, s, d, W.lstart.p, W.ord.p,
                           W.parent.p, W.lo.p, W.hi.p);
    // the auxiliary graph's components, then the cut vertices
    hipLaunchKernelGGL(k_vertex_ids, dim3(grid1(nv)), dim3(256), 0, s, nv, W.aux.p);
    hipLaunchKernelGGL(k_tv_cross, dim3(grid1(nv)), dim3(256), 0, s, nv, D.indptr.p, D.dst.p,
                       W.dist.p, W.parent.p, W.pre.p, W.nd.p, W.aux.p);
    hipLaunchKernelGGL(k_tv_tree, dim3(grid1(nv)), dim3(256), 0, s, nv, W.dist.p, W.parent.p,
                       W.pre.p, W.nd.p, W.lo.p, W.hi.p, W.aux.p);
    hipLaunchKernelGGL(k_cc_flatten, dim3(grid1(nv)), dim3(256), 0, s, nv, alive, W.aux.p);
    hipLaunchKernelGGL(k_tv_cut, dim3(grid1(nv)), dim3(256), 0, s, nv, W.dist.p, W.parent.p,
                       W.sfirst.p, W.ord.p, W.aux.p, W.cut.p);
    HIP_TRY(hipGetLastError());
    return HDFE_STATUS_OK;
}

// The leave-one-out connected set (core/connected.py:503-562), in place on
// the device mask alive: the largest component, workers (first dimension)
// with one row dropped, then rounds of cut workers, each round followed by
// the largest component again.  out = components before, workers with one
// row, cut workers, rounds.  One host read per round (the cut count) besides
// the search's level flags.
static int loo_device(hipStream_t s, Temp &tmp, const DevGraph &D, uint8_t *alive, int rank_by,
                      int64_t out[4]) {
    LooWork W;
    HIP_TRY(W.alloc(D.nv, s));
    // ctr: 0 components before, 1 workers with one row, 2 cut workers of the
    // round, 3 scratch components count, 7 the largest component's key
    HIP_TRY(hipMemsetAsync(W.ctr.p, 0, sizeof(unsigned long long) * 8, s));
    TRY(keep_largest_dev(s, D, alive, rank_by, W, W.ctr.p + 0));
    HIP_TRY(hipMemsetAsync(W.rows0.p, 0, sizeof(int) * D.n0, s));
    hipLaunchKernelGGL(k_rows_of_workers, dim3(grid1(D.np)), dim3(256), 0, s, D.np, D.pa.p,
                       D.pb.p, D.cnt.p, alive, W.rows0.p);
    hipLaunchKernelGGL(k_drop_marked, dim3(grid1(D.n0, 256, kSumBlocks)), dim3(256), 0, s, D.n0,
                       W.rows0.p, (const uint8_t *)nullptr, alive, W.ctr.p + 1);
    unsigned long long single = 0;
    HIP_TRY(scalar_to_host(&single, W.ctr.p + 1, s));
    // a pendant worker's removal never splits the rest: this only refreshes
    // the roots the search starts from
    if (single) TRY(keep_largest_dev(s, D, alive, rank_by, W, W.ctr.p + 3));
    int64_t cut_total = 0, rounds = 0;
    while (true) {
        TRY(cut_vertices_dev(s, tmp, D, alive, W));
        HIP_TRY(hipMemsetAsync(W.ctr.p + 2, 0, sizeof(unsigned long long), s));
        hipLaunchKernelGGL(k_drop_marked, dim3(grid1(D.n0, 256, kSumBlocks)), dim3(256), 0, s,
                           D.n0, (const int *)nullptr, W.cut.p, alive, W.ctr.p + 2);
        unsigned long long ncut = 0;
        HIP_TRY(scalar_to_host(&ncut, W.ctr.p + 2, s));
        if (!ncut) break;
        cut_total += (int64_t)ncut;
        ++rounds;
        TRY(keep_largest_dev(s, D, alive, rank_by, W, W.ctr.p + 3));
    }
    unsigned long long h[2];
    HIP_TRY(hipMemcpyAsync(h, W.ctr.p, sizeof(h), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    out[0] = (int64_t)h[0];
    out[1] = (int64_t)h[1];
    out[2] = cut_total;
    out[3] = rounds;
    return HDFE_STATUS_OK;
}

// What hdfe_components returns: per level of each dimension its id and
// component (-1: removed), the components' sizes, and the counts.
struct Components {
    std::vector<int64_t> levels[2], comp[2];
    std::vector<int64_t> sizes;  // (nc, 3): rows, levels of each dimension
    int64_t info[16] = {0};
};

}  // namespace hdfe

using namespace hdfe;

extern "C" {

// The leave-one-out connected set of the rows (a[i], b[i]) (device int64 id
// columns, n rows): keep (n, device uint8) marks the rows whose worker and
// firm survive; info (host, 8) = components before, workers with one row,
// cut workers, rounds, surviving workers, surviving firms, distinct pairs,
// workers + firms before.  The first column plays the workers.
int hdfe_loo_prune(int64_t n, const int64_t *a_ids, const int64_t *b_ids, uint8_t *keep,
                   int64_t *info) {
    ARG_CHECK(a_ids && b_ids && keep && info && n > 0 && n < INT32_MAX, "hdfe_loo_prune: bad argument");
    hipStream_t s = nullptr;
    HIP_TRY(stream_get(&s));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { stream_put(s); }
    } guard{s};
    Temp tmp;
    PairTables T;
    TRY(pair_tables(s, tmp, n, a_ids, b_ids, T, "hdfe_loo_prune", false));
    const DevGraph &D = T.D;
    DevBuf<uint8_t> alive;
    DEV_ALLOC(alive, D.nv, s);
    HIP_TRY(hipMemsetAsync(alive.p, 1, D.nv, s));
    int64_t lo[4];
    TRY(loo_device(s, tmp, D, alive.p, -1, lo));
    DevBuf<unsigned long long> cnt;
    DEV_ALLOC(cnt, 2, s);
    HIP_TRY(hipMemsetAsync(cnt.p, 0, 2 * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_count_alive, dim3(grid1(D.n0, 256, kSumBlocks)), dim3(256), 0, s,
                       (int64_t)0, D.n0, alive.p, cnt.p);
    hipLaunchKernelGGL(k_count_alive, dim3(grid1(D.n1, 256, kSumBlocks)), dim3(256), 0, s, D.n0,
                       D.nv, alive.p, cnt.p + 1);
    hipLaunchKernelGGL(k_keep_rows, dim3(grid1(n)), dim3(256), 0, s, n, T.ca.p, T.cb.p, D.n0,
                       alive.p, keep);
    HIP_TRY(hipGetLastError());
    unsigned long long a01[2];
    HIP_TRY(hipMemcpyAsync(a01, cnt.p, sizeof(a01), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    int64_t v[8] = {lo[0], lo[1], lo[2], lo[3], (int64_t)a01[0], (int64_t)a01[1], D.np, D.nv};
    for (int i = 0; i < 8; ++i) info[i] = v[i];
    return HDFE_STATUS_OK;
}

// The connected components of the graph whose vertices are the levels of two
// id columns (device int64, n rows) and whose edges are the rows
// (connected_components, core/connected.py:211-338; _find_components
// :378-493).  min_deg0/min_deg1 >= 1: the (k0, k1)-core first; leave_one_out:
// then the leave-one-out connected set, one component; rank_by: -1 rows,
// 0 or 1 the levels of that dimension.  Components are labelled by their
// smallest vertex (a first-dimension level), then ranked by size, stably.
// *out: a handle for hdfe_components_info/levels/sizes, freed by
// hdfe_components_destroy.
int hdfe_components(int64_t n, const int64_t *a_ids, const int64_t *b_ids, int min_deg0,
                    int min_deg1, int leave_one_out, int rank_by, void **out) {
    ARG_CHECK(a_ids && b_ids && out && n > 0 && n < INT32_MAX, "hdfe_components: bad argument");
    ARG_CHECK(min_deg0 >= 1 && min_deg1 >= 1, "hdfe_components: min_degree must be >= 1");
    ARG_CHECK(rank_by >= -1 && rank_by <= 1, "hdfe_components: rank_by must be -1, 0 or 1");
    hipStream_t s = nullptr;
    HIP_TRY(stream_get(&s));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { stream_put(s); }
    } guard{s};
    std::unique_ptr<Components> R(new Components());
    {
        Temp tmp;
        PairTables T;
        TRY(pair_tables(s, tmp, n, a_ids, b_ids, T, "hdfe_components", true));
        const Graph &G = T.G;
        const int64_t n0 = G.n0, n1 = G.n1, nv = n0 + n1;
        int64_t *info = R->info;
        info[0] = n0;
        info[1] = n1;
        info[3] = n;
        info[4] = (int64_t)G.keys.size();
        std::vector<uint8_t> alive(nv, 1);
        if (min_deg0 > 1 || min_deg1 > 1) {
            info[5] = peel(G, alive, min_deg0, min_deg1);
            int64_t a0 = 0, a1 = 0;
            for (int64_t v = 0; v < n0; ++v) a0 += alive[v];
            for (int64_t v = n0; v < nv; ++v) a1 += alive[v];
            info[6] = n0 - a0;
            info[7] = n1 - a1;
            info[8] = n - rows_alive(G, alive);
        }
        std::vector<int64_t> label;
        int64_t nc = 0;
        if (leave_one_out) {
            const std::vector<uint8_t> before = alive;
            int64_t lo[4];
            DevBuf<uint8_t> alive_d;
            DEV_ALLOC(alive_d, nv, s);
            HIP_TRY(hipMemcpyAsync(alive_d.p, alive.data(), nv, hipMemcpyHostToDevice, s));
            TRY(loo_device(s, tmp, T.D, alive_d.p, rank_by, lo));
            HIP_TRY(hipMemcpyAsync(alive.data(), alive_d.p, nv, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            for (int i = 0; i < 4; ++i) info[9 + i] = lo[i];
            int64_t d0 = 0, d1 = 0;
            for (int64_t v = 0; v < n0; ++v) d0 += before[v] && !alive[v];
            for (int64_t v = n0; v < nv; ++v) d1 += before[v] && !alive[v];
            info[13] = d0;
            info[14] = d1;
            info[15] = rows_alive(G, before) - rows_alive(G, alive);
            label.assign(nv, -1);
            bool any = false;
            for (int64_t v = 0; v < nv; ++v)
                if (alive[v]) {
                    label[v] = 0;
                    any = true;
                }
            nc = any ? 1 : 0;
        } else {
            nc = components(G, alive, label);
        }
        // sizes, then the ranking: stable, largest first
        std::vector<int64_t> rows(nc, 0), cnt0(nc, 0), cnt1(nc, 0);
        for (size_t e = 0; e < G.keys.size(); ++e) {
            const int64_t a = (int64_t)(G.keys[e] / (uint64_t)n1);
            const int64_t b = n0 + (int64_t)(G.keys[e] % (uint64_t)n1);
            if (label[a] >= 0 && label[b] >= 0) rows[label[a]] += G.counts[e];
        }
        for (int64_t v = 0; v < n0; ++v)
            if (label[v] >= 0) ++cnt0[label[v]];
        for (int64_t v = n0; v < nv; ++v)
            if (label[v] >= 0) ++cnt1[label[v]];
        const std::vector<int64_t> &key = rank_by < 0 ? rows : rank_by == 0 ? cnt0 : cnt1;
        std::vector<int64_t> order(nc);
        for (int64_t c = 0; c < nc; ++c) order[c] = c;
        std::stable_sort(order.begin(), order.end(),
                         [&](int64_t x, int64_t y) { return key[x] > key[y]; });
        std::vector<int64_t> rank(nc);
        for (int64_t r = 0; r < nc; ++r) rank[order[r]] = r;
        R->sizes.resize(3 * nc);
        for (int64_t r = 0; r < nc; ++r) {
            R->sizes[3 * r] = rows[order[r]];
            R->sizes[3 * r + 1] = cnt0[order[r]];
            R->sizes[3 * r + 2] = cnt1[order[r]];
        }
        info[2] = nc;
        for (int d = 0; d < 2; ++d) {
            const int64_t off = d == 0 ? 0 : n0, nd = d == 0 ? n0 : n1;
            R->comp[d].resize(nd);
            for (int64_t v = 0; v < nd; ++v)
                R->comp[d][v] = label[off + v] >= 0 ? rank[label[off + v]] : -1;
            HIP_TRY(to_host(R->levels[d], (d == 0 ? T.la : T.lb).p, nd, s));
        }
    }
    HIP_TRY(hipStreamSynchronize(s));
    *out = R.release();
    return HDFE_STATUS_OK;
}

// info (host, 16): levels of each dimension, components, rows, distinct
// pairs; min_degree's rounds, removed levels of each dimension, removed
// rows; leave-one-out's components before, workers with one row, cut
// workers, rounds, removed levels of each dimension, removed rows.
int hdfe_components_info(void *h, int64_t *info) {
    ARG_CHECK(h && info, "hdfe_components_info: null argument");
    const Components &R = *(const Components *)h;
    for (int i = 0; i < 16; ++i) info[i] = R.info[i];
    return HDFE_STATUS_OK;
}

// Dimension dim's level ids (sorted) and their components (-1: removed),
// host arrays of its number of levels.
int hdfe_components_levels(void *h, int dim, int64_t *levels, int64_t *comp) {
    ARG_CHECK(h && levels && comp && (dim == 0 || dim == 1), "hdfe_components_levels: bad argument");
    const Components &R = *(const Components *)h;
    std::copy(R.levels[dim].begin(), R.levels[dim].end(), levels);
    std::copy(R.comp[dim].begin(), R.comp[dim].end(), comp);
    return HDFE_STATUS_OK;
}

// The components' sizes, largest first: (components, 3) host, rows and the
// levels of each dimension.
int hdfe_components_sizes(void *h, int64_t *sizes) {
    ARG_CHECK(h && sizes, "hdfe_components_sizes: null argument");
    const Components &R = *(const Components *)h;
    std::copy(R.sizes.begin(), R.sizes.end(), sizes);
    return HDFE_STATUS_OK;
}

void hdfe_components_destroy(void *h) { delete (Components *)h; }

}  // extern "C"
