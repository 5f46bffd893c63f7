// The leave-one-out connected set of a (worker, firm) panel, for gfx950 and
// the host.
//
// Reference (hdfe_stream/): _find_components(leave_one_out=True)
// core/connected.py:378-493, _leave_one_out :503-562, and the kernels
// _nb_pairs_degree/_fill, _nb_bfs_components, _nb_articulation_points,
// _nb_pair_rows, _nb_level_rows (core/kernels_graph.py:22-352).
//
// Split by the shape of the work.  The data-parallel part runs on the device:
// level codes of both dimensions, the distinct (worker, firm) pairs with their
// row counts (one radix sort of the packed keys and a run-length pass), and
// the adjacency in both directions as a CSR (a second sort of the directed
// edges, so each vertex's neighbours are in a fixed order).  The pruning loop
// itself -- the largest component, workers with one row, and rounds of cut
// workers found by Tarjan's depth-first search -- is sequential graph search
// over a few million edges and runs in host C++ on the CSR copied back once.
// Every result is a set (components, cut vertices), so nothing depends on
// search order: the output is bit-exact against the reference's rule.
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

// Keep the component with the most rows; ties to the lowest label
// (core/connected.py:537-548).  Returns the number of components before.
static int64_t keep_largest(const Graph &G, std::vector<uint8_t> &alive) {
    std::vector<int64_t> label;
    const int64_t nc = components(G, alive, label);
    if (nc == 0) return 0;
    std::vector<int64_t> size(nc, 0);
    for (size_t e = 0; e < G.keys.size(); ++e) {
        const int64_t a = (int64_t)(G.keys[e] / (uint64_t)G.n1);
        const int64_t b = G.n0 + (int64_t)(G.keys[e] % (uint64_t)G.n1);
        if (label[a] >= 0 && label[b] >= 0) size[label[a]] += G.counts[e];
    }
    int64_t best = 0;
    for (int64_t c = 1; c < nc; ++c)
        if (size[c] > size[best]) best = c;
    for (int64_t v = 0; v < G.n0 + G.n1; ++v) alive[v] = alive[v] && label[v] == best;
    return nc;
}

// Vertices whose removal disconnects the alive subgraph, by an iterative
// Tarjan depth-first search (core/kernels_graph.py:22-86).
static void articulation(const Graph &G, const std::vector<uint8_t> &alive,
                         std::vector<uint8_t> &cut) {
    const int64_t nv = G.n0 + G.n1;
    cut.assign(nv, 0);
    std::vector<int64_t> disc(nv, -1), low(nv, 0), parent(nv, -1), sv(nv), sp(nv);
    int64_t timer = 0;
    for (int64_t start = 0; start < nv; ++start) {
        if (disc[start] != -1 || !alive[start]) continue;
        disc[start] = low[start] = timer++;
        sv[0] = start;
        sp[0] = G.indptr[start];
        int64_t top = 0, root_children = 0;
        while (top >= 0) {
            const int64_t u = sv[top];
            if (sp[top] < G.indptr[u + 1]) {
                const int64_t v = G.indices[sp[top]++];
                if (v == parent[u] || !alive[v]) continue;
                if (disc[v] == -1) {
                    parent[v] = u;
                    disc[v] = low[v] = timer++;
                    ++top;
                    sv[top] = v;
                    sp[top] = G.indptr[v];
                    if (u == start) ++root_children;
                } else if (disc[v] < low[u]) {
                    low[u] = disc[v];
                }
            } else {
                --top;
                if (top >= 0) {
                    const int64_t p = sv[top];
                    if (low[u] < low[p]) low[p] = low[u];
                    if (p != start && low[u] >= disc[p]) cut[p] = 1;
                }
            }
        }
        if (root_children > 1) cut[start] = 1;
    }
}

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
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    Temp tmp;
    DevBuf<int> ca, cb;
    DevBuf<int64_t> la, lb;
    int64_t n0 = 0, n1 = 0;
    int st = factorize_col(s, tmp, a_ids, n, ca, la, n0);
    if (st) return st;
    st = factorize_col(s, tmp, b_ids, n, cb, lb, n1);
    if (st) return st;
    ARG_CHECK(n0 + n1 < INT32_MAX, "hdfe_loo_prune: too many levels");
    // distinct pairs and their rows
    const int kbits = std::max(1, bits_for((uint64_t)n0 * (uint64_t)n1));
    DevBuf<uint64_t> k1, k2;
    DevBuf<int> head, rank;
    DEV_ALLOC(k1, n, s);
    DEV_ALLOC(k2, n, s);
    DEV_ALLOC(head, n, s);
    DEV_ALLOC(rank, n, s);
    hipLaunchKernelGGL(k_pair_keys, dim3(grid1(n)), dim3(256), 0, s, n, ca.p, cb.p, n1, k1.p);
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
    DevBuf<int64_t> indptr;
    DevBuf<int> dst;
    DEV_ALLOC(indptr, nv + 1, s);
    DEV_ALLOC(dst, ne, s);
    hipLaunchKernelGGL(k_csr_ptr, dim3(grid1(nv + 1)), dim3(256), 0, s, nv, ne, e2.p, indptr.p);
    hipLaunchKernelGGL(k_csr_dst, dim3(grid1(ne)), dim3(256), 0, s, ne, e2.p, dst.p);
    HIP_TRY(hipGetLastError());
    // to the host for the search
    Graph G;
    G.n0 = n0;
    G.n1 = n1;
    HIP_TRY(to_host(G.indptr, indptr.p, nv + 1, s));
    HIP_TRY(to_host(G.indices, dst.p, ne, s));
    HIP_TRY(to_host(G.keys, ukey.p, np, s));
    std::vector<int> us;
    HIP_TRY(to_host(us, ustart.p, np, s));
    G.counts.resize(np);
    for (int q = 0; q < np; ++q) G.counts[q] = (q + 1 < np ? us[q + 1] : (int)n) - us[q];
    std::vector<uint8_t> alive(nv, 1);
    const int64_t before = keep_largest(G, alive);
    // workers with one row (a pendant vertex is never a cut vertex, but its
    // leverage is 1)
    std::vector<int64_t> rows0(n0, 0);
    for (int q = 0; q < np; ++q) {
        const int64_t a = (int64_t)(G.keys[q] / (uint64_t)n1);
        const int64_t b = n0 + (int64_t)(G.keys[q] % (uint64_t)n1);
        if (alive[a] && alive[b]) rows0[a] += G.counts[q];
    }
    int64_t single = 0;
    for (int64_t v = 0; v < n0; ++v)
        if (alive[v] && rows0[v] == 1) {
            alive[v] = 0;
            ++single;
        }
    if (single) keep_largest(G, alive);
    int64_t cut_total = 0, rounds = 0;
    std::vector<uint8_t> cut;
    while (true) {
        articulation(G, alive, cut);
        int64_t ncut = 0;
        for (int64_t v = 0; v < n0; ++v)
            if (cut[v] && alive[v]) {
                alive[v] = 0;
                ++ncut;
            }
        if (!ncut) break;
        cut_total += ncut;
        ++rounds;
        keep_largest(G, alive);
    }
    int64_t a0 = 0, a1 = 0;
    for (int64_t v = 0; v < n0; ++v) a0 += alive[v];
    for (int64_t v = n0; v < nv; ++v) a1 += alive[v];
    DevBuf<uint8_t> alive_d;
    DEV_ALLOC(alive_d, nv, s);
    HIP_TRY(hipMemcpyAsync(alive_d.p, alive.data(), nv, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_keep_rows, dim3(grid1(n)), dim3(256), 0, s, n, ca.p, cb.p, n0, alive_d.p,
                       keep);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    int64_t v[8] = {before, single, cut_total, rounds, a0, a1, np, nv};
    for (int i = 0; i < 8; ++i) info[i] = v[i];
    return HDFE_STATUS_OK;
}

}  // extern "C"
