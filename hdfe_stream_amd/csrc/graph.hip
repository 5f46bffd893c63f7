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

// Keep the largest component, by rows (rank_by < 0) or by the levels of
// dimension rank_by; ties to the lowest label (core/connected.py:525-536).
// Returns the number of components before.
static int64_t keep_largest(const Graph &G, std::vector<uint8_t> &alive, int rank_by = -1) {
    std::vector<int64_t> label;
    const int64_t nc = components(G, alive, label);
    if (nc == 0) return 0;
    std::vector<int64_t> size(nc, 0);
    if (rank_by < 0) {
        for (size_t e = 0; e < G.keys.size(); ++e) {
            const int64_t a = (int64_t)(G.keys[e] / (uint64_t)G.n1);
            const int64_t b = G.n0 + (int64_t)(G.keys[e] % (uint64_t)G.n1);
            if (label[a] >= 0 && label[b] >= 0) size[label[a]] += G.counts[e];
        }
    } else {
        const int64_t v0 = rank_by == 0 ? 0 : G.n0, v1 = rank_by == 0 ? G.n0 : G.n0 + G.n1;
        for (int64_t v = v0; v < v1; ++v)
            if (label[v] >= 0) ++size[label[v]];
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

// The leave-one-out connected set (core/connected.py:503-562), in place on
// alive: the largest component, workers (first dimension) with one row
// dropped, then rounds of cut workers.  out = components before, workers
// with one row, cut workers, rounds.
static void loo_rounds(const Graph &G, std::vector<uint8_t> &alive, int rank_by, int64_t out[4]) {
    const int64_t n0 = G.n0, n1 = G.n1;
    out[0] = keep_largest(G, alive, rank_by);
    // workers with one row (a pendant vertex is never a cut vertex, but its
    // leverage is 1)
    std::vector<int64_t> rows0(n0, 0);
    for (size_t q = 0; q < G.keys.size(); ++q) {
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
    if (single) keep_largest(G, alive, rank_by);
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
        keep_largest(G, alive, rank_by);
    }
    out[1] = single;
    out[2] = cut_total;
    out[3] = rounds;
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

// Level codes of both columns (sorted level order), the distinct pairs with
// their rows, and the adjacency of the pairs in both directions, on the host.
struct PairTables {
    Graph G;
    DevBuf<int> ca, cb;       // row codes (device)
    DevBuf<int64_t> la, lb;   // the levels' ids, sorted (device)
};

static int pair_tables(hipStream_t s, Temp &tmp, int64_t n, const int64_t *a_ids,
                       const int64_t *b_ids, PairTables &T, const char *who) {
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
    DevBuf<int64_t> indptr;
    DevBuf<int> dst;
    DEV_ALLOC(indptr, nv + 1, s);
    DEV_ALLOC(dst, ne, s);
    hipLaunchKernelGGL(k_csr_ptr, dim3(grid1(nv + 1)), dim3(256), 0, s, nv, ne, e2.p, indptr.p);
    hipLaunchKernelGGL(k_csr_dst, dim3(grid1(ne)), dim3(256), 0, s, ne, e2.p, dst.p);
    HIP_TRY(hipGetLastError());
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
    TRY(pair_tables(s, tmp, n, a_ids, b_ids, T, "hdfe_loo_prune"));
    const Graph &G = T.G;
    const int64_t n0 = G.n0, nv = G.n0 + G.n1;
    std::vector<uint8_t> alive(nv, 1);
    int64_t lo[4];
    loo_rounds(G, alive, -1, lo);
    int64_t a0 = 0, a1 = 0;
    for (int64_t v = 0; v < n0; ++v) a0 += alive[v];
    for (int64_t v = n0; v < nv; ++v) a1 += alive[v];
    DevBuf<uint8_t> alive_d;
    DEV_ALLOC(alive_d, nv, s);
    HIP_TRY(hipMemcpyAsync(alive_d.p, alive.data(), nv, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_keep_rows, dim3(grid1(n)), dim3(256), 0, s, n, T.ca.p, T.cb.p, n0,
                       alive_d.p, keep);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    int64_t v[8] = {lo[0], lo[1], lo[2], lo[3], a0, a1, (int64_t)G.keys.size(), nv};
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
        TRY(pair_tables(s, tmp, n, a_ids, b_ids, T, "hdfe_components"));
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
            loo_rounds(G, alive, rank_by, lo);
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
