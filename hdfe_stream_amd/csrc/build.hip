// Passes 0 and 1 and the reduced system on the device, for gfx950.
//
// From the rows of a panel resident in HBM (fixed-effect id columns, the
// variables, optional weights) to the explicit S = D_o' M_0 D_o, its Jacobi
// diagonal and the right-hand side D_o' M_0 V, with everything kept on the
// device for the solve, the normalisation, the assembly of V' M_D V and the
// row pass that recovers the streamed effects.
//
// Reference steps restated (hdfe_stream/):
//   factorize: codes in sorted level order        core/passes.py:329-335
//   singletons, dimensions in turn                core/passes.py:859-904
//   streamed dimension: highest cardinality       core/estimator.py:371-385
//   rows sorted by (group, codes[, variables])    core/passes.py:340-394
//   cells, centring, identifying groups, obs/cnt  core/passes.py:397-589
//   diag / rhs                                    core/kernels_base.py:64-110
//   S                                             core/kernels_base.py:192-358
//   components and their ranking                  core/kernels_base.py:155-174,
//                                                 core/passes.py:933-979
//   normalisation                                 core/solve.py:383-399
//   V' M_D V from cells                           core/kernels_base.py:510-540
//   streamed effects and residuals                core/kernels_base.py:543-619
//
// Design (not the reference's numba loops): every step is a sort, a scan, a
// segmented reduction or a gather over struct-of-arrays columns, so the
// sums are formed in a fixed order and results do not depend on scheduling.
// The entries of S between two "sparse" dimensions (many levels) come from a
// sort-and-reduce of per-group level pairs; the entries that involve a
// "dense" dimension (few levels, e.g. years, whose rows touch every firm) are
// accumulated row by row into a dense panel, so that no level's row becomes a
// sort bucket of millions of entries.
#include <algorithm>
#include <numeric>

#include "devbuf.h"
#include "solver.h"

namespace hdfe {

constexpr int kDenseMax = 64;  // a dimension with at most this many levels is dense
constexpr int T256 = 256;

// ------------------------------------------------------------ kernels

__global__ void k_key_from_col(int64_t n, const int64_t *__restrict__ col, int64_t lo,
                               uint64_t *__restrict__ key, int *__restrict__ idx) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        key[i] = (uint64_t)(col[i] - lo);
        idx[i] = (int)i;
    }
}

__global__ void k_iota(int64_t n, int *__restrict__ idx) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        idx[i] = (int)i;
}

__global__ void k_heads_u64(int64_t n, const uint64_t *__restrict__ key, int *__restrict__ head) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// code[idx[i]] = rank[i] - 1; levels[rank - 1] = value at each head.
__global__ void k_scatter_codes(int64_t n, const uint64_t *__restrict__ key, int64_t lo,
                                const int *__restrict__ idx, const int *__restrict__ head,
                                const int *__restrict__ rank, int *__restrict__ code,
                                int64_t *__restrict__ levels) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int r = rank[i] - 1;
        code[idx[i]] = r;
        if (head[i] && levels) levels[r] = (int64_t)key[i] + lo;
    }
}

__global__ void k_minmax_partial(int64_t n, const int64_t *__restrict__ col,
                                 int64_t *__restrict__ part) {
    __shared__ int64_t lo[T256], hi[T256];
    int64_t a = INT64_MAX, b = INT64_MIN;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i + 3 * step < n; i += 4 * step) {  // four loads in flight
        const int64_t v0 = col[i], v1 = col[i + step], v2 = col[i + 2 * step], v3 = col[i + 3 * step];
        a = min(a, min(min(v0, v1), min(v2, v3)));
        b = max(b, max(max(v0, v1), max(v2, v3)));
    }
    for (; i < n; i += step) {
        a = min(a, col[i]);
        b = max(b, col[i]);
    }
    lo[threadIdx.x] = a;
    hi[threadIdx.x] = b;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            lo[threadIdx.x] = min(lo[threadIdx.x], lo[threadIdx.x + o]);
            hi[threadIdx.x] = max(hi[threadIdx.x], hi[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = lo[0];
        part[2 * blockIdx.x + 1] = hi[0];
    }
}

constexpr int kLdsBins = 8192;  // dimensions up to this many levels count in LDS

// Kernels over runs (a group's cells or rows) give each run W adjacent lanes,
// W near the mean run length: a wave then reads the consecutive runs of 64/W
// groups with coalesced loads, where a thread per run would stride.  Sums
// over a run's lanes use a fixed butterfly (deterministic).
template <int W>
__device__ __forceinline__ double run_sum(double x) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, W);
    return x;
}

static int run_width(int64_t items, int64_t runs) {
    const int64_t avg = runs > 0 ? (items + runs - 1) / runs : 1;
    return avg <= 4 ? 4 : avg <= 8 ? 8 : avg <= 16 ? 16 : 32;
}

#define HDFE_BY_WIDTH(w, LAUNCH) \
    switch (w) {                  \
    case 4: LAUNCH(4); break;     \
    case 8: LAUNCH(8); break;     \
    case 16: LAUNCH(16); break;   \
    default: LAUNCH(32); break;   \
    }

// Singletons: rows per level among the rows still kept.  A dimension with
// few levels (years: every row hits one of a handful of counters) counts in
// per-workgroup LDS bins first, then adds them to the global counters.
__global__ void k_level_count(int64_t n, const int *__restrict__ code,
                              const uint8_t *__restrict__ keep, int *__restrict__ cnt, int nl) {
    __shared__ int bins[kLdsBins];
    const bool lds = nl <= kLdsBins;
    if (lds) {
        for (int t = threadIdx.x; t < nl; t += blockDim.x) bins[t] = 0;
        __syncthreads();
    }
    // wave-uniform trip count: one atomic per run of equal codes
    const int lane = threadIdx.x & (kWave - 1);
    for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x - lane); i0 < n;
         i0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = i0 + lane;
        const bool in = i < n;
        const int key = in ? code[i] : -1;
        bool last;
        const int tot = run_total<int>(key, (in && keep[i]) ? 1 : 0, last);
        if (last && key >= 0 && tot) {
            if (lds) atomicAdd(&bins[key], tot);
            else atomicAdd(&cnt[key], tot);
        }
    }
    if (lds) {
        __syncthreads();
        for (int t = threadIdx.x; t < nl; t += blockDim.x)
            if (bins[t]) atomicAdd(&cnt[t], bins[t]);
    }
}

__global__ void k_drop_single(int64_t n, const int *__restrict__ code,
                              const int *__restrict__ cnt, uint8_t *__restrict__ keep,
                              unsigned long long *__restrict__ dropped) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (keep[i] && cnt[code[i]] == 1) {
            keep[i] = 0;
            atomicAdd(dropped, 1ull);
        }
}

// Levels still present after the restriction, for re-coding.
__global__ void k_level_present(int64_t n, const int *__restrict__ code,
                                const uint8_t *__restrict__ keep, int *__restrict__ present) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (keep[i]) present[code[i]] = 1;
}

__global__ void k_recode(int64_t n, int *__restrict__ code, const int *__restrict__ newcode) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        code[i] = newcode[code[i]];
}

__global__ void k_compact_levels(int64_t nl, const int *__restrict__ present,
                                 const int *__restrict__ newcode, const int64_t *__restrict__ lv,
                                 int64_t *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nl;
         i += (int64_t)gridDim.x * blockDim.x)
        if (present[i]) out[newcode[i]] = lv[i];
}

// Row sort keys: (g, code_1, ..., code_D) packed high to low; dropped rows
// sort last.
__global__ void k_pack_key(int64_t n, int D, const int *__restrict__ gcode,
                           const int *__restrict__ codes, int64_t ld, const int *__restrict__ shift,
                           int gshift, const uint8_t *__restrict__ keep,
                           uint64_t *__restrict__ key, int *__restrict__ idx) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t k = (uint64_t)gcode[i] << gshift;
        for (int d = 0; d < D; ++d) k |= (uint64_t)codes[d * ld + i] << shift[d];
        key[i] = keep[i] ? k : ~0ull;
        idx[i] = (int)i;
    }
}

// Order-preserving unsigned image of a double (ascending, -0 before +0).
__device__ __forceinline__ uint64_t f64_key(double v) {
    uint64_t u = __double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__global__ void k_var_key(int64_t n, const double *__restrict__ V, int m, int j,
                          const int *__restrict__ perm, uint64_t *__restrict__ key,
                          int *__restrict__ idx) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int r = perm ? perm[i] : (int)i;
        key[i] = f64_key(V[(int64_t)r * m + j]);
        idx[i] = r;
    }
}

// Sorted key of the rows in the current order perm: key[i] = full[perm[i]].
__global__ void k_gather_key(int64_t n, const uint64_t *__restrict__ full,
                             const int *__restrict__ perm, uint64_t *__restrict__ key) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        key[i] = full[perm[i]];
}

template <typename T>
__global__ void k_gather(int64_t n, const int *__restrict__ perm, const T *__restrict__ src,
                         T *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[perm[i]];
}

// V (n, m) row major gathered from m column pointers by the row order perm.
__global__ void k_gather_vars(int64_t n, int m, const double *const *__restrict__ cols,
                              const int *__restrict__ perm, double *__restrict__ V) {
    const int64_t nm = n * m;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nm;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / m;
        const int j = (int)(e - i * m);
        V[e] = cols[j][perm[i]];
    }
}

__global__ void k_gather_weights(int64_t n, const void *__restrict__ w, int f32,
                                 const int *__restrict__ perm, double *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int r = perm[i];
        out[i] = f32 ? (double)((const float *)w)[r] : ((const double *)w)[r];
    }
}

// Run heads over sorted rows: a new group (gh) or a new cell (ch).
__global__ void k_row_heads(int64_t n, int D, const int *__restrict__ g,
                            const int *__restrict__ codes, int64_t ld, int *__restrict__ gh,
                            int *__restrict__ ch) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int ng = (i == 0) || g[i] != g[i - 1];
        int nc = ng;
        for (int d = 0; d < D && !nc; ++d) nc = codes[d * ld + i] != codes[d * ld + i - 1];
        gh[i] = ng;
        ch[i] = nc;
    }
}

// starts[rank[i] - 1] = i at heads; starts[count] = n.
__global__ void k_starts(int64_t n, const int *__restrict__ head, const int *__restrict__ rank,
                         int *__restrict__ starts) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (head[i]) starts[rank[i] - 1] = (int)i;
    if (blockIdx.x == 0 && threadIdx.x == 0) starts[rank[n - 1]] = (int)n;
}

// Column sums of V (n, m) in fixed row blocks: part[b * m + j].
// MM: m at compile time (1..8; 0 runtime): the row's loads unguarded and
// four rows' loads in flight (each column still summed row by row in order)
template <int MM>
__global__ void k_colsum_rows(int64_t n, int m, const double *__restrict__ V,
                              double *__restrict__ part) {
    __shared__ double lds[T256];
    const int64_t r0 = part_begin(n, blockIdx.x, gridDim.x), r1 = part_begin(n, blockIdx.x + 1, gridDim.x);
    if (MM > 0) {
        constexpr int M1 = MM > 0 ? MM : 1;
        double sj[M1];
#pragma unroll
        for (int j = 0; j < M1; ++j) sj[j] = 0.0;
#pragma unroll 4
        for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x)
#pragma unroll
            for (int j = 0; j < M1; ++j) sj[j] += V[i * M1 + j];
#pragma unroll
        for (int j = 0; j < M1; ++j) {
            const double t = block_sum<double>(sj[j], lds);
            if (threadIdx.x == 0) part[(int64_t)blockIdx.x * M1 + j] = t;
        }
        return;
    }
    if (m <= 8) {
        // every column in one pass over the rows (each column's sums in the
        // same order as a pass per column)
        double sj[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < m) sj[j] += V[i * m + j];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (j >= m) break;
            const double t = block_sum<double>(sj[j], lds);
            if (threadIdx.x == 0) part[(int64_t)blockIdx.x * m + j] = t;
        }
        return;
    }
    for (int j = 0; j < m; ++j) {
        double s = 0.0;
        for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) s += V[i * m + j];
        s = block_sum<double>(s, lds);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * m + j] = s;
    }
}

// Squared deviations from the mean, same blocks.
template <int MM>
__global__ void k_colss_rows(int64_t n, int m, const double *__restrict__ V,
                             const double *__restrict__ mean, double *__restrict__ part) {
    __shared__ double lds[T256];
    const int64_t r0 = part_begin(n, blockIdx.x, gridDim.x), r1 = part_begin(n, blockIdx.x + 1, gridDim.x);
    if (MM > 0) {
        constexpr int M1 = MM > 0 ? MM : 1;
        double mu[M1], sj[M1];
#pragma unroll
        for (int j = 0; j < M1; ++j) {
            mu[j] = mean[j];
            sj[j] = 0.0;
        }
#pragma unroll 4
        for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x)
#pragma unroll
            for (int j = 0; j < M1; ++j) {
                const double d = V[i * M1 + j] - mu[j];
                sj[j] += d * d;
            }
#pragma unroll
        for (int j = 0; j < M1; ++j) {
            const double t = block_sum<double>(sj[j], lds);
            if (threadIdx.x == 0) part[(int64_t)blockIdx.x * M1 + j] = t;
        }
        return;
    }
    if (m <= 8) {
        double mu[8], sj[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            mu[j] = j < m ? mean[j] : 0.0;
            sj[j] = 0.0;
        }
        for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < m) {
                    const double d = V[i * m + j] - mu[j];
                    sj[j] += d * d;
                }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (j >= m) break;
            const double t = block_sum<double>(sj[j], lds);
            if (threadIdx.x == 0) part[(int64_t)blockIdx.x * m + j] = t;
        }
        return;
    }
    for (int j = 0; j < m; ++j) {
        double s = 0.0;
        for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
            const double d = V[i * m + j] - mean[j];
            s += d * d;
        }
        s = block_sum<double>(s, lds);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * m + j] = s;
    }
}

// Cells: weight n_c (or rows), the weighted sums of the centred variables,
// and per-block sums of the within-cell scatter (m <= 8).
__global__ void k_cells(int64_t C, int m, const int *__restrict__ cstart,
                        const double *__restrict__ V, const double *__restrict__ w,
                        const double *__restrict__ mean, double *__restrict__ cn,
                        double *__restrict__ crun, double *__restrict__ csums,
                        double *__restrict__ wpart) {
    __shared__ double lds[T256];
    double W[64];
    const int mm = m <= 8 ? m * m : 0;
    for (int t = 0; t < 64; ++t) W[t] = 0.0;
    const int64_t c0 = part_begin(C, blockIdx.x, gridDim.x), c1 = part_begin(C, blockIdx.x + 1, gridDim.x);
    for (int64_t c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
        const int s = cstart[c], e = cstart[c + 1];
        double nn = 0.0;
        double sm[8];
        for (int j = 0; j < 8; ++j) sm[j] = 0.0;
        for (int i = s; i < e; ++i) {
            const double wi = w ? w[i] : 1.0;
            nn += wi;
            if (m <= 8)
                for (int j = 0; j < m; ++j) sm[j] += wi * (V[(int64_t)i * m + j] - mean[j]);
        }
        cn[c] = nn;
        crun[c] = (double)(e - s);
        if (m <= 8) {
            for (int j = 0; j < m; ++j) csums[c * m + j] = sm[j];
            if (mm) {
                for (int i = s; i < e; ++i) {
                    const double wi = w ? w[i] : 1.0;
                    double dv[8];
                    for (int j = 0; j < m; ++j) dv[j] = V[(int64_t)i * m + j] - mean[j] - sm[j] / nn;
                    for (int a = 0; a < m; ++a)
                        for (int b = 0; b < m; ++b) W[a * m + b] += wi * dv[a] * dv[b];
                }
            }
        } else {
            for (int j = 0; j < m; ++j) {
                double t = 0.0;
                for (int i = s; i < e; ++i)
                    t += (w ? w[i] : 1.0) * (V[(int64_t)i * m + j] - mean[j]);
                csums[c * m + j] = t;
            }
        }
    }
    for (int t = 0; t < mm; ++t) {
        const double v = block_sum<double>(W[t], lds);
        if (threadIdx.x == 0) wpart[(int64_t)blockIdx.x * mm + t] = v;
    }
}

// k_cells for a compile-time m <= 8: the sums and the within-cell scatter
// stay in registers (the runtime-m form indexes a private array, which the
// compiler places in scratch memory).  The scatter is accumulated on its
// upper triangle, in the runtime-m form's order, and mirrored, so it is
// exactly symmetric (the runtime-m form rounds its two triangles apart).
// ONE: unweighted cells of one row each (cell c is row c): n = 1, one row,
// the sums the centred row itself and no scatter (every deviation is exactly
// zero), four cells' loads in flight a lane.
template <int M, bool ONE>
__global__ __launch_bounds__(256) void k_cells_m(int64_t C, const int *__restrict__ cstart,
                                                 const double *__restrict__ V,
                                                 const double *__restrict__ w,
                                                 const double *__restrict__ mean,
                                                 double *__restrict__ cn, double *__restrict__ crun,
                                                 double *__restrict__ csums,
                                                 double *__restrict__ wpart) {
    __shared__ double lds[T256];
    constexpr int NT = M * (M + 1) / 2;
    double W[NT];
    double mu[M];
#pragma unroll
    for (int t = 0; t < NT; ++t) W[t] = 0.0;
#pragma unroll
    for (int j = 0; j < M; ++j) mu[j] = mean[j];
    const int64_t c0 = part_begin(C, blockIdx.x, gridDim.x), c1 = part_begin(C, blockIdx.x + 1, gridDim.x);
    if (ONE) {
        constexpr int B = 4;
        for (int64_t c = c0 + threadIdx.x; c < c1; c += B * (int64_t)blockDim.x) {
            double v[B][M];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int64_t cb = min(c + b * (int64_t)blockDim.x, c1 - 1);
#pragma unroll
                for (int j = 0; j < M; ++j) v[b][j] = V[cb * M + j];
            }
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int64_t cb = c + b * (int64_t)blockDim.x;
                if (cb < c1) {
                    cn[cb] = 1.0;
                    crun[cb] = 1.0;
#pragma unroll
                    for (int j = 0; j < M; ++j) csums[cb * M + j] = 0.0 + 1.0 * (v[b][j] - mu[j]);
                }
            }
        }
    }
    for (int64_t c = c0 + threadIdx.x; !ONE && c < c1; c += blockDim.x) {
        const int s = cstart[c], e = cstart[c + 1];
        double nn = 0.0;
        double sm[M];
#pragma unroll
        for (int j = 0; j < M; ++j) sm[j] = 0.0;
        for (int i = s; i < e; ++i) {
            const double wi = w ? w[i] : 1.0;
            nn += wi;
#pragma unroll
            for (int j = 0; j < M; ++j) sm[j] += wi * (V[(int64_t)i * M + j] - mu[j]);
        }
        cn[c] = nn;
        crun[c] = (double)(e - s);
#pragma unroll
        for (int j = 0; j < M; ++j) csums[c * M + j] = sm[j];
        // an unweighted one-row cell: every deviation is exactly zero
        if (!w && e - s == 1) continue;
        for (int i = s; i < e; ++i) {
            const double wi = w ? w[i] : 1.0;
            double dv[M];
#pragma unroll
            for (int j = 0; j < M; ++j) dv[j] = V[(int64_t)i * M + j] - mu[j] - sm[j] / nn;
            int t = 0;
#pragma unroll
            for (int a = 0; a < M; ++a)
#pragma unroll
                for (int b = a; b < M; ++b) W[t++] += wi * dv[a] * dv[b];
        }
    }
    int t = 0;
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
        for (int b = a; b < M; ++b) {
            const double v = block_sum<double>(W[t++], lds);
            if (threadIdx.x == 0) {
                wpart[(int64_t)blockIdx.x * M * M + a * M + b] = v;
                wpart[(int64_t)blockIdx.x * M * M + b * M + a] = v;
            }
        }
}

// The S build's level tables, one thread per level: its dimension, whether
// the dimension is sparse, its index among the sparse or the dense levels
// (-1 in the other table) and the inverse lists.  Dimensions are contiguous
// in level order, so a level's index is its dimension's base plus its offset
// within the dimension.
struct LevelDims {
    int64_t off[8];
    int64_t base[8];      // sparse (or dense) levels in the dimensions before
    int sparse[8];
    int D;
};

__global__ void k_level_tables(int64_t L, LevelDims ld, uint8_t *__restrict__ sparse_dim,
                               uint8_t *__restrict__ sparse_lv,
                               int *__restrict__ sparse_idx, int *__restrict__ dense_idx,
                               int *__restrict__ lvl_dim, int *__restrict__ sparse_levels,
                               int *__restrict__ dense_levels) {
    if (blockIdx.x == 0 && (int)threadIdx.x < ld.D) sparse_dim[threadIdx.x] = (uint8_t)ld.sparse[threadIdx.x];
    for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a < L;
         a += (int64_t)gridDim.x * blockDim.x) {
        int t = 0;
        for (int d = 1; d < ld.D; ++d)
            if (a >= ld.off[d]) t = d;
        const int64_t idx = ld.base[t] + (a - ld.off[t]);
        lvl_dim[a] = t;
        sparse_lv[a] = (uint8_t)ld.sparse[t];
        if (ld.sparse[t]) {
            sparse_idx[a] = (int)idx;
            dense_idx[a] = -1;
            sparse_levels[idx] = (int)a;
        } else {
            dense_idx[a] = (int)idx;
            sparse_idx[a] = -1;
            dense_levels[idx] = (int)a;
        }
    }
}

// obs / cnt per level of one dimension over all cells (cnt in integer
// atomics: exact; obs, with weights, in float atomics).  Small dimensions
// count in LDS bins first.  ONE: unweighted cells of one row each (cell c is
// row c), so a run's count is its length, read off the wave's head mask
// without the run scan.  LDSB: the dimension counts in LDS bins (a larger one
// allocates none, which keeps the kernel's occupancy for its gathers).
template <bool ONE, bool LDSB>
__global__ void k_obs_cnt(int64_t C, const int *__restrict__ cstart,
                          const int *__restrict__ code, int64_t off, int nl,
                          const double *__restrict__ cn, unsigned long long *__restrict__ cnt,
                          double *__restrict__ obs, int weighted) {
    constexpr int NB = LDSB ? kLdsBins / 2 : 1;
    __shared__ unsigned bins[NB];
    __shared__ double wbins[NB];
    constexpr bool lds = LDSB;
    if (lds) {
        for (int t = threadIdx.x; t < nl; t += blockDim.x) {
            bins[t] = 0;
            wbins[t] = 0.0;
        }
        __syncthreads();
    }
    // wave-uniform trip count: one atomic per run of cells at one level
    const int lane = threadIdx.x & (kWave - 1);
    for (int64_t c0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x - lane); c0 < C;
         c0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = c0 + lane;
        const bool in = c < C;
        bool last;
        unsigned long long rt;
        double wt = 0.0;
        int a;
        if (ONE) {
            a = in ? code[c] : -1;
            const int up = __shfl_up(a, 1, kWave);
            const unsigned long long hm = __ballot(lane == 0 || up != a);
            last = lane == kWave - 1 || ((hm >> (lane + 1)) & 1ull);
            const int hl = 63 - __clzll((long long)(hm & ((2ull << lane) - 1ull)));
            rt = (unsigned long long)(lane - hl + 1);
        } else {
            const int s0 = in ? cstart[c] : 0;
            const unsigned long long r = in ? (unsigned long long)(cstart[c + 1] - s0) : 0ull;
            a = in ? code[s0] : -1;
            rt = run_total<unsigned long long>(a, r, last);
            wt = weighted ? run_total<double>(a, in ? cn[c] : 0.0, last) : 0.0;
        }
        if (last && a >= 0) {
            if (lds) {
                atomicAdd(&bins[a], (unsigned)rt);
                if (weighted) atomicAdd(&wbins[a], wt);
            } else {
                atomicAdd(&cnt[off + a], rt);
                if (weighted) atomicAdd(&obs[off + a], wt);
            }
        }
    }
    if (lds) {
        __syncthreads();
        for (int t = threadIdx.x; t < nl; t += blockDim.x)
            if (bins[t]) {
                atomicAdd(&cnt[off + t], (unsigned long long)bins[t]);
                if (weighted) atomicAdd(&obs[off + t], wbins[t]);
            }
    }
}

// obs / cnt of a dimension of at most NL levels (the years): every lane keeps
// its own counts and weights per level in registers (a select per level, no
// atomics and no run scan per cell; a wave's 64 cells spread over the few
// levels made 8-way LDS atomic conflicts), four cells' loads in flight at a
// time (addresses clamped to the last cell, their counts masked), then one
// block sum and one atomic per block and level.  ONE as in k_obs_cnt.
constexpr int kObsBatch = 4;

template <int NL, bool ONE>
__global__ __launch_bounds__(256) void k_obs_cnt_small(int64_t C, const int *__restrict__ cstart,
                                                       const int *__restrict__ code, int64_t off,
                                                       int nl, const double *__restrict__ cn,
                                                       unsigned long long *__restrict__ cnt,
                                                       double *__restrict__ obs, int weighted) {
    __shared__ unsigned long long ldsk[T256 / kWave];
    __shared__ double ldsw[T256 / kWave];
    unsigned long long k[NL];
    double wv[NL];
#pragma unroll
    for (int t = 0; t < NL; ++t) {
        k[t] = 0ull;
        wv[t] = 0.0;
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c0 < C;
         c0 += kObsBatch * stride) {
        int s[kObsBatch], e[kObsBatch], a[kObsBatch];
        double wc[kObsBatch];
#pragma unroll
        for (int b = 0; b < kObsBatch; ++b) {
            const int64_t c = min(c0 + b * stride, C - 1);
            s[b] = ONE ? (int)c : cstart[c];
            e[b] = ONE ? (int)c + 1 : cstart[c + 1];
            wc[b] = weighted ? cn[c] : 0.0;
        }
#pragma unroll
        for (int b = 0; b < kObsBatch; ++b) a[b] = code[s[b]];
#pragma unroll
        for (int b = 0; b < kObsBatch; ++b) {
            const bool in = c0 + b * stride < C;
            const unsigned long long r = in ? (unsigned long long)(e[b] - s[b]) : 0ull;
            const double x = in ? wc[b] : 0.0;
#pragma unroll
            for (int t = 0; t < NL; ++t) {
                const bool h = a[b] == t;
                k[t] += h ? r : 0ull;
                wv[t] += h ? x : 0.0;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NL; ++t) {
        if (t >= nl) break;
        const unsigned long long kt = block_sum<unsigned long long>(k[t], ldsk);
        const double wt = weighted ? block_sum<double>(wv[t], ldsw) : 0.0;
        if (threadIdx.x == 0 && kt) {
            atomicAdd(&cnt[off + t], kt);
            if (weighted) atomicAdd(&obs[off + t], wt);
        }
    }
}

__global__ void k_u64_to_f64(int64_t n, const unsigned long long *__restrict__ a,
                             double *__restrict__ b, double *__restrict__ b2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        b[i] = (double)a[i];
        if (b2) b2[i] = (double)a[i];
    }
}

// Cells per group (the group starts over cells), and the identifying flag.
__global__ void k_group_cells(int64_t G, const int *__restrict__ gcstart,
                              int *__restrict__ ident) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < G;
         g += (int64_t)gridDim.x * blockDim.x)
        ident[g] = (gcstart[g + 1] - gcstart[g]) > 1 ? 1 : 0;
}

// Cells per identifying group (0 for the others; entry G is 0).
__global__ void k_ident_ncells(int64_t G, const int *__restrict__ gcstart,
                               const int *__restrict__ ident, int *__restrict__ out) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g <= G;
         g += (int64_t)gridDim.x * blockDim.x)
        out[g] = (g < G && ident[g]) ? gcstart[g + 1] - gcstart[g] : 0;
}

// Scatter the identifying cells: for group g with ident, its cells go to
// [ipos[g], ...) (ipos = exclusive sum of cells over identifying groups).
template <int W>
__global__ void k_ident_cells(int64_t G, int D, int m, const int *__restrict__ gcstart,
                              const int *__restrict__ ident, const int *__restrict__ gpos,
                              const int *__restrict__ cpos, const int *__restrict__ cstart,
                              const int *__restrict__ codes, int64_t ld,
                              const double *__restrict__ cn, const double *__restrict__ csums,
                              int *__restrict__ icodes, int64_t ild, double *__restrict__ in,
                              double *__restrict__ isums, int *__restrict__ igstart,
                              int *__restrict__ igroup) {
    const int sub = threadIdx.x % W;
    for (int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W; g < G;
         g += (int64_t)gridDim.x * blockDim.x / W) {
        if (!ident[g]) continue;
        const int q = gpos[g], o0 = cpos[g], c0 = gcstart[g];
        if (sub == 0) {
            igstart[q] = o0;
            igroup[q] = (int)g;
        }
        for (int c = c0 + sub; c < gcstart[g + 1]; c += W) {
            const int r = cstart[c], o = o0 + (c - c0);
            for (int d = 0; d < D; ++d) icodes[d * ild + o] = codes[d * ld + r];
            in[o] = cn[c];
            for (int j = 0; j < m; ++j) isums[(int64_t)o * m + j] = csums[(int64_t)c * m + j];
        }
    }
}

// Each cell's codes (those of its first row), D x C, dimension-major.
__global__ void k_cell_codes(int64_t C, int D, const int *__restrict__ cstart,
                             const int *__restrict__ codes, int64_t ld, int *__restrict__ icodes) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < C;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int r = cstart[c];
        for (int d = 0; d < D; ++d) icodes[d * C + c] = codes[d * ld + r];
    }
}

// The sparse dimensions (those whose levels are rows of S's sparse part).
struct DimList {
    int d[8];
    int n;
};

// Incidence keys of the sparse dimensions: (group, level) for every
// (identifying cell, sparse dimension j), e = c * n + j.  The dense
// dimensions' (a handful of levels each: years) incidences are never formed:
// their diagonal, right-hand side and weights come from the cells directly.
__global__ void k_inc_keys_s(int64_t M, DimList dl, const int *__restrict__ cgroup,
                             const int *__restrict__ icodes, int64_t ild,
                             const int64_t *__restrict__ offs, int lbits,
                             uint64_t *__restrict__ key, int *__restrict__ val) {
    const int64_t n = M * dl.n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / dl.n;
        const int d = dl.d[(int)(e - c * dl.n)];
        key[e] = ((uint64_t)cgroup[c] << lbits) | (uint64_t)(offs[d] + icodes[d * ild + c]);
        val[e] = (int)c;
    }
}

// U[g, t] = the weight of group g's cells at dense level t: W lanes per
// group, per level a lane sum then a fixed butterfly.
template <int W>
__global__ void k_dense_u_cells(int64_t Gi, int Ld, const int *__restrict__ igstart,
                                const int *__restrict__ icodes, int64_t ild,
                                const int64_t *__restrict__ offs, DimList dense,
                                const int *__restrict__ dense_idx, const double *__restrict__ in,
                                double *__restrict__ U) {
    const int sub = threadIdx.x % W;
    for (int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W; g < Gi;
         g += (int64_t)gridDim.x * blockDim.x / W) {
        const int c0 = igstart[g], c1 = igstart[g + 1];
        for (int t = 0; t < Ld; ++t) {
            double u = 0.0;
            for (int c = c0 + sub; c < c1; c += W)
                for (int q = 0; q < dense.n; ++q) {
                    const int d = dense.d[q];
                    if (dense_idx[offs[d] + icodes[d * ild + c]] == t) u += in[c];
                }
            u = run_sum<W>(u);
            if (sub == 0) U[g * Ld + t] = u;
        }
    }
}

// The dense levels' cell sums: per workgroup over a contiguous range of
// identifying cells, for each level t, the sum of n_c (col 0) and of the
// cell sums s_cj (col j + 1) over the cells at level t; a block sum (fixed
// tree) per (t, col).  part[b, t, col].
__global__ __launch_bounds__(256) void k_dense_cell_sums(
    int64_t M, int Ld, int m, const int *__restrict__ icodes, int64_t ild,
    const int64_t *__restrict__ offs, DimList dense, const int *__restrict__ dense_idx,
    const double *__restrict__ in, const double *__restrict__ isums, double *__restrict__ part) {
    __shared__ double lds[T256 / 64];
    const int64_t c0 = part_begin(M, blockIdx.x, gridDim.x), c1 = part_begin(M, blockIdx.x + 1, gridDim.x);
    const int np = m + 1;
    for (int t = 0; t < Ld; ++t)
        for (int j0 = 0; j0 < np; j0 += 8) {
            const int nj = min(8, np - j0);
            double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            for (int64_t c = c0 + threadIdx.x; c < c1; c += blockDim.x)
                for (int q = 0; q < dense.n; ++q) {
                    const int d = dense.d[q];
                    if (dense_idx[offs[d] + icodes[d * ild + c]] != t) continue;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        if (j >= nj) break;
                        const int col = j0 + j;
                        acc[j] += col == 0 ? in[c] : isums[c * m + col - 1];
                    }
                }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (j >= nj) break;
                const double v = block_sum<double>(acc[j], lds);
                if (threadIdx.x == 0) part[((int64_t)blockIdx.x * Ld + t) * np + j0 + j] = v;
            }
        }
}

// diag and rhs of the dense levels (kernels_base.py:64-110 restricted to
// them): diag_t = sum_c n_c - sum_g U[g,t]^2 / N_g and rhs_tj = sum_c s_cj -
// sum_g U[g,t] vbar_gj (the cells' part from k_dense_cell_sums).  Here the
// groups' part: per pair (t, col) over a range of groups in nsub slices
// added in order (fixed), like k_dd.  part[b, t, col].
__global__ __launch_bounds__(256) void k_dense_diag(int64_t Gi, int Ld, int m,
                                                    const double *__restrict__ U,
                                                    const double *__restrict__ Ng,
                                                    const double *__restrict__ vbar,
                                                    double *__restrict__ part) {
    __shared__ double sv[T256];
    const int npair = Ld * (m + 1);
    const int nsub = npair * 2 <= (int)blockDim.x ? (int)blockDim.x / npair : 1;
    const int64_t g0 = part_begin(Gi, blockIdx.x, gridDim.x), g1 = part_begin(Gi, blockIdx.x + 1, gridDim.x);
    const int stride = nsub > 1 ? npair : (int)blockDim.x;
    const int sub = nsub > 1 ? (int)threadIdx.x / npair : 0;
    for (int pr0 = 0; pr0 < npair; pr0 += stride) {
        const int pr = pr0 + (nsub > 1 ? (int)threadIdx.x % npair : (int)threadIdx.x);
        const bool act = sub < nsub && pr < npair;
        double v = 0.0;
        if (act) {
            const int t = pr / (m + 1), col = pr - t * (m + 1);
            for (int64_t g = g0 + sub; g < g1; g += nsub) {
                const double u = U[g * Ld + t];
                v += col == 0 ? u * u / Ng[g] : u * vbar[g * m + col - 1];
            }
        }
        if (nsub > 1) {
            sv[threadIdx.x] = act ? v : 0.0;
            __syncthreads();
            if ((int)threadIdx.x < npair) {
                double tv = 0.0;
                for (int q = 0; q < nsub; ++q) tv += sv[q * npair + threadIdx.x];
                part[(int64_t)blockIdx.x * npair + threadIdx.x] = tv;
            }
            __syncthreads();
        } else if (act) {
            part[(int64_t)blockIdx.x * npair + pr] = v;
        }
    }
}

// The dense levels in one pass over the groups (m <= 8 variables, LDS slots
// for every sub-group's (level, column) pairs within 64 KB): W lanes per
// group find, per dense level t, the
// weight u and the cell sums s_j of the group's cells at t (lane sums, a
// fixed butterfly, which leaves the sums on every lane); lane t of the group
// (lane 0 for every t when there are more levels than lanes) writes U[g, t]
// and adds u - u^2 / N_g and s_j - u vbar_gj (kernels_base.py:64-110) to the
// sub-group's LDS slots; the slots are added in sub-group order into
// part[b, t, col] (fixed order).
template <int W, int MM>
__global__ __launch_bounds__(256) void k_dense_sums(
    int64_t Gi, int Ld, int m, const int *__restrict__ igstart, const int *__restrict__ icodes,
    int64_t ild, const int64_t *__restrict__ offs, DimList dense,
    const int *__restrict__ dense_idx, const double *__restrict__ in,
    const double *__restrict__ isums, const double *__restrict__ Ng,
    const double *__restrict__ vbar, double *__restrict__ U, double *__restrict__ part) {
    extern __shared__ double acc[];
    // MM: m at compile time (0: runtime), so the columns' loads carry no guards
    constexpr int MJ = MM > 0 ? MM : 8;
    const int mm = MM > 0 ? MM : m;
    const int np = mm + 1, npair = Ld * np;
    const int sub = threadIdx.x % W, sg = threadIdx.x / W, nsg = blockDim.x / W;
    for (int e = threadIdx.x; e < nsg * npair; e += blockDim.x) acc[e] = 0.0;
    __syncthreads();
    const int64_t g0 = part_begin(Gi, blockIdx.x, gridDim.x), g1 = part_begin(Gi, blockIdx.x + 1, gridDim.x);
    const bool own = Ld <= W;  // lane t keeps level t's sums
    for (int64_t g = g0 + sg; g < g1; g += nsg) {
        const int c0 = igstart[g], c1 = igstart[g + 1];
        double myu = 0.0, mys[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        // one dense dimension and at most two cells per lane (the usual
        // case): each lane's cells read once, into registers
        const bool simple = dense.n == 1 && c1 - c0 <= 2 * W;
        int ta = -1, tb = -1;
        double ina = 0.0, inb = 0.0, sa[8], sb[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) sa[j] = sb[j] = 0.0;
        if (simple) {
            // both cells' loads together (a lane without a cell reads the
            // group's first and drops it)
            const int d = dense.d[0];
            const int ca = c0 + sub, cb = ca + W;
            const bool ha = ca < c1, hb = cb < c1;
            const int xa = ha ? ca : c0, xb = hb ? cb : c0;
            const int ka = icodes[d * ild + xa], kb = icodes[d * ild + xb];
            const double na = in[xa], nb = in[xb];
            double va[MJ], vb[MJ];
#pragma unroll
            for (int j = 0; j < MJ; ++j) {
                const int jj = MM > 0 ? j : min(j, mm - 1);
                va[j] = isums[(int64_t)xa * mm + jj];
                vb[j] = isums[(int64_t)xb * mm + jj];
            }
            const int da = dense_idx[offs[d] + ka], db = dense_idx[offs[d] + kb];
            ta = ha ? da : -1;
            tb = hb ? db : -1;
            ina = ha ? na : 0.0;
            inb = hb ? nb : 0.0;
#pragma unroll
            for (int j = 0; j < MJ; ++j) {
                sa[j] = ha && j < mm ? va[j] : 0.0;
                sb[j] = hb && j < mm ? vb[j] : 0.0;
            }
        }
        // distinct levels among the group's cells (one cell a worker-year):
        // level t's sums are one cell's values, which lane t takes directly
        // (the same bits as the lane sums, whose other terms are zeros)
        bool fast = false;
        if (own && simple) {
            int bits = (ta >= 0 ? 1 << ta : 0) | (tb >= 0 ? 1 << tb : 0);
            int cnt = (ta >= 0) + (tb >= 0);
#pragma unroll
            for (int o = W / 2; o > 0; o >>= 1) {
                bits |= __shfl_xor(bits, o, W);
                cnt += __shfl_xor(cnt, o, W);
            }
            fast = __popc(bits) == cnt;
        }
        if (fast) {
            const int base = (int)(threadIdx.x & (kWave - 1)) - sub;
            const uint64_t own_bits = (W == 64 ? ~0ull : ((1ull << W) - 1)) << base;
            int srca = -1, srcb = -1;
            for (int t = 0; t < Ld; ++t) {
                const uint64_t ma = __ballot(ta == t) & own_bits, mb = __ballot(tb == t) & own_bits;
                if (sub == t) {
                    srca = ma ? __builtin_ctzll(ma) : -1;
                    srcb = mb ? __builtin_ctzll(mb) : -1;
                }
            }
            const int la = srca >= 0 ? srca : base, lb = srcb >= 0 ? srcb : base;
            const double ua = __shfl(ina, la), ub = __shfl(inb, lb);
            myu = srca >= 0 ? ua : (srcb >= 0 ? ub : 0.0);
#pragma unroll
            for (int j = 0; j < MJ; ++j) {
                if (j < mm) {
                    const double va = __shfl(sa[j], la), vb = __shfl(sb[j], lb);
                    mys[j] = srca >= 0 ? va : (srcb >= 0 ? vb : 0.0);
                }
            }
        }
        for (int t = 0; t < Ld && !fast; ++t) {
            double u = 0.0, sj[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if (simple) {
                // the cells in order, as the general loop adds them
                if (ta == t) {
                    u += ina;
#pragma unroll
                    for (int j = 0; j < 8; ++j) sj[j] += sa[j];
                }
                if (tb == t) {
                    u += inb;
#pragma unroll
                    for (int j = 0; j < 8; ++j) sj[j] += sb[j];
                }
            } else {
                for (int c = c0 + sub; c < c1; c += W)
                    for (int q = 0; q < dense.n; ++q) {
                        const int d = dense.d[q];
                        if (dense_idx[offs[d] + icodes[d * ild + c]] != t) continue;
                        u += in[c];
#pragma unroll
                        for (int j = 0; j < MJ; ++j)
                            if (j < mm) sj[j] += isums[(int64_t)c * mm + j];
                    }
            }
            u = run_sum<W>(u);
#pragma unroll
            for (int j = 0; j < MJ; ++j)
                if (j < mm) sj[j] = run_sum<W>(sj[j]);
            if (own) {
                if (sub == t) {
                    myu = u;
#pragma unroll
                    for (int j = 0; j < MJ; ++j) mys[j] = sj[j];
                }
            } else if (sub == 0) {
                U[g * Ld + t] = u;
                double *a = acc + (sg * Ld + t) * np;
                a[0] += u - u * u / Ng[g];
#pragma unroll
                for (int j = 0; j < MJ; ++j)
                    if (j < mm) a[1 + j] += sj[j] - u * vbar[g * mm + j];
            }
        }
        if (own && sub < Ld) {
            U[g * Ld + sub] = myu;
            double *a = acc + (sg * Ld + sub) * np;
            a[0] += myu - myu * myu / Ng[g];
#pragma unroll
            for (int j = 0; j < MJ; ++j)
                if (j < mm) a[1 + j] += mys[j] - myu * vbar[g * mm + j];
        }
    }
    __syncthreads();
    for (int pr = threadIdx.x; pr < npair; pr += blockDim.x) {
        double v = 0.0;
        for (int q = 0; q < nsg; ++q) v += acc[q * npair + pr];
        part[(int64_t)blockIdx.x * npair + pr] = v;
    }
}

// The dense levels' sums from k_dense_sums' partials: a wave per (level,
// column) pair, lanes strided over the workgroups, a fixed butterfly.
__global__ void k_dense_sums_finish(int nb, int Ld, int m, const double *__restrict__ part,
                                    const int *__restrict__ dense_levels,
                                    double *__restrict__ diag, double *__restrict__ rhs) {
    const int npair = Ld * (m + 1);
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x / 64);
    for (int pr = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; pr < npair; pr += waves) {
        double v = 0.0;
        for (int b = lane; b < nb; b += 64) v += part[(int64_t)b * npair + pr];
        v = wave_sum(v);
        if (lane == 0) {
            const int t = pr / (m + 1), col = pr - t * (m + 1);
            const int a = dense_levels[t];
            if (col == 0) diag[a] = v;
            else if (rhs) rhs[(int64_t)a * m + col - 1] = v;
        }
    }
}

__global__ void k_dense_diag_finish(int nb, int nbc, int Ld, int m, const double *__restrict__ part,
                                    const double *__restrict__ cpart,
                                    const int *__restrict__ dense_levels, double *__restrict__ diag,
                                    double *__restrict__ rhs) {
    const int npair = Ld * (m + 1);
    for (int pr = blockIdx.x * blockDim.x + threadIdx.x; pr < npair; pr += gridDim.x * blockDim.x) {
        double cs = 0.0, gs = 0.0;
        for (int b = 0; b < nbc; ++b) cs += cpart[(int64_t)b * npair + pr];
        for (int b = 0; b < nb; ++b) gs += part[(int64_t)b * npair + pr];
        const double v = cs - gs;
        const int t = pr / (m + 1), col = pr - t * (m + 1);
        const int a = dense_levels[t];
        if (col == 0) diag[a] = v;
        else if (rhs) rhs[(int64_t)a * m + col - 1] = v;
    }
}

// Each identifying group's incidence keys put in order within the group's
// segment (keys are emitted cell-major and cells are in group order, so a
// group's keys are contiguous): W lanes per group, every key's place counted
// against the segment's others, ties by position, so the order is the one a
// stable sort gives.  A segment longer than W raises `too_long` and is left
// for the caller's radix sort.
template <int W>
__global__ void k_seg_rank_sort(int64_t Gi, int D, const int *__restrict__ igstart,
                                const uint64_t *__restrict__ kin, const int *__restrict__ vin,
                                uint64_t *__restrict__ kout, int *__restrict__ vout,
                                int *__restrict__ too_long) {
    const int sub = threadIdx.x % W;
    for (int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W; g < Gi;
         g += (int64_t)gridDim.x * blockDim.x / W) {
        const int64_t s = (int64_t)igstart[g] * D, e = (int64_t)igstart[g + 1] * D;
        const int64_t len = e - s;
        if (len > W) {
            if (sub == 0) *too_long = 1;
            continue;
        }
        const bool in = sub < len;
        const uint64_t k = in ? kin[s + sub] : ~0ull;
        int rank = 0;
        for (int t = 0; t < W; ++t) {
            const uint64_t kt = __shfl(k, t, W);
            rank += (t < len) && (kt < k || (kt == k && t < sub));
        }
        if (in) {
            kout[s + rank] = k;
            vout[s + rank] = vin[s + sub];
        }
    }
}

// Whether the rows are already in group order, and each group's first row
// (group codes are dense, so in order every group has a first row).
__global__ void k_group_order(int64_t n, int64_t G, const int *__restrict__ g,
                              int *__restrict__ start, int *__restrict__ unsorted) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int gi = g[i];
        if (i > 0 && g[i - 1] > gi) *unsorted = 1;
        if (i == 0 || g[i - 1] != gi) start[gi] = (int)i;
        if (i == n - 1) start[G] = (int)n;
    }
}

// Ranks within segments of at most W keys (k_seg_rank_sort) at the widths a
// wave allows, trying the given width first and then 64 lanes; false when a
// segment is longer (the caller's radix sort then orders them).
static int seg_rank_sort(hipStream_t s, int64_t nseg, int mult, const int *start,
                         const uint64_t *kin, const int *vin, uint64_t *kout, int *vout,
                         int W, bool &done);

template <int W>
__global__ void k_cell_group(int64_t Gi, const int *__restrict__ igstart, int *__restrict__ cg) {
    const int sub = threadIdx.x % W;
    for (int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W; g < Gi;
         g += (int64_t)gridDim.x * blockDim.x / W)
        for (int c = igstart[g] + sub; c < igstart[g + 1]; c += W) cg[c] = (int)g;
}

// Per incidence (a run of equal keys): group, level, u = sum n, s = sums.
constexpr int kIncBatch = 8;   // a run's cells loaded together up to this many
// MM: m at compile time for the batched short runs (0: runtime m, no batch)
template <int MM>
__global__ void k_inc_reduce(int64_t E, int m, int lbits, const uint64_t *__restrict__ key,
                             const int *__restrict__ cell, const int *__restrict__ istart,
                             const double *__restrict__ in, const double *__restrict__ isums,
                             int *__restrict__ ig, int *__restrict__ il, double *__restrict__ iu,
                             double *__restrict__ is) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < E;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int s = istart[q], e = istart[q + 1];
        const uint64_t k = key[s];
        ig[q] = (int)(k >> lbits);
        il[q] = (int)(k & ((1ull << lbits) - 1));
        double u = 0.0;
        if (MM > 0 && e - s <= kIncBatch) {
            // a short run (the usual case): its cells' indices, then their
            // weights and sums, loaded together (cells past the run repeat
            // its last and are not added); each sum still in run order
            constexpr int M1 = MM > 0 ? MM : 1;
            int64_t cl[kIncBatch];
#pragma unroll
            for (int b = 0; b < kIncBatch; ++b) cl[b] = cell[s + min(b, e - s - 1)];
            double nb[kIncBatch], sb[kIncBatch][M1];
#pragma unroll
            for (int b = 0; b < kIncBatch; ++b) {
                nb[b] = in[cl[b]];
#pragma unroll
                for (int j = 0; j < M1; ++j) sb[b][j] = isums[cl[b] * M1 + j];
            }
            double v[M1];
#pragma unroll
            for (int j = 0; j < M1; ++j) v[j] = 0.0;
#pragma unroll
            for (int b = 0; b < kIncBatch; ++b)
                if (b < e - s) {
                    u += nb[b];
#pragma unroll
                    for (int j = 0; j < M1; ++j) v[j] += sb[b][j];
                }
            iu[q] = u;
#pragma unroll
            for (int j = 0; j < M1; ++j) is[q * M1 + j] = v[j];
            continue;
        }
        if (m <= 8) {
            // one pass over the run, the m sums in registers (each still
            // added in run order)
            double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            int t = s;
            for (; t + 1 < e; t += 2) {  // two cells' loads in flight
                const int64_t c0 = cell[t], c1 = cell[t + 1];
                const double n0 = in[c0], n1 = in[c1];
                double s0[8], s1[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    s0[j] = j < m ? isums[c0 * m + j] : 0.0;
                    s1[j] = j < m ? isums[c1 * m + j] : 0.0;
                }
                u += n0;
                u += n1;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    v[j] += s0[j];
                    v[j] += s1[j];
                }
            }
            if (t < e) {
                const int64_t c = cell[t];
                u += in[c];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (j < m) v[j] += isums[c * m + j];
            }
            iu[q] = u;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < m) is[q * m + j] = v[j];
            continue;
        }
        for (int t = s; t < e; ++t) u += in[cell[t]];
        iu[q] = u;
        for (int j = 0; j < m; ++j) {
            double v = 0.0;
            for (int t = s; t < e; ++t) v += isums[(int64_t)cell[t] * m + j];
            is[q * m + j] = v;
        }
    }
}

// Group weight N_g and vbar_g = sum of cell sums / N_g.
template <int W>
__global__ void k_group_totals(int64_t Gi, int m, const int *__restrict__ igstart,
                               const double *__restrict__ in, const double *__restrict__ isums,
                               double *__restrict__ Ng, double *__restrict__ vbar) {
    const int sub = threadIdx.x % W;
    for (int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W; g < Gi;
         g += (int64_t)gridDim.x * blockDim.x / W) {
        const int s = igstart[g], e = igstart[g + 1];
        double N = 0.0;
        for (int c = s + sub; c < e; c += W) N += in[c];
        N = run_sum<W>(N);
        if (sub == 0) Ng[g] = N;
        for (int j = 0; j < m; ++j) {
            double v = 0.0;
            for (int c = s + sub; c < e; c += W) v += isums[(int64_t)c * m + j];
            v = run_sum<W>(v);
            if (sub == 0) vbar[g * m + j] = v / N;
        }
    }
}

// First incidence of each group (incidences are sorted by group, level).
__global__ void k_group_inc_start(int64_t E, const int *__restrict__ ig,
                                  int *__restrict__ gis, int64_t Gi) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < E;
         q += (int64_t)gridDim.x * blockDim.x)
        if (q == 0 || ig[q] != ig[q - 1]) gis[ig[q]] = (int)q;
    if (blockIdx.x == 0 && threadIdx.x == 0) gis[Gi] = (int)E;
}

// lptr[a] = first position of level a in the level-sorted keys (level in
// the bits above gbits), by binary search; lptr[L] = E.
__global__ void k_level_ptr(int64_t L, int64_t E, const uint64_t *__restrict__ key, int gbits,
                            int *__restrict__ lptr) {
    for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a <= L;
         a += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = E;
        const uint64_t target = (uint64_t)a << gbits;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (key[mid] < target) lo = mid + 1;
            else hi = mid;
        }
        lptr[a] = (int)lo;
    }
}

__global__ void k_level_keys(int64_t E, const int *__restrict__ ig, const int *__restrict__ il,
                             int gbits, uint64_t *__restrict__ key, int *__restrict__ val) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < E;
         q += (int64_t)gridDim.x * blockDim.x) {
        key[q] = ((uint64_t)il[q] << gbits) | (uint64_t)ig[q];
        val[q] = (int)q;
    }
}

// diag[a] = sum over groups containing a of (u - u^2 / N_g)  (kernels_base.py:86-110)
// rhs[a, j] = sum of (s_j - u vbar_j)                          (kernels_base.py:64-83)
// Sums over a level's incidences in fixed chunks of kChunk incidences: one
// wave per chunk (lanes strided, then a fixed butterfly) writes a partial; a
// level in millions of groups (a year) is spread over many waves.  The
// partials of a level are then added in chunk order.
constexpr int kChunk = 2048;

// Each incidence's terms, in incidence order (its group's N_g and vbar are
// read next to its neighbours'): contrib[q] = (u - u^2 / N_g, s_j - u vbar_j).
__global__ void k_inc_contrib(int64_t E, int m, const int *__restrict__ ig,
                              const double *__restrict__ iu, const double *__restrict__ is,
                              const double *__restrict__ Ng, const double *__restrict__ vbar,
                              double *__restrict__ contrib) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < E;
         q += (int64_t)gridDim.x * blockDim.x) {
        const double u = iu[q];
        const int gq = ig[q];
        double *o = contrib + q * (m + 1);
        o[0] = u - u * u / Ng[gq];
        for (int j = 0; j < m; ++j) o[1 + j] = is[q * m + j] - u * vbar[(int64_t)gq * m + j];
    }
}

// The level sums of the contributions: one wave per fixed chunk of a level's
// incidences (lanes strided, then a fixed butterfly) writes a partial, one
// gather of a contiguous (m + 1)-row per incidence.
__global__ void k_diag_rhs_part_c(int64_t nchunk, int m, const int *__restrict__ chunk_lvl,
                                  const int *__restrict__ chunk_beg, const int *__restrict__ lptr,
                                  const int *__restrict__ lq, const double *__restrict__ contrib,
                                  double *__restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t c = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; c < nchunk;
         c += waves) {
        const int a = chunk_lvl[c];
        const int b = chunk_beg[c], e = min(b + kChunk, lptr[a + 1]);
        for (int j0 = 0; j0 <= m; j0 += 8) {
            const int nj = min(8, m + 1 - j0);
            double r[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) r[j] = 0.0;
            for (int t = b + lane; t < e; t += 64) {
                const double *o = contrib + (int64_t)lq[t] * (m + 1) + j0;
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (j < nj) r[j] += o[j];
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (j >= nj) break;
                const double v = wave_sum(r[j]);
                if (lane == 0) part[c * (m + 1) + j0 + j] = v;
            }
        }
    }
}

// One wave per level adds its chunks' partials (lanes strided, a fixed
// butterfly): a level of millions of incidences (a year) has thousands of
// chunks, which one thread would add in a long dependent chain.
__global__ void k_diag_rhs_sum(int64_t L, int m, const int *__restrict__ cptr,
                               const double *__restrict__ part, double *__restrict__ diag,
                               double *__restrict__ rhs) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t a = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; a < L;
         a += waves) {
        const int c0 = cptr[a], c1 = cptr[a + 1];
        double d = 0.0;
        for (int c = c0 + lane; c < c1; c += 64) d += part[(int64_t)c * (m + 1)];
        d = wave_sum(d);
        if (lane == 0) diag[a] = d;
        if (rhs)
            for (int j = 0; j < m; ++j) {
                double r = 0.0;
                for (int c = c0 + lane; c < c1; c += 64) r += part[(int64_t)c * (m + 1) + 1 + j];
                r = wave_sum(r);
                if (lane == 0) rhs[a * m + j] = r;
            }
    }
}

// Chunks per level (at least one, so every level gets a partial).
__global__ void k_level_nchunk(int64_t L, const int *__restrict__ lptr, int *__restrict__ nch) {
    for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a <= L;
         a += (int64_t)gridDim.x * blockDim.x)
        nch[a] = a < L ? max(1, (lptr[a + 1] - lptr[a] + kChunk - 1) / kChunk) : 0;
}

__global__ void k_level_chunks(int64_t L, const int *__restrict__ lptr, const int *__restrict__ cptr,
                               int *__restrict__ chunk_lvl, int *__restrict__ chunk_beg) {
    for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a < L;
         a += (int64_t)gridDim.x * blockDim.x)
        for (int c = cptr[a], b = lptr[a]; c < cptr[a + 1]; ++c, b += kChunk) {
            chunk_lvl[c] = (int)a;
            chunk_beg[c] = b;
        }
}

// Entries of S between sparse levels: per group, every ordered pair (a, b)
// of its sparse incidences gives [a == b] u_a - u_a u_b / N_g; for sparse
// dimensions d1 != d2 each cell adds n_c at its (a, b).  Count, then emit.
__global__ void k_ss_count(int64_t Gi, const int *__restrict__ gis,
                           const int *__restrict__ il, const uint8_t *__restrict__ sparse_lv,
                           const int *__restrict__ igstart, int ncross,
                           int64_t *__restrict__ cnt) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < Gi;
         g += (int64_t)gridDim.x * blockDim.x) {
        int64_t q = 0;
        for (int t = gis[g]; t < gis[g + 1]; ++t) q += sparse_lv[il[t]];
        cnt[g] = q * q + (int64_t)ncross * (igstart[g + 1] - igstart[g]);
    }
}

__global__ void k_ss_emit(int64_t Gi, int D, const int *__restrict__ gis,
                          const int *__restrict__ il, const double *__restrict__ iu,
                          const uint8_t *__restrict__ sparse_lv, const double *__restrict__ Ng,
                          const int *__restrict__ igstart, const int *__restrict__ icodes,
                          int64_t ild, const int64_t *__restrict__ offs,
                          const uint8_t *__restrict__ sparse_dim, int lbits,
                          const int64_t *__restrict__ pos, const double *__restrict__ in,
                          uint64_t *__restrict__ key, double *__restrict__ val) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < Gi;
         g += (int64_t)gridDim.x * blockDim.x) {
        int64_t o = pos[g];
        const double N = Ng[g];
        for (int t1 = gis[g]; t1 < gis[g + 1]; ++t1) {
            const int a = il[t1];
            if (!sparse_lv[a]) continue;
            const double ua = iu[t1];
            for (int t2 = gis[g]; t2 < gis[g + 1]; ++t2) {
                const int b = il[t2];
                if (!sparse_lv[b]) continue;
                key[o] = ((uint64_t)a << lbits) | (uint64_t)b;
                val[o] = (a == b ? ua : 0.0) - ua * iu[t2] / N;
                ++o;
            }
        }
        for (int c = igstart[g]; c < igstart[g + 1]; ++c)
            for (int d1 = 0; d1 < D; ++d1) {
                if (!sparse_dim[d1]) continue;
                for (int d2 = 0; d2 < D; ++d2) {
                    if (d2 == d1 || !sparse_dim[d2]) continue;
                    key[o] = ((uint64_t)(offs[d1] + icodes[d1 * ild + c]) << lbits) |
                             (uint64_t)(offs[d2] + icodes[d2 * ild + c]);
                    val[o] = in[c];
                    ++o;
                }
            }
    }
}

// Sum runs of equal keys (sorted); out row/col/val at the run's rank.
__global__ void k_reduce_runs(int64_t n, const uint64_t *__restrict__ key,
                              const double *__restrict__ val, const int *__restrict__ head,
                              const int *__restrict__ rank, int lbits, int *__restrict__ row,
                              int *__restrict__ col, double *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (!head[i]) continue;
        const uint64_t k = key[i];
        double s = val[i];
        int64_t t = i + 1;
        // four keys and values loaded together, added in order up to the run's end
        bool more = true;
        while (more && t + 3 < n) {
            const uint64_t k0 = key[t], k1 = key[t + 1], k2 = key[t + 2], k3 = key[t + 3];
            const double v0 = val[t], v1 = val[t + 1], v2 = val[t + 2], v3 = val[t + 3];
            if (k0 != k) { more = false; break; }
            s += v0;
            if (k1 != k) { more = false; break; }
            s += v1;
            if (k2 != k) { more = false; break; }
            s += v2;
            if (k3 != k) { more = false; break; }
            s += v3;
            t += 4;
        }
        if (more)
            for (; t < n && key[t] == k; ++t) s += val[t];
        const int r = rank[i] - 1;
        row[r] = (int)(k >> lbits);
        col[r] = (int)(k & ((1ull << lbits) - 1));
        out[r] = s;
    }
}

// U[g, t] = u of dense level t in group g (0 when absent).

// Sparse x dense block, first per incidence: for incidence q = (group g,
// sparse level a) and dense level t, C[q, t] = n_g(a, t) - u_a U[g, t] / N_g
// (n_g(a, t): the weight of g's cells holding both a and t), with bit t of
// F[q] set when t occurs in g.  One scan of g's cells per incidence.
__global__ void k_sd_inc(int64_t E, int D, int Ld, const int *__restrict__ ig,
                         const int *__restrict__ il, const double *__restrict__ iu,
                         const int *__restrict__ sidx, const uint8_t *__restrict__ sparse_lv,
                         const int *__restrict__ lvl_dim, const int64_t *__restrict__ offs,
                         const uint8_t *__restrict__ sparse_dim, const int *__restrict__ dense_idx,
                         const double *__restrict__ U, const double *__restrict__ Ng,
                         const int *__restrict__ igstart, const int *__restrict__ icodes,
                         int64_t ild, const double *__restrict__ in, double *__restrict__ Cq,
                         uint64_t *__restrict__ Fq, int dd1) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < E;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int a = il[q];
        if (!sparse_lv[a]) continue;
        const int64_t r = sidx[q];
        const int g = ig[q];
        const int da = lvl_dim[a];
        const int la = a - (int)offs[da];
        const double f = iu[q] / Ng[g];
        uint64_t bits = 0;
        if (Ld <= 8) {
            // the row in registers (the same terms in the same order)
            double rr[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const double ut = t < Ld ? U[(int64_t)g * Ld + t] : 0.0;
                rr[t] = -f * ut;
                if (ut != 0.0) bits |= 1ull << t;
            }
            const int c0 = igstart[g], c1 = igstart[g + 1];
            if (dd1 >= 0 && c1 - c0 <= kIncBatch) {
                // one dense dimension (dd1) and a short group (the usual
                // case): the cells' codes and weights, then their dense
                // slots, loaded together; added in cell order as below
                int ca[kIncBatch], ct[kIncBatch];
                double cv[kIncBatch];
#pragma unroll
                for (int b = 0; b < kIncBatch; ++b) {
                    const int c = c0 + min(b, c1 - c0 - 1);
                    ca[b] = icodes[da * ild + c];
                    ct[b] = icodes[dd1 * ild + c];
                    cv[b] = in[c];
                }
#pragma unroll
                for (int b = 0; b < kIncBatch; ++b) ct[b] = dense_idx[offs[dd1] + ct[b]];
#pragma unroll
                for (int b = 0; b < kIncBatch; ++b)
                    if (b < c1 - c0 && ca[b] == la)
#pragma unroll
                        for (int t = 0; t < 8; ++t)
                            if (ct[b] == t) rr[t] += cv[b];
            } else {
                for (int c = c0; c < c1; ++c) {
                    if (icodes[da * ild + c] != la) continue;
                    for (int d = 0; d < D; ++d) {
                        if (sparse_dim[d]) continue;
                        const int tc = dense_idx[offs[d] + icodes[d * ild + c]];
                        const double v = in[c];
#pragma unroll
                        for (int t = 0; t < 8; ++t)
                            if (tc == t) rr[t] += v;
                    }
                }
            }
            double *row = Cq + r * Ld;
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (t < Ld) row[t] = rr[t];
            Fq[r] = bits;
            continue;
        }
        double *row = Cq + r * Ld;
        for (int t = 0; t < Ld; ++t) {
            const double ut = U[(int64_t)g * Ld + t];
            row[t] = -f * ut;
            if (ut != 0.0) bits |= 1ull << t;
        }
        for (int c = igstart[g]; c < igstart[g + 1]; ++c) {
            if (icodes[da * ild + c] != la) continue;
            for (int d = 0; d < D; ++d) {
                if (sparse_dim[d]) continue;
                const int t = dense_idx[offs[d] + icodes[d * ild + c]];
                row[t] += in[c];
            }
        }
        Fq[r] = bits;
    }
}

// Then per (sparse level a, dense level t): the sum over a's incidences, in
// level-CSR order (fixed), and the presence.
__global__ void k_sd_sum(int64_t Ls, int Ld, const int *__restrict__ sparse_levels,
                         const int *__restrict__ lptr, const int *__restrict__ lq,
                         const int *__restrict__ sidx, const double *__restrict__ Cq,
                         const uint64_t *__restrict__ Fq, double *__restrict__ PD,
                         uint8_t *__restrict__ PP) {
    const int64_t n = Ls * (int64_t)Ld;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = e / Ld;
        const int t = (int)(e - s * Ld);
        const int a = sparse_levels[s];
        double v = 0.0;
        uint64_t pres = 0;
        int p = lptr[a];
        const int pe = lptr[a + 1];
        for (; p + 3 < pe; p += 4) {  // four incidences' loads in flight, added in order
            const int64_t r0 = sidx[lq[p]], r1 = sidx[lq[p + 1]], r2 = sidx[lq[p + 2]],
                          r3 = sidx[lq[p + 3]];
            const double c0 = Cq[r0 * Ld + t], c1 = Cq[r1 * Ld + t], c2 = Cq[r2 * Ld + t],
                         c3 = Cq[r3 * Ld + t];
            pres |= Fq[r0] | Fq[r1] | Fq[r2] | Fq[r3];
            v += c0;
            v += c1;
            v += c2;
            v += c3;
        }
        for (; p < pe; ++p) {
            const int64_t r = sidx[lq[p]];
            v += Cq[r * Ld + t];
            pres |= Fq[r];
        }
        PD[e] = v;
        PP[e] = (uint8_t)((pres >> t) & 1ull);
    }
}

__global__ void k_sparse_flag(int64_t E, const int *__restrict__ il,
                              const uint8_t *__restrict__ sparse_lv, int *__restrict__ flag) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q <= E;
         q += (int64_t)gridDim.x * blockDim.x)
        flag[q] = q < E ? (int)sparse_lv[il[q]] : 0;
}

// Dense x dense block: per workgroup of groups, thread per (t1, t2) pair:
// sum of n_g(t1, t2) - U_t1 U_t2 / N_g and a presence count; partials
// reduced in a fixed order by k_dd_finish.
__global__ __launch_bounds__(256) void k_dd(int64_t Gi, int Ld, const int *__restrict__ dense_levels,
                     const int *__restrict__ lvl_dim, const int64_t *__restrict__ offs,
                     const double *__restrict__ U, const double *__restrict__ Ng,
                     const int *__restrict__ igstart, const int *__restrict__ icodes,
                     int64_t ild, const double *__restrict__ in, double *__restrict__ part,
                     int *__restrict__ ppart) {
    __shared__ double sv[T256];
    __shared__ int sp[T256];
    const int npair = Ld * Ld;
    // with few pairs, nsub slices of the block's groups run side by side
    // and are added in slice order (fixed)
    const int nsub = npair * 2 <= (int)blockDim.x ? (int)blockDim.x / npair : 1;
    const int64_t g0 = part_begin(Gi, blockIdx.x, gridDim.x), g1 = part_begin(Gi, blockIdx.x + 1, gridDim.x);
    const int stride = nsub > 1 ? npair : (int)blockDim.x;
    const int sub = nsub > 1 ? (int)threadIdx.x / npair : 0;
    for (int pr0 = 0; pr0 < npair; pr0 += stride) {
        const int pr = pr0 + (nsub > 1 ? (int)threadIdx.x % npair : (int)threadIdx.x);
        const bool act = sub < nsub && pr < npair;
        double v = 0.0;
        int pres = 0;
        if (act) {
            const int t1 = pr / Ld, t2 = pr % Ld;
            const int b1 = dense_levels[t1], b2 = dense_levels[t2];
            const int d1 = lvl_dim[b1], d2 = lvl_dim[b2];
            const int l1 = b1 - (int)offs[d1], l2 = b2 - (int)offs[d2];
            int64_t g = g0 + sub;
            if (d1 == d2) {
                // eight groups' loads issued together, added in group order
                for (; g + 7 * nsub < g1; g += 8 * nsub) {
                    double a[8], b[8], n[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const int64_t gi = g + i * nsub;
                        a[i] = U[gi * Ld + t1];
                        b[i] = U[gi * Ld + t2];
                        n[i] = Ng[gi];
                    }
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        if (a[i] == 0.0 || b[i] == 0.0) continue;
                        pres = 1;
                        const double nat = (t1 == t2) ? a[i] : 0.0;
                        v += nat - a[i] * b[i] / n[i];
                    }
                }
            }
            for (; g < g1; g += nsub) {
                const double u1 = U[g * Ld + t1], u2 = U[g * Ld + t2];
                if (u1 == 0.0 || u2 == 0.0) continue;
                pres = 1;
                double nat;
                if (d1 == d2) {
                    nat = (t1 == t2) ? u1 : 0.0;
                } else {
                    nat = 0.0;
                    for (int c = igstart[g]; c < igstart[g + 1]; ++c)
                        if (icodes[d1 * ild + c] == l1 && icodes[d2 * ild + c] == l2) nat += in[c];
                }
                v += nat - u1 * u2 / Ng[g];
            }
        }
        if (nsub > 1) {
            sv[threadIdx.x] = v;
            sp[threadIdx.x] = pres;
            __syncthreads();
            if ((int)threadIdx.x < npair) {
                double tv = 0.0;
                int tp = 0;
                for (int q = 0; q < nsub; ++q) {
                    tv += sv[q * npair + threadIdx.x];
                    tp |= sp[q * npair + threadIdx.x];
                }
                part[(int64_t)blockIdx.x * npair + threadIdx.x] = tv;
                ppart[(int64_t)blockIdx.x * npair + threadIdx.x] = tp;
            }
            __syncthreads();
        } else if (act) {
            part[(int64_t)blockIdx.x * npair + pr] = v;
            ppart[(int64_t)blockIdx.x * npair + pr] = pres;
        }
    }
}

// Dense x dense block when every dense level belongs to one dimension (the
// years) and there are LD of them: a lane per group reads the group's LD
// sums U_g (one row) and Ng once and adds every pair (t1 <= t2) of
// [t1 == t2] U_t1 - U_t1 U_t2 / N_g in registers (each group's row was read
// once per pair by k_dd), the groups of a lane in order, then a fixed block
// tree per pair; presence as a bit mask per lane, OR-ed in LDS.  The
// partials' layout is k_dd's (both triangles written).
template <int LD>
__global__ __launch_bounds__(256) void k_dd_rows(int64_t Gi, const double *__restrict__ U,
                                                 const double *__restrict__ Ng,
                                                 double *__restrict__ part,
                                                 int *__restrict__ ppart) {
    constexpr int NP = LD * (LD + 1) / 2;
    __shared__ double lds[T256 / kWave];
    __shared__ unsigned long long pm;
    double acc[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[q] = 0.0;
    unsigned long long mask = 0ull;
    if (threadIdx.x == 0) pm = 0ull;
    const int64_t g0 = part_begin(Gi, blockIdx.x, gridDim.x), g1 = part_begin(Gi, blockIdx.x + 1, gridDim.x);
    for (int64_t g = g0 + threadIdx.x; g < g1; g += blockDim.x) {
        double u[LD];
#pragma unroll
        for (int t = 0; t < LD; ++t) u[t] = U[g * LD + t];
        const double n = Ng[g];
        int q = 0;
#pragma unroll
        for (int t1 = 0; t1 < LD; ++t1)
#pragma unroll
            for (int t2 = t1; t2 < LD; ++t2, ++q) {
                const bool h = u[t1] != 0.0 && u[t2] != 0.0;
                const double nat = t1 == t2 ? u[t1] : 0.0;
                acc[q] += h ? nat - u[t1] * u[t2] / n : 0.0;
                mask |= h ? 1ull << q : 0ull;
            }
    }
    __syncthreads();
    if (mask) atomicOr(&pm, mask);
    int q = 0;
#pragma unroll
    for (int t1 = 0; t1 < LD; ++t1)
#pragma unroll
        for (int t2 = t1; t2 < LD; ++t2, ++q) {
            const double v = block_sum<double>(acc[q], lds);
            if (threadIdx.x == 0) {
                const int pres = (int)((pm >> q) & 1ull);
                part[(int64_t)blockIdx.x * LD * LD + t1 * LD + t2] = v;
                part[(int64_t)blockIdx.x * LD * LD + t2 * LD + t1] = v;
                ppart[(int64_t)blockIdx.x * LD * LD + t1 * LD + t2] = pres;
                ppart[(int64_t)blockIdx.x * LD * LD + t2 * LD + t1] = pres;
            }
        }
}

// A wave per pair: lanes strided over the workgroups' partials, a fixed
// butterfly.
__global__ void k_dd_finish(int nb, int npair, const double *__restrict__ part,
                            const int *__restrict__ ppart, double *__restrict__ DD,
                            uint8_t *__restrict__ DP) {
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x / 64);
    for (int pr = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; pr < npair; pr += waves) {
        double v = 0.0;
        int p = 0;
        for (int b = lane; b < nb; b += 64) {
            v += part[(int64_t)b * npair + pr];
            p |= ppart[(int64_t)b * npair + pr];
        }
        v = wave_sum(v);
        p = __any(p) ? 1 : 0;
        if (lane == 0) {
            DD[pr] = v;
            DP[pr] = (uint8_t)p;
        }
    }
}

// Row lengths of S.
__global__ void k_row_nnz(int64_t L, int Ld, const int *__restrict__ ss_rowcnt,
                          const int *__restrict__ sparse_idx, const int *__restrict__ dense_idx,
                          const uint8_t *__restrict__ PP, const int *__restrict__ colpres,
                          const uint8_t *__restrict__ DP, int *__restrict__ nnz) {
    for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a < L;
         a += (int64_t)gridDim.x * blockDim.x) {
        int c = 0;
        const int s = sparse_idx[a];
        if (s >= 0) {
            c = ss_rowcnt[s];
            for (int t = 0; t < Ld; ++t) c += PP[(int64_t)s * Ld + t];
        } else {
            const int t = dense_idx[a];
            c = colpres[t];
            for (int t2 = 0; t2 < Ld; ++t2) c += DP[t * Ld + t2];
        }
        nnz[a] = c;
    }
}

// Sparse rows: merge the sorted sparse columns with the present dense ones.
__global__ void k_fill_sparse_rows(int64_t Ls, int Ld, const int *__restrict__ sparse_levels,
                                   const int *__restrict__ dense_levels,
                                   const int *__restrict__ ss_rowptr, const int *__restrict__ ss_col,
                                   const double *__restrict__ ss_val, const uint8_t *__restrict__ PP,
                                   const double *__restrict__ PD, const int *__restrict__ indptr,
                                   int *__restrict__ indices, double *__restrict__ data) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < Ls;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int a = sparse_levels[s];
        int o = indptr[a];
        int p = ss_rowptr[s];
        const int pe = ss_rowptr[s + 1];
        int t = 0;
        while (t < Ld && !PP[s * Ld + t]) ++t;
        while (p < pe || t < Ld) {
            const bool take_dense = t < Ld && (p >= pe || dense_levels[t] < ss_col[p]);
            if (take_dense) {
                indices[o] = dense_levels[t];
                data[o] = PD[s * Ld + t];
                ++t;
                while (t < Ld && !PP[s * Ld + t]) ++t;
            } else {
                indices[o] = ss_col[p];
                data[o] = ss_val[p];
                ++p;
            }
            ++o;
        }
    }
}

// The same rows with W lanes per row: each entry's place is found, not
// walked to.  A sparse entry lands after the row's sparse entries before it
// and the present dense columns below its column; a dense entry after the
// present dense columns before it and the sparse columns below it (a binary
// search of the row).  Dense and sparse columns never coincide, so the
// places are the merge's, and the sparse entries' stores are coalesced.
template <int W>
__global__ void k_fill_sparse_rows_w(int64_t Ls, int Ld, const int *__restrict__ sparse_levels,
                                     const int *__restrict__ dense_levels,
                                     const int *__restrict__ ss_rowptr,
                                     const int *__restrict__ ss_col,
                                     const double *__restrict__ ss_val,
                                     const uint8_t *__restrict__ PP, const double *__restrict__ PD,
                                     const int *__restrict__ indptr, int *__restrict__ indices,
                                     double *__restrict__ data) {
    const int sub = threadIdx.x % W;
    for (int64_t s = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W; s < Ls;
         s += (int64_t)gridDim.x * blockDim.x / W) {
        const int o0 = indptr[sparse_levels[s]];
        const int p0 = ss_rowptr[s], pe = ss_rowptr[s + 1];
        const uint8_t *pp = PP + s * Ld;
        for (int p = p0 + sub; p < pe; p += W) {
            const int col = ss_col[p];
            int below = 0;
            for (int t = 0; t < Ld; ++t) below += (pp[t] && dense_levels[t] < col) ? 1 : 0;
            indices[o0 + (p - p0) + below] = col;
            data[o0 + (p - p0) + below] = ss_val[p];
        }
        for (int t = sub; t < Ld; t += W) {
            if (!pp[t]) continue;
            const int dl = dense_levels[t];
            int rank = 0;
            for (int u = 0; u < t; ++u) rank += pp[u] ? 1 : 0;
            int lo = p0, hi = pe;
            while (lo < hi) {
                const int mid = lo + (hi - lo) / 2;
                if (ss_col[mid] < dl) lo = mid + 1;
                else hi = mid;
            }
            const int o = o0 + rank + (lo - p0);
            indices[o] = dl;
            data[o] = PD[s * Ld + t];
        }
    }
}

// Dense rows, sparse columns: entry (t, a) at the rank of a among the
// present sparse columns plus the present dense columns below a.
__global__ void k_fill_dense_sparse(int64_t Ls, int Ld, const int *__restrict__ sparse_levels,
                                    const int *__restrict__ dense_levels,
                                    const uint8_t *__restrict__ PP, const double *__restrict__ PD,
                                    const int *__restrict__ cpos, const uint8_t *__restrict__ DP,
                                    const int *__restrict__ indptr, int *__restrict__ indices,
                                    double *__restrict__ data) {
    const int64_t n = Ls * (int64_t)Ld;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = e / Ld;
        const int t = (int)(e - s * Ld);
        if (!PP[e]) continue;
        const int a = sparse_levels[s];
        int below = 0;
        for (int t2 = 0; t2 < Ld; ++t2) below += (DP[t * Ld + t2] && dense_levels[t2] < a);
        const int o = indptr[dense_levels[t]] + cpos[(int64_t)t * (Ls + 1) + s] + below;
        indices[o] = a;
        data[o] = PD[e];
    }
}

// Dense rows, dense columns.
__global__ void k_fill_dense_dense(int64_t Ls, int Ld, const int *__restrict__ sparse_levels,
                                   const int *__restrict__ dense_levels,
                                   const int *__restrict__ cpos, const uint8_t *__restrict__ DP,
                                   const double *__restrict__ DD, const int *__restrict__ indptr,
                                   int *__restrict__ indices, double *__restrict__ data) {
    const int npair = Ld * Ld;
    for (int pr = blockIdx.x * blockDim.x + threadIdx.x; pr < npair; pr += gridDim.x * blockDim.x) {
        const int t = pr / Ld, t2 = pr % Ld;
        if (!DP[pr]) continue;
        const int b = dense_levels[t2];
        // present sparse columns below b: sparse levels are ascending
        int lo = 0, hi = (int)Ls;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sparse_levels[mid] < b) lo = mid + 1;
            else hi = mid;
        }
        int below = cpos[(int64_t)t * (Ls + 1) + lo];
        for (int u = 0; u < t2; ++u) below += DP[t * Ld + u];
        const int o = indptr[dense_levels[t]] + below;
        indices[o] = b;
        data[o] = DD[pr];
    }
}

__global__ void k_jacobi(int64_t L, const double *__restrict__ d, double *__restrict__ dinv) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < L;
         i += (int64_t)gridDim.x * blockDim.x)
        dinv[i] = d[i] > 1e-12 ? 1.0 / d[i] : 0.0;  // core/solve.py:142-146
}

__global__ void k_pp_col(int64_t Ls, int Ld, int t, const uint8_t *__restrict__ PP,
                         int *__restrict__ col) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < Ls;
         s += (int64_t)gridDim.x * blockDim.x)
        col[s] = PP[s * Ld + t];
}

__global__ void k_ss_rowcnt(int64_t n, const int *__restrict__ row, const int *__restrict__ sparse_idx,
                            int *__restrict__ cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[sparse_idx[row[i]]], 1);
}

// ------------------------------------------------------- components

// Union-find on the levels of dimension 0 linked through identifying groups
// (core/kernels_base.py:155-174) with uf_root / uf_join (common.h); roots end
// as each component's smallest level, whatever the order of the unions.
__global__ void k_uf_init(int64_t n, int *parent) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        parent[i] = (int)i;
}

// One pass of unions over the groups: a group's cells (W lanes per group)
// join their level to the first cell's.  Each union walks both roots and
// hooks the larger root under the smaller with a CAS, retrying from the
// values a failed CAS returns until both walks meet, so when the pass ends
// every group's levels share a root, and each set's root is its smallest
// level whatever the order the unions ran in.  Every parent pointer, however
// stale the copy a lane reads, points within its set, so a walk that meets
// proves the union.  A level equal to the cell before it is skipped (a
// group's cells repeat its levels, one cell per level of every other
// dimension).
template <int W>
__global__ void k_uf_union(int64_t Gi, const int *__restrict__ igstart,
                           const int *__restrict__ icodes, int *parent) {
    const int sub = threadIdx.x % W;
    for (int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W; g < Gi;
         g += (int64_t)gridDim.x * blockDim.x / W) {
        const int s = igstart[g], e = igstart[g + 1];
        const int a0 = icodes[s];
        for (int c = s + 1 + sub; c < e; c += W) {
            const int lc = icodes[c];
            if (lc == a0 || lc == icodes[c - 1]) continue;
            uf_join(parent, a0, lc);
        }
    }
}

// Unions of the rows' pairs of levels (a_i, nb0 + b_i) of two non-streamed
// dimensions; a row repeating the pair of the row before it is skipped.
__global__ void k_uf_rows(int64_t n, const int *__restrict__ ca, const int *__restrict__ cb,
                          int nb0, int *parent) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int x = ca[i], y = nb0 + cb[i];
        if (i > 0 && ca[i - 1] == x && nb0 + cb[i - 1] == y) continue;
        uf_join(parent, x, y);
    }
}

// Sets bit x * ny + y of a bitmap.  A bit already set is only read, so the
// many rows of a few-level pair do not contend on atomics.
__device__ __forceinline__ void set_edge(unsigned *bits, int x, int y, int64_t ny) {
    const uint64_t b = (uint64_t)x * (uint64_t)ny + (uint64_t)y;
    unsigned *wd = bits + (b >> 5);
    const unsigned mk = 1u << (b & 31);
    if (!(*wd & mk)) atomicOr(wd, mk);
}

// The distinct edges (ca[i], cb[i]) of two non-streamed dimensions' rows.
__global__ void k_pair_bits(int64_t n, const int *__restrict__ ca, const int *__restrict__ cb,
                            int64_t ny, unsigned *bits) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        set_edge(bits, ca[i], cb[i], ny);
}

// The streamed dimension's groups join their levels of cb: the edges
// (level on the group's first row, level on each other row), W lanes per
// group.
template <int W>
__global__ void k_group_pair_bits(int64_t G, const int *__restrict__ gstart,
                                  const int *__restrict__ cb, int64_t ny, unsigned *bits) {
    const int sub = threadIdx.x % W;
    for (int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W; g < G;
         g += (int64_t)gridDim.x * blockDim.x / W) {
        const int s = gstart[g], e = gstart[g + 1];
        const int x = cb[s];
        for (int i = s + 1 + sub; i < e; i += W) {
            const int y = cb[i];
            if (y != x) set_edge(bits, x, y, ny);
        }
    }
}

// A pair with a side of at most 64 levels (years, say): its components are
// those of that side's levels, two of them linked when some level of the
// other side (or group of the streamed dimension) has rows in both.  Each
// such level's 64-bit mask of small-side levels m is ORed into adj[y] for
// every y in m: first into the workgroup's copy in LDS, then into adj once
// per workgroup (integer ORs: the result is exact).
__device__ __forceinline__ void adj_add(unsigned long long *lds, unsigned long long m) {
    for (unsigned long long v = m; v; v &= v - 1) atomicOr(&lds[__builtin_ctzll(v)], m);
}

__device__ __forceinline__ void adj_flush(unsigned long long *lds, unsigned long long *adj) {
    __syncthreads();
    for (int y = threadIdx.x; y < 64; y += blockDim.x)
        if (lds[y]) atomicOr(&adj[y], lds[y]);
}

// The streamed dimension's groups: each group's mask of its levels of cb.
template <int W>
__global__ void k_group_masks(int64_t G, const int *__restrict__ gstart,
                              const int *__restrict__ cb, unsigned long long *adj) {
    __shared__ unsigned long long lds[64];
    for (int y = threadIdx.x; y < 64; y += blockDim.x) lds[y] = 0;
    __syncthreads();
    const int sub = threadIdx.x % W;
    for (int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W; g < G;
         g += (int64_t)gridDim.x * blockDim.x / W) {
        unsigned long long m = 0;
        for (int i = gstart[g] + sub; i < gstart[g + 1]; i += W) m |= 1ull << cb[i];
#pragma unroll
        for (int o = W / 2; o > 0; o >>= 1) m |= __shfl_xor(m, o, W);
        if (sub == 0 && (m & (m - 1))) adj_add(lds, m);
    }
    adj_flush(lds, adj);
}

// Another dimension: each of its levels' mask of small-side levels, from
// the rows (masks (na), zeroed first), then the masks into adj.
__global__ void k_row_masks(int64_t n, const int *__restrict__ ca, const int *__restrict__ cb,
                            unsigned long long *masks) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long bit = 1ull << cb[i];
        unsigned long long *mp = masks + ca[i];
        if (!(*mp & bit)) atomicOr(mp, bit);
    }
}

__global__ void k_mask_adj(int64_t na, const unsigned long long *__restrict__ masks,
                           unsigned long long *adj) {
    __shared__ unsigned long long lds[64];
    for (int y = threadIdx.x; y < 64; y += blockDim.x) lds[y] = 0;
    __syncthreads();
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < na;
         x += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long m = masks[x];
        if (m & (m - 1)) adj_add(lds, m);
    }
    adj_flush(lds, adj);
}

// Unions of the edges a bitmap marks (bit x * ny + y joins x and yoff + y).
__global__ void k_uf_bits(int64_t nwords, const unsigned *__restrict__ bits, int64_t ny,
                          int yoff, int *parent) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nwords;
         q += (int64_t)gridDim.x * blockDim.x) {
        unsigned v = bits[q];
        while (v) {
            const int t = __builtin_ctz(v);
            v &= v - 1;
            const uint64_t b = ((uint64_t)q << 5) + t;
            uf_join(parent, (int)(b / ny), yoff + (int)(b % ny));
        }
    }
}

// The roots among n union-find entries (one per wave's lanes, summed with
// integer atomics: exact).
__global__ void k_uf_count_roots(int64_t n, const int *__restrict__ parent,
                                 unsigned long long *count) {
    unsigned long long c = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        c += parent[i] == (int)i;
    c = wave_sum(c);
    if ((threadIdx.x & (kWave - 1)) == 0 && c) atomicAdd(count, c);
}

__global__ void k_uf_flatten(int64_t n, int *parent) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        parent[i] = uf_root(parent, (int)i);
}

// rows per component root (exact: integer counts)
// A wave whose active lanes share one root adds their sum with one atomic
// (a giant component would otherwise serialise every level on one address).
__global__ void k_comp_rows(int64_t n, const int *__restrict__ root,
                            const double *__restrict__ cnt, unsigned long long *__restrict__ rows,
                            int *__restrict__ is_root) {
    const int64_t span = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += span) {
        const int64_t i = base + threadIdx.x;
        const bool ok = i < n;
        const int r = ok ? root[i] : -1;
        const unsigned long long c = ok ? (unsigned long long)cnt[i] : 0ull;
        if (ok) is_root[i] = r == (int)i;
        const int r0 = __shfl(r, 0, 64);
        const bool uniform = __all(!ok || r == r0) && r0 >= 0;
        if (uniform) {
            unsigned long long s = c;
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if ((threadIdx.x & 63) == 0) atomicAdd(&rows[r0], s);
        } else if (ok) {
            atomicAdd(&rows[r], c);
        }
    }
}

// sort key of a root: rows descending, then root ascending (stable
// argsort of -rows over roots in level order, core/passes.py:946-948)
__global__ void k_comp_keys(int64_t n, const int *__restrict__ is_root,
                            const unsigned long long *__restrict__ rows,
                            const int *__restrict__ rpos, uint64_t *__restrict__ key,
                            int *__restrict__ val) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (!is_root[i]) continue;
        const int r = rpos[i];
        key[r] = ((uint64_t)(0xFFFFFFFFull - (rows[i] & 0xFFFFFFFFull)) << 32) | (uint64_t)i;
        val[r] = (int)i;
    }
}

__global__ void k_comp_rank(int64_t nc, const int *__restrict__ sorted_roots,
                            int *__restrict__ rank_of_root) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nc;
         r += (int64_t)gridDim.x * blockDim.x)
        rank_of_root[sorted_roots[r]] = (int)r;
}

__global__ void k_comp_label(int64_t n, const int *__restrict__ root,
                             const int *__restrict__ rank_of_root, int *__restrict__ comp) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        comp[i] = rank_of_root[root[i]];
}

__global__ void k_comp_level_keys(int64_t nl, int shift, const int *__restrict__ comp,
                                  uint64_t *__restrict__ key, int *__restrict__ val) {
    for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a < nl;
         a += (int64_t)gridDim.x * blockDim.x) {
        key[a] = ((uint64_t)comp[a] << shift) | (uint64_t)a;
        val[a] = (int)a;
    }
}

// cptr (nc + 1) from the levels sorted by component: where each component's
// run starts (every component has a level, its root).
__global__ void k_comp_counts(int64_t nl, int shift, const uint64_t *__restrict__ skey,
                              int *__restrict__ cptr) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nl;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(skey[p] >> shift);
        const int prev = p > 0 ? (int)(skey[p - 1] >> shift) : -1;
        for (int q = prev + 1; q <= c; ++q) cptr[q] = (int)p;
        if (p == nl - 1) cptr[c + 1] = (int)nl;
    }
}

// Chunks of at most kCompChunk sorted levels within one component.
constexpr int kCompChunk = 4096;

__global__ void k_comp_nchunk(int64_t nc, const int *__restrict__ cptr, int *__restrict__ nch) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c <= nc;
         c += (int64_t)gridDim.x * blockDim.x)
        nch[c] = c < nc ? max(1, (cptr[c + 1] - cptr[c] + kCompChunk - 1) / kCompChunk) : 0;
}

__global__ void k_comp_chunk_ends(int64_t nc, const int *__restrict__ cptr,
                                  const int *__restrict__ chp, int *__restrict__ ch_end) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc;
         c += (int64_t)gridDim.x * blockDim.x)
        for (int q = chp[c]; q < chp[c + 1]; ++q) ch_end[q] = cptr[c + 1];
}

__global__ void k_comp_chunks(int64_t nc, const int *__restrict__ cptr,
                              const int *__restrict__ chp, int *__restrict__ ch_beg) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc;
         c += (int64_t)gridDim.x * blockDim.x)
        for (int q = chp[c], b = cptr[c]; q < chp[c + 1]; ++q, b += kCompChunk) ch_beg[q] = b;
}

// --------------------------------------------------- normalisation

// Component means of Gamma's dimension-0 block, weighted by obs, from the
// levels sorted by component: a workgroup per chunk (at most kCompChunk
// levels of one component) writes (sum obs, sum obs G_j), then each
// component adds its chunks in order.
__global__ void k_comp_chunk_sums(int64_t nchunk, int k, const int *__restrict__ ch_beg,
                                  const int *__restrict__ chp_end, const int *__restrict__ clev,
                                  const double *__restrict__ obs, const double *__restrict__ Gm,
                                  double *__restrict__ part) {
    __shared__ double lds[T256 / 64];
    for (int64_t q = blockIdx.x; q < nchunk; q += gridDim.x) {
        const int b = ch_beg[q], e = min(b + kCompChunk, chp_end[q]);
        for (int j = -1; j < k; ++j) {
            double v = 0.0;
            for (int p = b + threadIdx.x; p < e; p += blockDim.x) {
                const int a = clev[p];
                v += j < 0 ? obs[a] : obs[a] * Gm[(int64_t)a * k + j];
            }
            v = block_sum<double>(v, lds);
            if (threadIdx.x == 0) part[q * (k + 1) + j + 1] = v;
        }
    }
}

__global__ void k_comp_means(int64_t nc, int k, const int *__restrict__ chp,
                             const double *__restrict__ part, double *__restrict__ mean) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nc * k;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / k;
        const int j = (int)(e - c * k);
        double sw = 0.0, s = 0.0;
        for (int q = chp[c]; q < chp[c + 1]; ++q) {
            sw += part[(int64_t)q * (k + 1)];
            s += part[(int64_t)q * (k + 1) + 1 + j];
        }
        mean[e] = s / (sw == 0.0 ? 1.0 : sw);
    }
}

__global__ void k_sub_comp_means(int64_t nl, int k, const int *__restrict__ comp,
                                 const double *__restrict__ mean, double *__restrict__ Gm) {
    const int64_t n = nl * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = e / k;
        const int j = (int)(e - a * k);
        Gm[e] -= mean[(int64_t)comp[a] * k + j];
    }
}

// Weighted column sums of a block of levels [o, o + nl) in fixed blocks.
__global__ void k_wsum_block(int64_t nl, int k, int64_t o, const double *__restrict__ obs,
                             const double *__restrict__ Gm, double *__restrict__ part) {
    __shared__ double lds[T256];
    const int64_t r0 = part_begin(nl, blockIdx.x, gridDim.x), r1 = part_begin(nl, blockIdx.x + 1, gridDim.x);
    for (int j = 0; j <= k; ++j) {
        double s = 0.0;
        for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x)
            s += j < k ? obs[o + i] * Gm[(o + i) * k + j] : obs[o + i];
        s = block_sum<double>(s, lds);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * (k + 1) + j] = s;
    }
}

__global__ void k_sub_global_mean(int64_t nl, int k, int64_t o, int nb,
                                  const double *__restrict__ part, double *__restrict__ Gm) {
    // every thread forms the means from the partials in the same order
    const int64_t n = nl * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e % k);
        double s = 0.0, sw = 0.0;
        for (int b = 0; b < nb; ++b) {
            s += part[(int64_t)b * (k + 1) + j];
            sw += part[(int64_t)b * (k + 1) + k];
        }
        Gm[o * k + e] -= s / (sw > 1.0 ? sw : 1.0);
    }
}

// --------------------------------------------------- assembly / pass 2

// Row i's covariates col[c] less their fitted effects, x - sum_d Gm[level_d,
// col] (0 where col[c] < 0): the row's first kRowDims dimensions' codes are
// loaded together, then their effects -- two waits, not two per dimension
// (a runtime loop over the dimensions waited on each code and each effect
// in turn).
constexpr int kRowDims = 3;
template <int K>
__device__ __forceinline__ void partial_row(int64_t i, int D, int m, const int *__restrict__ codes,
                                            int64_t ld, const int64_t *__restrict__ offs,
                                            const double *__restrict__ V,
                                            const double *__restrict__ Gm, const int (&col)[K],
                                            double (&u)[K]) {
    int64_t lv[kRowDims];
#pragma unroll
    for (int d = 0; d < kRowDims; ++d) {
        const int dd = min(d, D - 1);   // a valid dimension (its effect weighted 0 past D)
        lv[d] = (offs[dd] + codes[dd * ld + i]) * m;
    }
    // every column's value and effects loaded before any is used
    double x[K], g[K][kRowDims];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const int cc = max(col[c], 0);
        x[c] = V[i * m + cc];
#pragma unroll
        for (int d = 0; d < kRowDims; ++d) g[c][d] = Gm[lv[d] + cc];
    }
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const int cc = max(col[c], 0);
#pragma unroll
        for (int d = 0; d < kRowDims; ++d) x[c] -= d < D ? g[c][d] : 0.0;
        for (int d = kRowDims; d < D; ++d) x[c] -= Gm[(offs[d] + codes[d * ld + i]) * m + cc];
        u[c] = col[c] < 0 ? 0.0 : x[c];
    }
}

// v[j] -= sum_d Gm[level_d(r), j] for the MM columns of row (or cell) r,
// its codes first and then every column's effects in flight together, as
// partial_row.
template <int MM>
__device__ __forceinline__ void minus_effects(int64_t r, int D, const int *__restrict__ codes,
                                              int64_t ld, const int64_t *__restrict__ offs,
                                              const double *__restrict__ Gm, double (&v)[MM]) {
    int64_t lv[kRowDims];
#pragma unroll
    for (int d = 0; d < kRowDims; ++d) {
        const int dd = min(d, D - 1);
        lv[d] = (offs[dd] + codes[dd * ld + r]) * MM;
    }
    double g[MM][kRowDims];
#pragma unroll
    for (int j = 0; j < MM; ++j)
#pragma unroll
        for (int d = 0; d < kRowDims; ++d) g[j][d] = Gm[lv[d] + j];
#pragma unroll
    for (int j = 0; j < MM; ++j)
#pragma unroll
        for (int d = 0; d < kRowDims; ++d) v[j] -= d < D ? g[j][d] : 0.0;
    for (int d = kRowDims; d < D; ++d) {
        const int64_t l = (offs[d] + codes[d * ld + r]) * MM;
#pragma unroll
        for (int j = 0; j < MM; ++j) v[j] -= Gm[l + j];
    }
}


// V' M_D V between cells (core/kernels_base.py:510-540), m <= 8: per group
// q_c = sums_c / n_c - sum_d Gamma[level] less the group mean.  MM is m,
// fixed at compile time so the m x m sums stay in registers.
// A group's cells run over W adjacent lanes.
template <int MM, int W>
__global__ __launch_bounds__(T256) void k_assemble(
    int64_t Gi, int D, const int *__restrict__ igstart, const int *__restrict__ icodes,
    int64_t ild, const int64_t *__restrict__ offs, const double *__restrict__ in,
    const double *__restrict__ isums, const double *__restrict__ Gm, double *__restrict__ part) {
    __shared__ double lds[T256 / 64];
    double A[MM * MM];
#pragma unroll
    for (int t = 0; t < MM * MM; ++t) A[t] = 0.0;
    const int sub = threadIdx.x % W;
    const int64_t g0 = part_begin(Gi, blockIdx.x, gridDim.x), g1 = part_begin(Gi, blockIdx.x + 1, gridDim.x);
    for (int64_t g = g0 + threadIdx.x / W; g < g1; g += blockDim.x / W) {
        const int s = igstart[g], e = igstart[g + 1];
        double qg[MM], v0[MM], N = 0.0;
#pragma unroll
        for (int j = 0; j < MM; ++j) qg[j] = v0[j] = 0.0;
        for (int c = s + sub; c < e; c += W) {
            const double ic = in[c];
            N += ic;
            double v[MM];
#pragma unroll
            for (int j = 0; j < MM; ++j) v[j] = isums[(int64_t)c * MM + j] / ic;
            minus_effects<MM>(c, D, icodes, ild, offs, Gm, v);
#pragma unroll
            for (int j = 0; j < MM; ++j) {
                if (c == s + sub) v0[j] = v[j];
                qg[j] += ic * v[j];
            }
        }
        N = run_sum<W>(N);
#pragma unroll
        for (int j = 0; j < MM; ++j) qg[j] = run_sum<W>(qg[j]) / N;
        // a lane's first cell: its values kept from the pass above
        for (int c = s + sub; c < e; c += W) {
            const double ic = in[c];
            double q[MM], v[MM];
            if (c == s + sub) {
#pragma unroll
                for (int j = 0; j < MM; ++j) v[j] = v0[j];
            } else {
#pragma unroll
                for (int j = 0; j < MM; ++j) v[j] = isums[(int64_t)c * MM + j] / ic;
                minus_effects<MM>(c, D, icodes, ild, offs, Gm, v);
            }
#pragma unroll
            for (int j = 0; j < MM; ++j) q[j] = v[j] - qg[j];
#pragma unroll
            for (int a = 0; a < MM; ++a)
#pragma unroll
                for (int b = 0; b < MM; ++b) A[a * MM + b] += ic * q[a] * q[b];
        }
    }
#pragma unroll
    for (int t = 0; t < MM * MM; ++t) {
        const double v = block_sum<double>(A[t], lds);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * MM * MM + t] = v;
    }
}

// V'M_D V for more than eight variables, by 8 x 8 tiles (columns ja + a,
// jb + b for a < na, b < nb): k_assemble's between-cell part and k_cells'
// within-cell scatter for one tile, the same products in the same order as
// the one-tile kernels take them per entry.
template <int W>
__global__ __launch_bounds__(T256) void k_assemble_tile(
    int64_t Gi, int D, int m, int ja, int na, int jb, int nbc, const int *__restrict__ igstart,
    const int *__restrict__ icodes, int64_t ild, const int64_t *__restrict__ offs,
    const double *__restrict__ in, const double *__restrict__ isums,
    const double *__restrict__ Gm, double *__restrict__ part) {
    __shared__ double lds[T256 / 64];
    double A[64];
#pragma unroll
    for (int t = 0; t < 64; ++t) A[t] = 0.0;
    const int sub = threadIdx.x % W;
    const int64_t g0 = part_begin(Gi, blockIdx.x, gridDim.x), g1 = part_begin(Gi, blockIdx.x + 1, gridDim.x);
    for (int64_t g = g0 + threadIdx.x / W; g < g1; g += blockDim.x / W) {
        const int s = igstart[g], e = igstart[g + 1];
        double qa[8], qb[8], N = 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) qa[j] = qb[j] = 0.0;
        for (int c = s + sub; c < e; c += W) {
            const double ic = in[c];
            N += ic;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (j < na) {
                    double v = isums[(int64_t)c * m + ja + j] / ic;
                    for (int d = 0; d < D; ++d) v -= Gm[(offs[d] + icodes[d * ild + c]) * m + ja + j];
                    qa[j] += ic * v;
                }
                if (j < nbc) {
                    double v = isums[(int64_t)c * m + jb + j] / ic;
                    for (int d = 0; d < D; ++d) v -= Gm[(offs[d] + icodes[d * ild + c]) * m + jb + j];
                    qb[j] += ic * v;
                }
            }
        }
        N = run_sum<W>(N);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            qa[j] = run_sum<W>(qa[j]) / N;
            qb[j] = run_sum<W>(qb[j]) / N;
        }
        for (int c = s + sub; c < e; c += W) {
            const double ic = in[c];
            double pa[8], pb[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                pa[j] = pb[j] = 0.0;
                if (j < na) {
                    double v = isums[(int64_t)c * m + ja + j] / ic;
                    for (int d = 0; d < D; ++d) v -= Gm[(offs[d] + icodes[d * ild + c]) * m + ja + j];
                    pa[j] = v - qa[j];
                }
                if (j < nbc) {
                    double v = isums[(int64_t)c * m + jb + j] / ic;
                    for (int d = 0; d < D; ++d) v -= Gm[(offs[d] + icodes[d * ild + c]) * m + jb + j];
                    pb[j] = v - qb[j];
                }
            }
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) A[a * 8 + b] += ic * pa[a] * pb[b];
        }
    }
#pragma unroll
    for (int t = 0; t < 64; ++t) {
        const double v = block_sum<double>(A[t], lds);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * 64 + t] = v;
    }
}

// The within-cell scatter of one 8 x 8 tile (k_cells' W for m > 8): per
// cell, sum over its rows of w (V_a - mean_a - s_a / n)(V_b - mean_b - s_b / n).
__global__ __launch_bounds__(T256) void k_cells_tile(int64_t C, int m, int ja, int na, int jb,
                                                     int nbc, const int *__restrict__ cstart,
                                                     const double *__restrict__ V,
                                                     const double *__restrict__ w,
                                                     const double *__restrict__ mean,
                                                     const double *__restrict__ cn,
                                                     const double *__restrict__ csums,
                                                     double *__restrict__ wpart) {
    __shared__ double lds[T256];
    double Wt[64];
#pragma unroll
    for (int t = 0; t < 64; ++t) Wt[t] = 0.0;
    const int64_t c0 = part_begin(C, blockIdx.x, gridDim.x), c1 = part_begin(C, blockIdx.x + 1, gridDim.x);
    for (int64_t c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
        const int s = cstart[c], e = cstart[c + 1];
        if (e - s < 2) continue;  // one row: every deviation is zero
        const double nn = cn[c];
        for (int i = s; i < e; ++i) {
            const double wi = w ? w[i] : 1.0;
            double da[8], db[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                da[j] = j < na ? V[(int64_t)i * m + ja + j] - mean[ja + j] - csums[c * m + ja + j] / nn : 0.0;
                db[j] = j < nbc ? V[(int64_t)i * m + jb + j] - mean[jb + j] - csums[c * m + jb + j] / nn : 0.0;
            }
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) Wt[a * 8 + b] += wi * da[a] * db[b];
        }
    }
#pragma unroll
    for (int t = 0; t < 64; ++t) {
        const double v = block_sum<double>(Wt[t], lds);
        if (threadIdx.x == 0) wpart[(int64_t)blockIdx.x * 64 + t] = v;
    }
}

// Streamed effects and residuals (core/kernels_base.py:543-619): per group
// the weighted mean of y - x'beta - sum_d gam_y[level], the residual the
// rest.  V column 0 is y, columns xcols[...] the regressors.
// A group's rows run over W adjacent lanes; a lane keeps its first row's
// value in a register and corrects any further rows in place.
template <int W>
__global__ void k_pass2(int64_t G, int D, int m, const int *__restrict__ gstart,
                        const int *__restrict__ codes, int64_t ld,
                        const int64_t *__restrict__ offs, const double *__restrict__ V,
                        const double *__restrict__ w, const double *__restrict__ gam_y,
                        const double *__restrict__ beta, const int *__restrict__ xcols, int kx,
                        double *__restrict__ geff, double *__restrict__ resid) {
    const int sub = threadIdx.x % W;
    for (int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W; g < G;
         g += (int64_t)gridDim.x * blockDim.x / W) {
        const int s = gstart[g], e = gstart[g + 1], i0 = s + sub;
        double Wg = 0.0, mu = 0.0, u0 = 0.0;
        for (int i = i0; i < e; i += W) {
            double u = V[(int64_t)i * m];
            for (int j = 0; j < kx; ++j) u -= V[(int64_t)i * m + xcols[j]] * beta[j];
            for (int d = 0; d < D; ++d) u -= gam_y[offs[d] + codes[d * ld + i]];
            const double wi = w ? w[i] : 1.0;
            Wg += wi;
            mu += wi * u;
            if (i == i0) u0 = u;
            else resid[i] = u;
        }
        Wg = run_sum<W>(Wg);
        mu = run_sum<W>(mu) / Wg;
        if (sub == 0) geff[g] = mu;
        if (i0 < e) resid[i0] = u0 - mu;
        for (int i = i0 + W; i < e; i += W) resid[i] -= mu;
    }
}

// The rows' columns cols (kc) of V less their fixed-effect fit: for each
// group, u_i = V[i, col] - sum_d Gm[level_d(i), col], and out[i, c] = u_i
// less the group's weighted mean of u (the covariates partialled on the
// fixed effects, M_F x, from their reduced solutions in Gm (L, m)).
template <int W>
__global__ void k_partial(int64_t G, int D, int m, const int *__restrict__ gstart,
                          const int *__restrict__ codes, int64_t ld,
                          const int64_t *__restrict__ offs, const double *__restrict__ V,
                          const double *__restrict__ w, const double *__restrict__ Gm,
                          const int *__restrict__ cols, int kc, double *__restrict__ out) {
    const int sub = threadIdx.x % W;
    for (int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W; g < G;
         g += (int64_t)gridDim.x * blockDim.x / W) {
        const int s = gstart[g], e = gstart[g + 1];
        for (int c = 0; c < kc; ++c) {
            const int col = cols[c];
            double Wg = 0.0, mu = 0.0;
            for (int i = s + sub; i < e; i += W) {
                double u = V[(int64_t)i * m + col];
                for (int d = 0; d < D; ++d) u -= Gm[(offs[d] + codes[d * ld + i]) * m + col];
                const double wi = w ? w[i] : 1.0;
                Wg += wi;
                mu += wi * u;
                out[(int64_t)i * kc + c] = u;
            }
            Wg = run_sum<W>(Wg);
            mu = run_sum<W>(mu) / Wg;
            for (int i = s + sub; i < e; i += W) out[(int64_t)i * kc + c] -= mu;
        }
    }
}

// The coefficients' scores u_i = w_i e_i x~_i (x~ the covariates cols,
// kc <= K of them, partialled on the fixed effects as k_partial makes them;
// e the residuals) and their sums in one pass over the groups
// (core/inference.py:430-668, the meats of the residual pass): each
// workgroup writes part[b] = [sum w e^2, sum_i u_i u_i' (upper triangle),
// sum_g s_g s_g' (upper triangle; s_g = the sum of u over group g's rows)],
// summed in a fixed order (lanes by a butterfly, waves in order), and U
// (n, kc) holds the scores when not null.  Covariates c >= kc are zero.
template <int K>
struct ScoreAcc {
    static constexpr int T = K * (K + 1) / 2;
    static constexpr int NV = 1 + 2 * T;
};

// The covariates' columns, by value (no host-to-device copy before the
// launch).
struct ColSel {
    int c[8];
};

template <int W, int K>
__global__ void __launch_bounds__(256)
    k_scores(int64_t G, int D, int m, const int *__restrict__ gstart, const int *__restrict__ codes,
             int64_t ld, const int64_t *__restrict__ offs, const double *__restrict__ V,
             const double *__restrict__ w, const double *__restrict__ Gm,
             ColSel cols, int kc, const double *__restrict__ e,
             double *__restrict__ U, double *__restrict__ part) {
    constexpr int T = ScoreAcc<K>::T, NV = ScoreAcc<K>::NV;
    const int sub = threadIdx.x % W;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    int col[K];
#pragma unroll
    for (int c = 0; c < K; ++c) col[c] = c < kc ? cols.c[c] : -1;
    for (int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W; g < G;
         g += (int64_t)gridDim.x * blockDim.x / W) {
        const int s = gstart[g], en = gstart[g + 1];
        // the group's weighted means of the covariates less their fits
        double Wg = 0.0, mu[K];
#pragma unroll
        for (int c = 0; c < K; ++c) mu[c] = 0.0;
        // the lane's first row kept from this sweep for the next
        double u0[K];
        for (int i = s + sub; i < en; i += W) {
            const double wi = w ? w[i] : 1.0;
            Wg += wi;
            double u[K];
            partial_row<K>(i, D, m, codes, ld, offs, V, Gm, col, u);
#pragma unroll
            for (int c = 0; c < K; ++c) {
                mu[c] += wi * u[c];
                if (i == s + sub) u0[c] = u[c];
            }
        }
        Wg = run_sum<W>(Wg);
#pragma unroll
        for (int c = 0; c < K; ++c) mu[c] = run_sum<W>(mu[c]) / Wg;
        double sg[K];
#pragma unroll
        for (int c = 0; c < K; ++c) sg[c] = 0.0;
        for (int i = s + sub; i < en; i += W) {
            const double wi = w ? w[i] : 1.0, f = wi * e[i];
            acc[0] += f * e[i];
            double x[K], u[K];
            if (i == s + sub) {
#pragma unroll
                for (int c = 0; c < K; ++c) u[c] = u0[c];
            } else {
                partial_row<K>(i, D, m, codes, ld, offs, V, Gm, col, u);
            }
#pragma unroll
            for (int c = 0; c < K; ++c) {
                x[c] = col[c] < 0 ? 0.0 : f * (u[c] - mu[c]);
                sg[c] += x[c];
                if (U && col[c] >= 0) U[(int64_t)i * kc + c] = x[c];
            }
            int t = 1;
#pragma unroll
            for (int a = 0; a < K; ++a)
#pragma unroll
                for (int b = a; b < K; ++b) acc[t++] += x[a] * x[b];
        }
#pragma unroll
        for (int c = 0; c < K; ++c) sg[c] = run_sum<W>(sg[c]);
        if (sub == 0) {
            int t = 1 + T;
#pragma unroll
            for (int a = 0; a < K; ++a)
#pragma unroll
                for (int b = a; b < K; ++b) acc[t++] += sg[a] * sg[b];
        }
    }
    // the workgroup's sums: each wave by a butterfly, then the waves in order
    __shared__ double red[256 / kWave][NV];
    const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double t = wave_sum(acc[v]);
        if (lane == 0) red[wv][v] = t;
    }
    __syncthreads();
    for (int v = threadIdx.x; v < NV; v += blockDim.x) {
        double t = 0.0;
        for (int q = 0; q < (int)(blockDim.x / kWave); ++q) t += red[q][v];
        part[(int64_t)blockIdx.x * NV + v] = t;
    }
}

// gam_y = Gamma[:, 0] - Gamma[:, xcols] beta
__global__ void k_gam_y(int64_t L, int m, const double *__restrict__ Gm,
                        const double *__restrict__ beta, const int *__restrict__ xcols, int kx,
                        double *__restrict__ out) {
    for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a < L;
         a += (int64_t)gridDim.x * blockDim.x) {
        double v = Gm[a * m];
        for (int j = 0; j < kx; ++j) v -= Gm[a * m + xcols[j]] * beta[j];
        out[a] = v;
    }
}

// ------------------------------------------------------------ host

struct Panel {
    hipStream_t s = nullptr;
    int dev = 0;  // the device of s
    hipStream_t side = nullptr;  // copies to the host beside the panel's work
    hipEvent_t side_ev = nullptr;
    Temp tmp;
    int64_t n_in = 0, n = 0;  // rows given / kept
    int nfe = 0, D = 0, m = 0, gdim = 0;
    std::vector<int> dims;         // user index of each non-streamed dimension
    std::vector<int64_t> nlev;     // their levels
    std::vector<int64_t> offs_h;
    int64_t s_budget = -1;   // bytes the explicit S build may take (-1: no limit)
    int64_t s_skipped = 0;   // the estimate of an S not built for the budget (0: built)
    int64_t G = 0, L = 0, C = 0, M = 0, Gi = 0, E = 0;
    int weighted = 0;
    int64_t singletons = 0, single_rounds = 0;
    DevBuf<int> g, codes, gstart, cstart, cgstart;  // sorted rows; codes (D, n)
    DevBuf<double> V, w;
    DevBuf<int> perm;                      // sorted row -> input row
    DevBuf<int64_t> glevels;               // streamed level values
    std::vector<DevBuf<int64_t>> levels;   // other dimensions' level values
    DevBuf<int64_t> offs;
    DevBuf<double> means, raw_ss, W;
    DevBuf<double> obs, cnt;
    // identifying cells
    DevBuf<int> icodes, igstart, igroup, cgroup;
    DevBuf<double> in, isums;
    // incidences
    DevBuf<int> ig, il, gis, lptr, lq;
    DevBuf<double> iu, is, Ng, vbar;
    DevBuf<double> diag, rhs;
    // components
    DevBuf<int> comp, cptr, clev;
    DevBuf<int> cchp, cch_beg, cch_end;  // chunks of the component-sorted levels
    int64_t ncomp = 0, ncchunk = 0;
    // S
    void *matrix = nullptr;
    int64_t nnz = 0;
    int n_sparse_ss = 0;
    double t_build = 0.0;
};

// Dense-range factorization: values within a range of at most a few times
// n are coded through a presence table and its prefix sum (no sort).
__global__ void k_mark_present(int64_t n, const int64_t *__restrict__ col, int64_t lo,
                               int *__restrict__ present) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        present[col[i] - lo] = 1;
}

__global__ void k_dense_codes(int64_t n, const int64_t *__restrict__ col, int64_t lo,
                              const int *__restrict__ rank, int *__restrict__ code) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        code[i] = rank[col[i] - lo];
}

__global__ void k_dense_levels(int64_t range, const int *__restrict__ present,
                               const int *__restrict__ rank, int64_t lo,
                               int64_t *__restrict__ levels) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < range;
         r += (int64_t)gridDim.x * blockDim.x)
        if (present[r]) levels[rank[r]] = lo + r;
}

static int seg_rank_sort(hipStream_t s, int64_t nseg, int mult, const int *start,
                         const uint64_t *kin, const int *vin, uint64_t *kout, int *vout,
                         int W, bool &done) {
    done = false;
    DevBuf<int> flag;
    DEV_ALLOC(flag, 1, s);
    for (int attempt = 0; attempt < 2 && !done; ++attempt) {
        if (attempt == 1) {
            if (W >= 64) break;
            W = 64;
        }
        HIP_TRY(hipMemsetAsync(flag.p, 0, 4, s));
#define HDFE_SEG(WW)                                                                             \
    hipLaunchKernelGGL(k_seg_rank_sort<WW>, dim3(grid1(nseg * WW)), dim3(T256), 0, s, nseg, mult, \
                       start, kin, vin, kout, vout, flag.p)
        switch (W) {
        case 4: HDFE_SEG(4); break;
        case 8: HDFE_SEG(8); break;
        case 16: HDFE_SEG(16); break;
        case 32: HDFE_SEG(32); break;
        default: HDFE_SEG(64); break;
        }
#undef HDFE_SEG
        int tl = 0;
        HIP_TRY(scalar_to_host(&tl, flag.p, s));
        done = !tl;
    }
    return HDFE_STATUS_OK;
}

// Codes of an int64 column in sorted level order (core/passes.py:329-335):
// code (n) and the level values (nl), on stream s.  A column whose values
// span at most max(2n, 2^22) integers is coded through a presence table
// (the usual case for ids, years, and ids already coded); any other by a
// radix sort of the values.  Both give every value its rank among the
// distinct values.
int factorize_col(hipStream_t s, Temp &tmp, const int64_t *col, int64_t n, DevBuf<int> &code,
                  DevBuf<int64_t> &levels, int64_t &nl) {
    // min / max for the key range
    const int nb = 1024;
    DevBuf<int64_t> part;
    DEV_ALLOC(part, 2 * nb, s);
    hipLaunchKernelGGL(k_minmax_partial, dim3(nb), dim3(T256), 0, s, n, col, part.p);
    std::vector<int64_t> h;
    HIP_TRY(to_host(h, part.p, 2 * nb, s));
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int b = 0; b < nb; ++b) {
        lo = std::min(lo, h[2 * b]);
        hi = std::max(hi, h[2 * b + 1]);
    }
    const uint64_t span = (uint64_t)hi - (uint64_t)lo;
    if (n > 0 && span < (uint64_t)std::max<int64_t>(2 * n, (int64_t)1 << 22) &&
        span < (uint64_t)INT32_MAX - 1) {
        const int64_t range = (int64_t)span + 1;
        DevBuf<int> present, rank;
        DEV_ALLOC(present, range + 1, s);
        DEV_ALLOC(rank, range + 1, s);
        HIP_TRY(hipMemsetAsync(present.p, 0, (range + 1) * 4, s));
        hipLaunchKernelGGL(k_mark_present, dim3(grid1(n)), dim3(T256), 0, s, n, col, lo,
                           present.p);
        HIP_TRY(exclusive_sum(tmp, present.p, rank.p, range + 1, s));
        int cnt = 0;
        HIP_TRY(scalar_to_host(&cnt, rank.p + range, s));
        nl = cnt;
        DEV_ALLOC(code, n, s);
        DEV_ALLOC(levels, nl, s);
        hipLaunchKernelGGL(k_dense_codes, dim3(grid1(n)), dim3(T256), 0, s, n, col, lo, rank.p,
                           code.p);
        hipLaunchKernelGGL(k_dense_levels, dim3(grid1(range)), dim3(T256), 0, s, range,
                           present.p, rank.p, lo, levels.p);
        HIP_TRY(hipGetLastError());
        return HDFE_STATUS_OK;
    }
    const int bits = std::max(1, bits_for((uint64_t)(hi - lo)));
    DevBuf<uint64_t> k1, k2;
    DevBuf<int> i1, i2, head, rank;
    DEV_ALLOC(k1, n, s);
    DEV_ALLOC(k2, n, s);
    DEV_ALLOC(i1, n, s);
    DEV_ALLOC(i2, n, s);
    DEV_ALLOC(head, n, s);
    DEV_ALLOC(rank, n, s);
    hipLaunchKernelGGL(k_key_from_col, dim3(grid1(n)), dim3(T256), 0, s, n, col, lo, k1.p, i1.p);
    HIP_TRY(sort_pairs(tmp, k1.p, k2.p, i1.p, i2.p, n, 0, bits, s));
    hipLaunchKernelGGL(k_heads_u64, dim3(grid1(n)), dim3(T256), 0, s, n, k2.p, head.p);
    HIP_TRY(inclusive_sum(tmp, head.p, rank.p, n, s));
    int last = 0;
    HIP_TRY(scalar_to_host(&last, rank.p + n - 1, s));
    nl = last;
    DEV_ALLOC(code, n, s);
    DEV_ALLOC(levels, nl, s);
    hipLaunchKernelGGL(k_scatter_codes, dim3(grid1(n)), dim3(T256), 0, s, n, k2.p, lo, i2.p,
                       head.p, rank.p, code.p, levels.p);
    HIP_TRY(hipGetLastError());
    return HDFE_STATUS_OK;
}

static int factorize(Panel &P, const int64_t *col, int64_t n, DevBuf<int> &code,
                     DevBuf<int64_t> &levels, int64_t &nl) {
    return factorize_col(P.s, P.tmp, col, n, code, levels, nl);
}

// Re-code dimension values to the levels present among kept rows.
static int recode(Panel &P, int64_t n, DevBuf<int> &code, DevBuf<int64_t> &levels, int64_t &nl,
                  const uint8_t *keep) {
    hipStream_t s = P.s;
    DevBuf<int> present, newcode;
    DEV_ALLOC(present, nl + 1, s);
    DEV_ALLOC(newcode, nl + 1, s);
    HIP_TRY(hipMemsetAsync(present.p, 0, (nl + 1) * 4, s));
    hipLaunchKernelGGL(k_level_present, dim3(grid1(n)), dim3(T256), 0, s, n, code.p, keep, present.p);
    HIP_TRY(exclusive_sum(P.tmp, present.p, newcode.p, nl + 1, s));
    int cnt = 0;
    HIP_TRY(scalar_to_host(&cnt, newcode.p + nl, s));
    DevBuf<int64_t> lv;
    DEV_ALLOC(lv, cnt, s);
    hipLaunchKernelGGL(k_compact_levels, dim3(grid1(nl)), dim3(T256), 0, s, nl, present.p,
                       newcode.p, levels.p, lv.p);
    hipLaunchKernelGGL(k_recode, dim3(grid1(n)), dim3(T256), 0, s, n, code.p, newcode.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    levels = std::move(lv);
    nl = cnt;
    return HDFE_STATUS_OK;
}

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace hdfe

#include <chrono>

using namespace hdfe;

extern "C" {

void hdfe_destroy(void *h);  // pcg.hip

// Passes 0 and 1 on the device.  fe: nfe int64 id columns (device), vars:
// m float64 columns (device, y first), weights: NULL or float32/float64
// (device).  stream_dim: -1 to choose the dimension with the most levels.
// sort_vars: also order rows by the variables (the canonical row order of a
// fit whose intermediates are kept, core/passes.py:340-368).
// Pass 1 over the panel's sorted rows (V, w): means, cells, the weighted
// cell sums, obs/cnt, identifying groups and cells (core/passes.py:397-589).
// Run at creation and again by hdfe_panel_reweight.
static int run_pass1(Panel &P) {
    hipStream_t s = P.s;
    const int64_t nk = P.n;
    const int m = P.m;
    const int D = P.D;
    // pass 1: means, cells, groups
    const int nb = 512;
    DevBuf<double> part;
    DEV_ALLOC(part, (int64_t)nb * std::max(m * m, m) + 64, s);
    std::vector<double> hp;
    DEV_ALLOC(P.means, m, s);
    DEV_ALLOC(P.raw_ss, m, s);
#define HDFE_CS(MM) hipLaunchKernelGGL(k_colsum_rows<MM>, dim3(nb), dim3(T256), 0, s, nk, m, P.V.p, part.p)
    switch (m) {
        case 1: HDFE_CS(1); break;
        case 2: HDFE_CS(2); break;
        case 3: HDFE_CS(3); break;
        case 4: HDFE_CS(4); break;
        default: HDFE_CS(0); break;
    }
#undef HDFE_CS
    HIP_TRY(to_host(hp, part.p, (int64_t)nb * m, s));
    std::vector<double> mean(m, 0.0), ss(m, 0.0);
    for (int b = 0; b < nb; ++b)
        for (int j = 0; j < m; ++j) mean[j] += hp[(size_t)b * m + j];
    for (int j = 0; j < m; ++j) mean[j] /= (double)nk;
    HIP_TRY(hipMemcpyAsync(P.means.p, mean.data(), m * 8, hipMemcpyHostToDevice, s));
#define HDFE_CS(MM) \
    hipLaunchKernelGGL(k_colss_rows<MM>, dim3(nb), dim3(T256), 0, s, nk, m, P.V.p, P.means.p, part.p)
    switch (m) {
        case 1: HDFE_CS(1); break;
        case 2: HDFE_CS(2); break;
        case 3: HDFE_CS(3); break;
        case 4: HDFE_CS(4); break;
        default: HDFE_CS(0); break;
    }
#undef HDFE_CS
    HIP_TRY(to_host(hp, part.p, (int64_t)nb * m, s));
    for (int b = 0; b < nb; ++b)
        for (int j = 0; j < m; ++j) ss[j] += hp[(size_t)b * m + j];
    HIP_TRY(hipMemcpyAsync(P.raw_ss.p, ss.data(), m * 8, hipMemcpyHostToDevice, s));
    DevBuf<int> gh, ch, grank, crank;
    DEV_ALLOC(gh, nk, s);
    DEV_ALLOC(ch, nk, s);
    DEV_ALLOC(grank, nk, s);
    DEV_ALLOC(crank, nk, s);
    hipLaunchKernelGGL(k_row_heads, dim3(grid1(nk)), dim3(T256), 0, s, nk, D, P.g.p, P.codes.p,
                       nk, gh.p, ch.p);
    HIP_TRY(inclusive_sum(P.tmp, gh.p, grank.p, nk, s));
    HIP_TRY(inclusive_sum(P.tmp, ch.p, crank.p, nk, s));
    int ng = 0, nc = 0;
    HIP_TRY(scalar_to_host(&ng, grank.p + nk - 1, s));
    HIP_TRY(scalar_to_host(&nc, crank.p + nk - 1, s));
    ARG_CHECK(ng == P.G, "internal: group count mismatch after sorting");
    P.C = nc;
    DEV_ALLOC(P.gstart, ng + 1, s);
    DEV_ALLOC(P.cstart, nc + 1, s);
    hipLaunchKernelGGL(k_starts, dim3(grid1(nk)), dim3(T256), 0, s, nk, gh.p, grank.p, P.gstart.p);
    DEV_ALLOC(P.cgstart, ng + 1, s);
    if (nc == nk) {
        // every row its own cell: the cells' starts are the rows, and the
        // groups' starts over cells are their starts over rows
        hipLaunchKernelGGL(k_iota, dim3(grid1(nk + 1)), dim3(T256), 0, s, nk + 1, P.cstart.p);
        HIP_TRY(hipMemcpyAsync(P.cgstart.p, P.gstart.p, (size_t)(ng + 1) * 4,
                               hipMemcpyDeviceToDevice, s));
    } else {
        hipLaunchKernelGGL(k_starts, dim3(grid1(nk)), dim3(T256), 0, s, nk, ch.p, crank.p,
                           P.cstart.p);
        // group starts over cells: the cell rank of each group's first row
        DevBuf<int> cghead, cgrank;
        DEV_ALLOC(cghead, nc, s);
        DEV_ALLOC(cgrank, nc, s);
        // a cell starts a group when its first row does
        hipLaunchKernelGGL(k_gather<int>, dim3(grid1(nc)), dim3(T256), 0, s, (int64_t)nc,
                           P.cstart.p, gh.p, cghead.p);
        HIP_TRY(inclusive_sum(P.tmp, cghead.p, cgrank.p, nc, s));
        hipLaunchKernelGGL(k_starts, dim3(grid1(nc)), dim3(T256), 0, s, (int64_t)nc, cghead.p,
                           cgrank.p, P.cgstart.p);
    }
    gh.reset();
    ch.reset();
    grank.reset();
    crank.reset();
    DevBuf<double> cn, crun, csums, wpart;
    DEV_ALLOC(cn, nc, s);
    DEV_ALLOC(crun, nc, s);
    DEV_ALLOC(csums, (int64_t)nc * m, s);
    const int mm = m <= 8 ? m * m : 0;
    DEV_ALLOC(wpart, (int64_t)nb * std::max(mm, 1), s);
    {
        const double *wp = P.weighted ? P.w.p : (const double *)nullptr;
        const bool one = nc == nk && !P.weighted;
#define HDFE_CELLS(MM)                                                                          \
    case MM:                                                                                    \
        if (one)                                                                                \
            hipLaunchKernelGGL((k_cells_m<MM, true>), dim3(nb), dim3(T256), 0, s, (int64_t)nc,  \
                               P.cstart.p, P.V.p, wp, P.means.p, cn.p, crun.p, csums.p,         \
                               wpart.p);                                                        \
        else                                                                                    \
            hipLaunchKernelGGL((k_cells_m<MM, false>), dim3(nb), dim3(T256), 0, s, (int64_t)nc, \
                               P.cstart.p, P.V.p, wp, P.means.p, cn.p, crun.p, csums.p,         \
                               wpart.p);                                                        \
        break;
        switch (m) {
            HDFE_CELLS(1)
            HDFE_CELLS(2)
            HDFE_CELLS(3)
            HDFE_CELLS(4)
            HDFE_CELLS(5)
            HDFE_CELLS(6)
            HDFE_CELLS(7)
            HDFE_CELLS(8)
        default:
            hipLaunchKernelGGL(k_cells, dim3(nb), dim3(T256), 0, s, (int64_t)nc, m, P.cstart.p,
                               P.V.p, wp, P.means.p, cn.p, crun.p, csums.p, wpart.p);
        }
#undef HDFE_CELLS
    }
    DEV_ALLOC(P.W, std::max(m * m, 1), s);
    if (mm) {
        HIP_TRY(to_host(hp, wpart.p, (int64_t)nb * mm, s));
        std::vector<double> Wh(mm, 0.0);
        for (int b = 0; b < nb; ++b)
            for (int t = 0; t < mm; ++t) Wh[t] += hp[(size_t)b * mm + t];
        HIP_TRY(hipMemcpyAsync(P.W.p, Wh.data(), mm * 8, hipMemcpyHostToDevice, s));
    } else {
        // more than eight variables: the scatter by 8 x 8 tiles
        const double *wp = P.weighted ? P.w.p : (const double *)nullptr;
        DevBuf<double> tpart;
        DEV_ALLOC(tpart, (int64_t)nb * 64, s);
        std::vector<double> Wh((size_t)m * m, 0.0), th;
        const int nt = (m + 7) / 8;
        for (int ta = 0; ta < nt; ++ta)
            for (int tb = ta; tb < nt; ++tb) {
                const int ja = 8 * ta, jb = 8 * tb;
                const int na = std::min(8, m - ja), nbc = std::min(8, m - jb);
                hipLaunchKernelGGL(k_cells_tile, dim3(nb), dim3(T256), 0, s, (int64_t)nc, m, ja, na,
                                   jb, nbc, P.cstart.p, P.V.p, wp, P.means.p, cn.p, csums.p,
                                   tpart.p);
                HIP_TRY(hipGetLastError());
                HIP_TRY(to_host(th, tpart.p, (int64_t)nb * 64, s));
                for (int a = 0; a < na; ++a)
                    for (int b = 0; b < nbc; ++b) {
                        double v = 0.0;
                        for (int q = 0; q < nb; ++q) v += th[(size_t)q * 64 + a * 8 + b];
                        Wh[(size_t)(ja + a) * m + jb + b] = v;
                        Wh[(size_t)(jb + b) * m + ja + a] = v;
                    }
            }
        HIP_TRY(hipMemcpyAsync(P.W.p, Wh.data(), (size_t)m * m * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    const int64_t L = P.L;
    DEV_ALLOC(P.obs, std::max<int64_t>(L, 1), s);
    DEV_ALLOC(P.cnt, std::max<int64_t>(L, 1), s);
    {
        DevBuf<unsigned long long> ic;
        DEV_ALLOC(ic, std::max<int64_t>(L, 1), s);
        HIP_TRY(hipMemsetAsync(ic.p, 0, std::max<int64_t>(L, 1) * 8, s));
        HIP_TRY(hipMemsetAsync(P.obs.p, 0, std::max<int64_t>(L, 1) * 8, s));
        for (int t = 0; t < D; ++t) {
            const int nl = (int)P.nlev[t];
            const int *ct = P.codes.p + (int64_t)t * nk;
            const bool one = nc == nk && !P.weighted;
#define HDFE_OCS(NL, ONE)                                                                       \
    hipLaunchKernelGGL((k_obs_cnt_small<NL, ONE>), dim3(grid1(nc, T256, 1024)), dim3(T256), 0, s, \
                       (int64_t)nc, P.cstart.p, ct, P.offs_h[t], nl, cn.p, ic.p, P.obs.p,        \
                       P.weighted)
            if (nl <= 8 && one)
                HDFE_OCS(8, true);
            else if (nl <= 8)
                HDFE_OCS(8, false);
            else if (nl <= 16 && one)
                HDFE_OCS(16, true);
            else if (nl <= 16)
                HDFE_OCS(16, false);
#undef HDFE_OCS
#define HDFE_OC(ONE, LDSB)                                                                   \
    hipLaunchKernelGGL((k_obs_cnt<ONE, LDSB>), dim3(grid1(nc, T256, 2048)), dim3(T256), 0, s, \
                       (int64_t)nc, P.cstart.p, ct, P.offs_h[t], nl, cn.p, ic.p, P.obs.p,     \
                       P.weighted)
            else if (nl <= kLdsBins / 2 && one)
                HDFE_OC(true, true);
            else if (nl <= kLdsBins / 2)
                HDFE_OC(false, true);
            else if (one)
                HDFE_OC(true, false);
            else
                HDFE_OC(false, false);
#undef HDFE_OC
        }
        hipLaunchKernelGGL(k_u64_to_f64, dim3(grid1(L)), dim3(T256), 0, s, L, ic.p, P.cnt.p,
                           P.weighted ? (double *)nullptr : P.obs.p);
        HIP_TRY(hipStreamSynchronize(s));
    }
    // identifying groups and their cells
    DevBuf<int> ident, gpos, gcells, cpos;
    DEV_ALLOC(ident, ng + 1, s);
    DEV_ALLOC(gpos, ng + 1, s);
    DEV_ALLOC(gcells, ng + 1, s);
    DEV_ALLOC(cpos, ng + 1, s);
    HIP_TRY(hipMemsetAsync(ident.p, 0, (ng + 1) * 4, s));
    hipLaunchKernelGGL(k_group_cells, dim3(grid1(ng)), dim3(T256), 0, s, (int64_t)ng,
                       P.cgstart.p, ident.p);
    HIP_TRY(exclusive_sum(P.tmp, ident.p, gpos.p, ng + 1, s));
    int Gi = 0;
    HIP_TRY(scalar_to_host(&Gi, gpos.p + ng, s));
    P.Gi = Gi;
    {
        // cells of identifying groups, as an exclusive sum over groups
        hipLaunchKernelGGL(k_ident_ncells, dim3(grid1(ng + 1)), dim3(T256), 0, s, (int64_t)ng,
                           P.cgstart.p, ident.p, gcells.p);
        HIP_TRY(exclusive_sum(P.tmp, gcells.p, cpos.p, ng + 1, s));
        int Mi = 0;
        HIP_TRY(scalar_to_host(&Mi, cpos.p + ng, s));
        P.M = Mi;
    }
    const int64_t M = P.M;
    DEV_ALLOC(P.icodes, std::max<int64_t>(1, (int64_t)D * M), s);
    DEV_ALLOC(P.igstart, Gi + 1, s);
    DEV_ALLOC(P.igroup, std::max(Gi, 1), s);
    if (Gi == ng && M == nc) {
        // every group identifies: the identifying cells are the cells, whose
        // weights and sums become the panel's as they are; only the codes
        // at the cells' first rows are gathered
        P.in = std::move(cn);
        P.isums = std::move(csums);
        hipLaunchKernelGGL(k_cell_codes, dim3(grid1(M)), dim3(T256), 0, s, M, D, P.cstart.p,
                           P.codes.p, nk, P.icodes.p);
        HIP_TRY(hipMemcpyAsync(P.igstart.p, P.cgstart.p, (size_t)(ng + 1) * 4,
                               hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_iota, dim3(grid1(Gi)), dim3(T256), 0, s, (int64_t)Gi, P.igroup.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));
        return HDFE_STATUS_OK;
    }
    DEV_ALLOC(P.in, std::max<int64_t>(M, 1), s);
    DEV_ALLOC(P.isums, std::max<int64_t>(M * m, 1), s);
#define HDFE_L(W)                                                                                \
    hipLaunchKernelGGL(k_ident_cells<W>, dim3(grid1((int64_t)ng * W)), dim3(T256), 0, s,        \
                       (int64_t)ng, D, m, P.cgstart.p, ident.p, gpos.p, cpos.p, P.cstart.p,     \
                       P.codes.p, nk, cn.p, csums.p, P.icodes.p, M, P.in.p, P.isums.p,         \
                       P.igstart.p, P.igroup.p)
    HDFE_BY_WIDTH(run_width(P.C, ng), HDFE_L)
#undef HDFE_L
    HIP_TRY(hipMemcpyAsync(P.igstart.p + Gi, cpos.p + ng, 4, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return HDFE_STATUS_OK;
}

int hdfe_panel_create(int64_t n_rows, int nfe, const int64_t *const *fe, int m,
                      const double *const *vars, const void *weights, int weights_f32,
                      int fixef_rm_singleton, int stream_dim, int sort_vars, void **out) {
    ARG_CHECK(out && fe && vars, "hdfe_panel_create: null argument");
    ARG_CHECK(n_rows > 0 && n_rows < INT32_MAX, "hdfe_panel_create: rows must be in [1, 2^31)");
    ARG_CHECK(nfe >= 1 && nfe <= 8, "hdfe_panel_create: 1 to 8 fixed effects");
    ARG_CHECK(m >= 1 && m <= 64, "hdfe_panel_create: 1 to 64 variables");
    std::unique_ptr<Panel> P(new Panel());
    const auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipGetDevice(&P->dev));
    HIP_TRY(stream_get(&P->s));
    hipStream_t s = P->s;
    const int64_t n = n_rows;
    P->n_in = n;
    P->nfe = nfe;
    P->m = m;
    P->weighted = weights != nullptr;
    // pass 0: codes for every dimension
    std::vector<DevBuf<int>> code(nfe);
    std::vector<DevBuf<int64_t>> lv(nfe);
    std::vector<int64_t> nl(nfe);
    for (int d = 0; d < nfe; ++d) TRY(factorize(*P, fe[d], n, code[d], lv[d], nl[d]));
    DevBuf<uint8_t> keep;
    DEV_ALLOC(keep, n, s);
    HIP_TRY(hipMemsetAsync(keep.p, 1, n, s));
    if (fixef_rm_singleton) {
        // dimensions in turn until each has been checked since the last drop
        DevBuf<int> cnt;
        int64_t maxl = *std::max_element(nl.begin(), nl.end());
        DEV_ALLOC(cnt, maxl, s);
        DevBuf<unsigned long long> dropped;
        DEV_ALLOC(dropped, 1, s);
        bool found = false;
        int quiet = 0;
        int64_t i = 0;
        std::vector<int64_t> rounds;
        while (quiet < nfe - (found ? 1 : 0)) {
            const int d = (int)(i % nfe);
            HIP_TRY(hipMemsetAsync(cnt.p, 0, nl[d] * 4, s));
            HIP_TRY(hipMemsetAsync(dropped.p, 0, 8, s));
            hipLaunchKernelGGL(k_level_count, dim3(grid1(n, T256, 2048)), dim3(T256), 0, s, n,
                               code[d].p, keep.p, cnt.p, (int)nl[d]);
            hipLaunchKernelGGL(k_drop_single, dim3(grid1(n)), dim3(T256), 0, s, n, code[d].p,
                               cnt.p, keep.p, dropped.p);
            unsigned long long dr = 0;
            HIP_TRY(scalar_to_host(&dr, dropped.p, s));
            if (dr) {
                found = true;
                quiet = 0;
                P->singletons += (int64_t)dr;
                const int64_t rd = i / nfe;
                if (rounds.empty() || rounds.back() != rd) rounds.push_back(rd);
            } else {
                ++quiet;
            }
            ++i;
        }
        P->single_rounds = (int64_t)rounds.size();
        if (P->singletons)
            for (int d = 0; d < nfe; ++d) TRY(recode(*P, n, code[d], lv[d], nl[d], keep.p));
    }
    P->n = n - P->singletons;
    ARG_CHECK(P->n > 0, "no complete observations left after dropping singletons");
    // the streamed dimension: the most levels, ties to the first listed
    int gd = stream_dim;
    if (gd < 0) {
        gd = 0;
        for (int d = 1; d < nfe; ++d)
            if (nl[d] > nl[gd]) gd = d;
    }
    ARG_CHECK(gd < nfe, "hdfe_panel_create: stream_dim out of range");
    P->gdim = gd;
    for (int d = 0; d < nfe; ++d)
        if (d != gd) P->dims.push_back(d);
    const int D = (int)P->dims.size();
    P->D = D;
    P->G = nl[gd];
    int64_t o = 0;
    for (int t = 0; t < D; ++t) {
        P->nlev.push_back(nl[P->dims[t]]);
        P->offs_h.push_back(o);
        o += nl[P->dims[t]];
    }
    P->L = o;
    DEV_ALLOC(P->offs, std::max(D, 1), s);
    if (D) HIP_TRY(hipMemcpyAsync(P->offs.p, P->offs_h.data(), D * 8, hipMemcpyHostToDevice, s));
    // row order: (group, codes) packed, optionally after stable sorts by the
    // variables (least significant first)
    const int gbits = std::max(1, bits_for((uint64_t)P->G));
    std::vector<int> shift(D);
    int tot = 0;
    for (int t = D - 1; t >= 0; --t) {
        shift[t] = tot;
        tot += std::max(1, bits_for((uint64_t)P->nlev[t]));
    }
    ARG_CHECK(tot + gbits <= 63, "hdfe_panel_create: level codes need more than 63 key bits");
    DevBuf<int> codes_in, shift_d;
    DEV_ALLOC(codes_in, std::max<int64_t>(1, (int64_t)D * n), s);
    DEV_ALLOC(shift_d, std::max(D, 1), s);
    for (int t = 0; t < D; ++t)
        HIP_TRY(hipMemcpyAsync(codes_in.p + t * n, code[P->dims[t]].p, n * 4,
                               hipMemcpyDeviceToDevice, s));
    if (D) HIP_TRY(hipMemcpyAsync(shift_d.p, shift.data(), D * 4, hipMemcpyHostToDevice, s));
    DevBuf<uint64_t> full, k1, k2;
    DevBuf<int> i1, i2;
    DEV_ALLOC(full, n, s);
    DEV_ALLOC(k1, n, s);
    DEV_ALLOC(k2, n, s);
    DEV_ALLOC(i1, n, s);
    DEV_ALLOC(i2, n, s);
    hipLaunchKernelGGL(k_pack_key, dim3(grid1(n)), dim3(T256), 0, s, n, D, code[gd].p, codes_in.p,
                       n, shift_d.p, tot, keep.p, full.p, i1.p);
    // rows stored in group order (a panel kept by worker) need only each
    // group's rows ranked by their codes: the order the radix sort gives
    bool rows_sorted = false;
    if (!sort_vars && P->singletons == 0 && P->G > 0 && n < INT32_MAX) {
        DevBuf<int> gst, unsorted;
        DEV_ALLOC(gst, P->G + 1, s);
        DEV_ALLOC(unsorted, 1, s);
        HIP_TRY(hipMemsetAsync(unsorted.p, 0, 4, s));
        hipLaunchKernelGGL(k_group_order, dim3(grid1(n)), dim3(T256), 0, s, n, P->G, code[gd].p,
                           gst.p, unsorted.p);
        int un = 0;
        HIP_TRY(scalar_to_host(&un, unsorted.p, s));
        if (!un) {
            int W = 4;
            while (W < 64 && (int64_t)W * P->G < n) W *= 2;
            TRY(seg_rank_sort(s, P->G, 1, gst.p, full.p, i1.p, k2.p, i2.p, W, rows_sorted));
        }
    }
    DevBuf<double> Vin;
    DevBuf<const double *> colp;
    DEV_ALLOC(colp, m, s);
    HIP_TRY(hipMemcpyAsync(colp.p, vars, m * sizeof(double *), hipMemcpyHostToDevice, s));
    if (sort_vars) {
        DEV_ALLOC(Vin, n * m, s);
        DevBuf<int> ident;
        DEV_ALLOC(ident, n, s);
        // V in input order, row major, for the variable keys
        hipLaunchKernelGGL(k_iota, dim3(grid1(n)), dim3(T256), 0, s, n, ident.p);
        hipLaunchKernelGGL(k_gather_vars, dim3(grid1(n * m)), dim3(T256), 0, s, n, m, colp.p,
                           ident.p, Vin.p);
        bool first = true;
        for (int j = m - 1; j >= 0; --j) {
            hipLaunchKernelGGL(k_var_key, dim3(grid1(n)), dim3(T256), 0, s, n, Vin.p, m, j,
                               first ? (const int *)nullptr : (const int *)i1.p, k1.p, i2.p);
            HIP_TRY(sort_pairs(P->tmp, k1.p, k2.p, i2.p, i1.p, n, 0, 64, s));
            first = false;
        }
        hipLaunchKernelGGL(k_gather_key, dim3(grid1(n)), dim3(T256), 0, s, n, full.p, i1.p, k1.p);
    } else if (!rows_sorted) {
        HIP_TRY(hipMemcpyAsync(k1.p, full.p, n * 8, hipMemcpyDeviceToDevice, s));
    }
    // kept keys are below 2^(tot + gbits); a dropped row's key has that bit set
    if (!rows_sorted)
        HIP_TRY(sort_pairs(P->tmp, k1.p, k2.p, i1.p, i2.p, n, 0, tot + gbits + 1, s));
    const int64_t nk = P->n;
    P->perm = std::move(i2);
    // gather the sorted columns (kept rows come first)
    DEV_ALLOC(P->g, nk, s);
    hipLaunchKernelGGL(k_gather<int>, dim3(grid1(nk)), dim3(T256), 0, s, nk, P->perm.p,
                       code[gd].p, P->g.p);
    DEV_ALLOC(P->codes, std::max<int64_t>(1, (int64_t)D * nk), s);
    for (int t = 0; t < D; ++t)
        hipLaunchKernelGGL(k_gather<int>, dim3(grid1(nk)), dim3(T256), 0, s, nk, P->perm.p,
                           code[P->dims[t]].p, P->codes.p + t * nk);
    DEV_ALLOC(P->V, nk * m, s);
    hipLaunchKernelGGL(k_gather_vars, dim3(grid1(nk * m)), dim3(T256), 0, s, nk, m, colp.p,
                       P->perm.p, P->V.p);
    if (weights) {
        DEV_ALLOC(P->w, nk, s);
        hipLaunchKernelGGL(k_gather_weights, dim3(grid1(nk)), dim3(T256), 0, s, nk, weights,
                           weights_f32, P->perm.p, P->w.p);
    }
    P->glevels = std::move(lv[gd]);
    for (int t = 0; t < D; ++t) P->levels.push_back(std::move(lv[P->dims[t]]));
    HIP_TRY(hipGetLastError());
    // free pass-0 scratch
    code.clear();
    full.reset();
    k1.reset();
    k2.reset();
    i1.reset();
    Vin.reset();
    keep.reset();
    codes_in.reset();

    TRY(run_pass1(*P));
    P->t_build = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *out = P.release();
    return HDFE_STATUS_OK;
}

// Integer facts of the panel: [n_in, n, G, C, M, Gi, L, D, gdim, singletons,
// single_rounds, E, nnz, ncomp, m, weighted].
int hdfe_panel_info(void *h, int64_t *info, int64_t *nlev, int *dims) {
    ARG_CHECK(h && info, "hdfe_panel_info: null argument");
    Panel &P = *(Panel *)h;
    int64_t v[16] = {P.n_in, P.n, P.G, P.C, P.M, P.Gi, P.L, P.D, P.gdim, P.singletons,
                     P.single_rounds, P.E, P.nnz, P.ncomp, P.m, P.weighted};
    for (int i = 0; i < 16; ++i) info[i] = v[i];
    for (int t = 0; t < P.D; ++t) {
        if (nlev) nlev[t] = P.nlev[t];
        if (dims) dims[t] = P.dims[t];
    }
    return HDFE_STATUS_OK;
}

// Device pointers of the panel's arrays, by name.
// A copy of `bytes` of device memory at src (a panel array) into page-locked
// host memory at dst, on a stream of its own, after the work queued on the
// panel's stream so far: it overlaps the panel's later kernels.
int hdfe_panel_copy_async(void *h, const void *src, void *dst, int64_t bytes) {
    ARG_CHECK(h && src && dst && bytes >= 0, "hdfe_panel_copy_async: bad argument");
    Panel &P = *(Panel *)h;
    if (!P.side) {
        HIP_TRY(stream_get(&P.side));
        HIP_TRY(hipEventCreateWithFlags(&P.side_ev, hipEventDisableTiming));
    }
    HIP_TRY(hipEventRecord(P.side_ev, P.s));
    HIP_TRY(hipStreamWaitEvent(P.side, P.side_ev, 0));
    if (bytes) HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, P.side));
    return HDFE_STATUS_OK;
}

// Waits for the panel's copies to the host.
int hdfe_panel_copy_wait(void *h) {
    ARG_CHECK(h, "hdfe_panel_copy_wait: null handle");
    Panel &P = *(Panel *)h;
    if (P.side) HIP_TRY(hipStreamSynchronize(P.side));
    return HDFE_STATUS_OK;
}

// The explicit S build's budget in bytes (-1: none), and the estimate of the
// last S that was not built for it (0: built).
int hdfe_panel_budget(void *h, int64_t set, int64_t *budget, int64_t *skipped) {
    ARG_CHECK(h, "hdfe_panel_budget: null handle");
    Panel &P = *(Panel *)h;
    if (set >= -1) P.s_budget = set;
    if (budget) *budget = P.s_budget;
    if (skipped) *skipped = P.s_skipped;
    return HDFE_STATUS_OK;
}

int hdfe_panel_array(void *h, const char *name, void **ptr, int64_t *count) {
    ARG_CHECK(h && name && ptr, "hdfe_panel_array: null argument");
    Panel &P = *(Panel *)h;
    std::string nm(name);
    void *p = nullptr;
    int64_t c = 0;
    if (nm == "g") { p = P.g.p; c = P.n; }
    else if (nm == "codes") { p = P.codes.p; c = P.D * P.n; }
    else if (nm == "V") { p = P.V.p; c = P.n * P.m; }
    else if (nm == "w") { p = P.w.p; c = P.weighted ? P.n : 0; }
    else if (nm == "perm") { p = P.perm.p; c = P.n; }
    else if (nm == "gstart") { p = P.gstart.p; c = P.G + 1; }
    else if (nm == "cstart") { p = P.cstart.p; c = P.C + 1; }
    else if (nm == "glevels") { p = P.glevels.p; c = P.G; }
    else if (nm == "means") { p = P.means.p; c = P.m; }
    else if (nm == "raw_ss") { p = P.raw_ss.p; c = P.m; }
    else if (nm == "W") { p = P.W.p; c = (int64_t)P.m * P.m; }
    else if (nm == "obs") { p = P.obs.p; c = P.L; }
    else if (nm == "cnt") { p = P.cnt.p; c = P.L; }
    else if (nm == "icodes") { p = P.icodes.p; c = P.D * P.M; }
    else if (nm == "in") { p = P.in.p; c = P.M; }
    else if (nm == "isums") { p = P.isums.p; c = P.M * P.m; }
    else if (nm == "igstart") { p = P.igstart.p; c = P.Gi + 1; }
    else if (nm == "igroup") { p = P.igroup.p; c = P.Gi; }
    else if (nm == "diag") { p = P.diag.p; c = P.diag.p ? P.L : 0; }
    else if (nm == "rhs") { p = P.rhs.p; c = P.rhs.p ? P.L * P.m : 0; }
    else if (nm == "comp") { p = P.comp.p; c = P.comp.p ? (P.D ? P.nlev[0] : 0) : 0; }
    else if (nm == "offs") { p = P.offs.p; c = P.D; }
    else if (nm.rfind("levels", 0) == 0) {
        const int t = std::atoi(nm.c_str() + 6);
        ARG_CHECK(t >= 0 && t < P.D, "hdfe_panel_array: no such levels array");
        p = P.levels[t].p;
        c = P.nlev[t];
    } else {
        set_error("hdfe_panel_array: unknown array " + nm);
        return HDFE_STATUS_ARG;
    }
    *ptr = p;
    if (count) *count = c;
    return HDFE_STATUS_OK;
}

// Step 2's system: incidences, diag, rhs, the explicit S and its solver
// handle (*matrix: owned by the panel; do not destroy).  with_rhs: build the
// right-hand side of the panel's own variables.
int hdfe_panel_system(void *h, int with_rhs, void **matrix, double *seconds) {
    ARG_CHECK(h, "hdfe_panel_system: null handle");
    Panel &P = *(Panel *)h;
    hipStream_t s = P.s;
    const auto t0 = std::chrono::steady_clock::now();
    const int D = P.D, m = P.m;
    const int64_t L = P.L, M = P.M, Gi = P.Gi;
    ARG_CHECK(D >= 1 && L >= 1, "hdfe_panel_system: no dimension besides the streamed one");
    // dense / sparse dimensions and their levels (ascending global index)
    std::vector<uint8_t> sparse_dim(D);
    {
        // smallest dimensions first while the dense levels fit kDenseMax
        std::vector<int> order(D);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return P.nlev[a] < P.nlev[b]; });
        int64_t ld = 0;
        for (int t = 0; t < D; ++t) sparse_dim[t] = 1;
        for (int t : order) {
            if (P.nlev[t] <= kDenseMax && ld + P.nlev[t] <= kDenseMax && D > 1) {
                sparse_dim[t] = 0;
                ld += P.nlev[t];
            }
        }
        // keep at least one sparse dimension
        bool any_sparse = false;
        for (int t = 0; t < D; ++t) any_sparse |= (bool)sparse_dim[t];
        if (!any_sparse) sparse_dim[order.back()] = 1;
    }
    LevelDims ldims{};
    ldims.D = D;
    int64_t Ls = 0, Ld64 = 0;
    for (int t = 0; t < D; ++t) {
        ldims.off[t] = P.offs_h[t];
        ldims.sparse[t] = sparse_dim[t];
        ldims.base[t] = sparse_dim[t] ? Ls : Ld64;
        (sparse_dim[t] ? Ls : Ld64) += P.nlev[t];
    }
    const int Ld = (int)Ld64;
    int ncross = 0;
    for (int d1 = 0; d1 < D; ++d1)
        for (int d2 = 0; d2 < D; ++d2) ncross += (d1 != d2 && sparse_dim[d1] && sparse_dim[d2]);
    DevBuf<uint8_t> d_sparse_lv, d_sparse_dim;
    DevBuf<int> d_sparse_idx, d_dense_idx, d_lvl_dim, d_sparse_levels, d_dense_levels;
    DEV_ALLOC(d_sparse_lv, L, s);
    DEV_ALLOC(d_sparse_dim, D, s);
    DEV_ALLOC(d_sparse_idx, L, s);
    DEV_ALLOC(d_dense_idx, L, s);
    DEV_ALLOC(d_lvl_dim, L, s);
    DEV_ALLOC(d_sparse_levels, std::max<int64_t>(Ls, 1), s);
    DEV_ALLOC(d_dense_levels, std::max(Ld, 1), s);
    // (the dimensions' flags written by the kernel from its arguments: a
    // pageable copy here would wait behind the level values' copy to the host)
    hipLaunchKernelGGL(k_level_tables, dim3(grid1(L)), dim3(T256), 0, s, (int64_t)L, ldims,
                       d_sparse_dim.p, d_sparse_lv.p, d_sparse_idx.p, d_dense_idx.p, d_lvl_dim.p,
                       d_sparse_levels.p, d_dense_levels.p);

    // incidences (group, level) of the identifying cells
    const int lbits = std::max(1, bits_for((uint64_t)L));
    const int gbits = std::max(1, bits_for((uint64_t)std::max<int64_t>(Gi, 1)));
    ARG_CHECK(lbits + gbits <= 64, "hdfe_panel_system: incidence keys need more than 64 bits");
    DimList sdl{}, ddl{};
    for (int t = 0; t < D; ++t) {
        if (sparse_dim[t]) sdl.d[sdl.n++] = t;
        else ddl.d[ddl.n++] = t;
    }
    const int Ds = sdl.n;
    const int64_t nE = M * Ds;
    DEV_ALLOC(P.cgroup, std::max<int64_t>(M, 1), s);
    const int cw = run_width(M, Gi);
#define HDFE_L(W)                                                                       \
    hipLaunchKernelGGL(k_cell_group<W>, dim3(grid1(Gi * W)), dim3(T256), 0, s, Gi,      \
                       P.igstart.p, P.cgroup.p)
    HDFE_BY_WIDTH(cw, HDFE_L)
#undef HDFE_L
    DevBuf<uint64_t> k1, k2;
    DevBuf<int> v1, v2, head, rank;
    DEV_ALLOC(k1, std::max<int64_t>(nE, 1), s);
    DEV_ALLOC(k2, std::max<int64_t>(nE, 1), s);
    DEV_ALLOC(v1, std::max<int64_t>(nE, 1), s);
    DEV_ALLOC(v2, std::max<int64_t>(nE, 1), s);
    DEV_ALLOC(head, std::max<int64_t>(nE, 1), s);
    DEV_ALLOC(rank, std::max<int64_t>(nE, 1), s);
    int E = 0;
    if (nE) {
        hipLaunchKernelGGL(k_inc_keys_s, dim3(grid1(nE)), dim3(T256), 0, s, M, sdl, P.cgroup.p,
                           P.icodes.p, M, P.offs.p, lbits, k1.p, v1.p);
        // within-group ranks when every group's segment fits a few lanes,
        // else the radix sort (the same order either way)
        int W = 4;
        while (W < 64 && W < Ds * cw) W *= 2;
        bool sorted = false;
        TRY(seg_rank_sort(s, Gi, Ds, P.igstart.p, k1.p, v1.p, k2.p, v2.p, W, sorted));
        if (!sorted) HIP_TRY(sort_pairs(P.tmp, k1.p, k2.p, v1.p, v2.p, nE, 0, lbits + gbits, s));
        hipLaunchKernelGGL(k_heads_u64, dim3(grid1(nE)), dim3(T256), 0, s, nE, k2.p, head.p);
        HIP_TRY(inclusive_sum(P.tmp, head.p, rank.p, nE, s));
        HIP_TRY(scalar_to_host(&E, rank.p + nE - 1, s));
    }
    P.E = E;
    DevBuf<int> istart;
    DEV_ALLOC(istart, E + 1, s);
    if (nE)
        hipLaunchKernelGGL(k_starts, dim3(grid1(nE)), dim3(T256), 0, s, nE, head.p, rank.p, istart.p);
    DEV_ALLOC(P.ig, std::max(E, 1), s);
    DEV_ALLOC(P.il, std::max(E, 1), s);
    DEV_ALLOC(P.iu, std::max(E, 1), s);
    DEV_ALLOC(P.is, std::max<int64_t>((int64_t)E * m, 1), s);
    DEV_ALLOC(P.Ng, std::max<int64_t>(Gi, 1), s);
    DEV_ALLOC(P.vbar, std::max<int64_t>(Gi * m, 1), s);
    DEV_ALLOC(P.gis, Gi + 1, s);
#define HDFE_IR(MM)                                                                             \
    hipLaunchKernelGGL(k_inc_reduce<MM>, dim3(grid1(E)), dim3(T256), 0, s, (int64_t)E, m, lbits, \
                       k2.p, v2.p, istart.p, P.in.p, P.isums.p, P.ig.p, P.il.p, P.iu.p, P.is.p)
    switch (m) {
        case 1: HDFE_IR(1); break;
        case 2: HDFE_IR(2); break;
        case 3: HDFE_IR(3); break;
        case 4: HDFE_IR(4); break;
        default: HDFE_IR(0); break;
    }
#undef HDFE_IR
#define HDFE_L(W)                                                                        \
    hipLaunchKernelGGL(k_group_totals<W>, dim3(grid1(Gi * W)), dim3(T256), 0, s, Gi, m,  \
                       P.igstart.p, P.in.p, P.isums.p, P.Ng.p, P.vbar.p)
    HDFE_BY_WIDTH(cw, HDFE_L)
#undef HDFE_L
    hipLaunchKernelGGL(k_group_inc_start, dim3(grid1(E)), dim3(T256), 0, s, (int64_t)E, P.ig.p,
                       P.gis.p, Gi);
    // level CSR of the incidences (sorted by level, then group)
    DEV_ALLOC(P.lptr, L + 1, s);
    DEV_ALLOC(P.lq, std::max(E, 1), s);
    if (E) {
        hipLaunchKernelGGL(k_level_keys, dim3(grid1(E)), dim3(T256), 0, s, (int64_t)E, P.ig.p,
                           P.il.p, gbits, k1.p, v1.p);
        // the incidences are in (group, level) order: a stable sort on the
        // level bits alone leaves each level's groups in order
        HIP_TRY(sort_pairs(P.tmp, k1.p, k2.p, v1.p, P.lq.p, E, gbits, lbits + gbits, s));
    }
    hipLaunchKernelGGL(k_level_ptr, dim3(grid1(L + 1)), dim3(T256), 0, s, L, (int64_t)E, k2.p,
                       gbits, P.lptr.p);
    DEV_ALLOC(P.diag, L, s);
    if (with_rhs) DEV_ALLOC(P.rhs, L * m, s);
    {
        DevBuf<int> nch, cptr, clvl, cbeg;
        DEV_ALLOC(nch, L + 1, s);
        DEV_ALLOC(cptr, L + 1, s);
        hipLaunchKernelGGL(k_level_nchunk, dim3(grid1(L + 1)), dim3(T256), 0, s, L, P.lptr.p, nch.p);
        HIP_TRY(exclusive_sum(P.tmp, nch.p, cptr.p, L + 1, s));
        int nchunk = 0;
        HIP_TRY(scalar_to_host(&nchunk, cptr.p + L, s));
        DEV_ALLOC(clvl, nchunk, s);
        DEV_ALLOC(cbeg, nchunk, s);
        DevBuf<double> part;
        DEV_ALLOC(part, (int64_t)nchunk * (m + 1), s);
        hipLaunchKernelGGL(k_level_chunks, dim3(grid1(L)), dim3(T256), 0, s, L, P.lptr.p, cptr.p,
                           clvl.p, cbeg.p);
        {
            DevBuf<double> contrib;
            DEV_ALLOC(contrib, std::max<int64_t>((int64_t)P.E * (m + 1), 1), s);
            hipLaunchKernelGGL(k_inc_contrib, dim3(grid1(P.E)), dim3(T256), 0, s, (int64_t)P.E, m,
                               P.ig.p, P.iu.p, P.is.p, P.Ng.p, P.vbar.p, contrib.p);
            hipLaunchKernelGGL(k_diag_rhs_part_c, dim3(grid1((int64_t)nchunk * 64)), dim3(T256), 0,
                               s, (int64_t)nchunk, m, clvl.p, cbeg.p, P.lptr.p, P.lq.p, contrib.p,
                               part.p);
        }
        hipLaunchKernelGGL(k_diag_rhs_sum, dim3(grid1(L * 64)), dim3(T256), 0, s, L, m, cptr.p, part.p,
                           P.diag.p, with_rhs ? P.rhs.p : (double *)nullptr);
        HIP_TRY(hipGetLastError());
    }
    // the dense levels (no incidences): weights per group, then their sums
    DevBuf<double> U;
    if (Ld) {
        DEV_ALLOC(U, std::max<int64_t>(Gi * Ld, 1), s);
        HIP_TRY(hipMemsetAsync(U.p, 0, std::max<int64_t>(Gi * Ld, 1) * 8, s));
        const int npd = Ld * (m + 1);
        if (m <= 8 && (int64_t)(T256 / cw) * npd * 8 <= 65536) {
            const int nbs = 2048;
            DevBuf<double> spart;
            DEV_ALLOC(spart, (int64_t)nbs * npd, s);
#define HDFE_L2(W, MM)                                                                              \
    hipLaunchKernelGGL((k_dense_sums<W, MM>), dim3(nbs), dim3(T256), (T256 / W) * npd * sizeof(double), \
                       s, Gi, Ld, m, P.igstart.p, P.icodes.p, M, P.offs.p, ddl, d_dense_idx.p,       \
                       P.in.p, P.isums.p, P.Ng.p, P.vbar.p, U.p, spart.p)
#define HDFE_L(W)                       \
    switch (m) {                        \
        case 1: HDFE_L2(W, 1); break;   \
        case 2: HDFE_L2(W, 2); break;   \
        case 3: HDFE_L2(W, 3); break;   \
        case 4: HDFE_L2(W, 4); break;   \
        default: HDFE_L2(W, 0); break;  \
    }
            HDFE_BY_WIDTH(cw, HDFE_L)
#undef HDFE_L
#undef HDFE_L2
            hipLaunchKernelGGL(k_dense_sums_finish, dim3(grid1((int64_t)npd * 64)), dim3(T256), 0,
                               s, nbs, Ld, m, spart.p, d_dense_levels.p, P.diag.p,
                               with_rhs ? P.rhs.p : (double *)nullptr);
            HIP_TRY(hipGetLastError());
        } else {
#define HDFE_L(W)                                                                          \
    hipLaunchKernelGGL(k_dense_u_cells<W>, dim3(grid1(Gi * W)), dim3(T256), 0, s, Gi, Ld,  \
                       P.igstart.p, P.icodes.p, M, P.offs.p, ddl, d_dense_idx.p, P.in.p, U.p)
        HDFE_BY_WIDTH(cw, HDFE_L)
#undef HDFE_L
        const int nbd = 1024, nbc = 1024;
        DevBuf<double> dpart, cpart;
        DEV_ALLOC(dpart, (int64_t)nbd * npd, s);
        DEV_ALLOC(cpart, (int64_t)nbc * npd, s);
        hipLaunchKernelGGL(k_dense_cell_sums, dim3(nbc), dim3(T256), 0, s, M, Ld, m, P.icodes.p,
                           M, P.offs.p, ddl, d_dense_idx.p, P.in.p, P.isums.p, cpart.p);
        hipLaunchKernelGGL(k_dense_diag, dim3(nbd), dim3(T256), 0, s, Gi, Ld, m, U.p, P.Ng.p,
                           P.vbar.p, dpart.p);
        hipLaunchKernelGGL(k_dense_diag_finish, dim3(grid1(npd)), dim3(T256), 0, s, nbd, nbc, Ld,
                           m, dpart.p, cpart.p, d_dense_levels.p, P.diag.p,
                           with_rhs ? P.rhs.p : (double *)nullptr);
        }
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipGetLastError());

    // sparse x sparse: emit, sort, reduce
    DevBuf<int64_t> cnt, pos;
    DEV_ALLOC(cnt, Gi + 1, s);
    DEV_ALLOC(pos, Gi + 1, s);
    HIP_TRY(hipMemsetAsync(cnt.p, 0, (Gi + 1) * 8, s));
    hipLaunchKernelGGL(k_ss_count, dim3(grid1(Gi)), dim3(T256), 0, s, Gi, P.gis.p, P.il.p,
                       d_sparse_lv.p, P.igstart.p, ncross, cnt.p);
    HIP_TRY(exclusive_sum(P.tmp, cnt.p, pos.p, Gi + 1, s));
    int64_t nent = 0;
    HIP_TRY(scalar_to_host(&nent, pos.p + Gi, s));
    // the explicit build's size from here: the emitted pairs sorted with their
    // values (two key and two value arrays) and S itself (at most one entry
    // per pair, plus the dense levels' rows); over the budget, or past what
    // this build's 32-bit counts take (2^31 emitted pairs or entries of S),
    // the panel keeps diag and rhs and no S (the caller's stream_cg
    // fallback: hdfe_stream_create)
    const int64_t est = nent * 32 + (nent + L * Ld) * 12 + (L + 1) * 4;
    const int64_t nnz_max = nent + 2 * (int64_t)Ls * Ld + (int64_t)Ld * Ld;
    if ((P.s_budget >= 0 && est > P.s_budget) || nnz_max >= INT32_MAX) {
        if (P.matrix) {
            hdfe_destroy(P.matrix);
            P.matrix = nullptr;
        }
        P.s_skipped = est;
        if (matrix) *matrix = nullptr;
        if (seconds)
            *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return HDFE_STATUS_OK;
    }
    P.s_skipped = 0;
    DevBuf<uint64_t> ek1, ek2;
    DevBuf<double> ev1, ev2;
    DEV_ALLOC(ek1, std::max<int64_t>(nent, 1), s);
    DEV_ALLOC(ek2, std::max<int64_t>(nent, 1), s);
    DEV_ALLOC(ev1, std::max<int64_t>(nent, 1), s);
    DEV_ALLOC(ev2, std::max<int64_t>(nent, 1), s);
    k1.reset();
    k2.reset();
    v1.reset();
    v2.reset();
    hipLaunchKernelGGL(k_ss_emit, dim3(grid1(Gi)), dim3(T256), 0, s, Gi, D, P.gis.p, P.il.p,
                       P.iu.p, d_sparse_lv.p, P.Ng.p, P.igstart.p, P.icodes.p, M, P.offs.p,
                       d_sparse_dim.p, lbits, pos.p, P.in.p, ek1.p, ev1.p);
    int nss = 0;
    DevBuf<int> ss_row, ss_col, ss_rowcnt, ss_rowptr;
    DevBuf<double> ss_val;
    if (nent) {
        HIP_TRY(sort_pairs(P.tmp, ek1.p, ek2.p, ev1.p, ev2.p, nent, 0, 2 * lbits, s));
        head.reset();
        rank.reset();
        DEV_ALLOC(head, nent, s);
        DEV_ALLOC(rank, nent, s);
        hipLaunchKernelGGL(k_heads_u64, dim3(grid1(nent)), dim3(T256), 0, s, nent, ek2.p, head.p);
        HIP_TRY(inclusive_sum(P.tmp, head.p, rank.p, nent, s));
        HIP_TRY(scalar_to_host(&nss, rank.p + nent - 1, s));
    }
    DEV_ALLOC(ss_row, std::max(nss, 1), s);
    DEV_ALLOC(ss_col, std::max(nss, 1), s);
    DEV_ALLOC(ss_val, std::max(nss, 1), s);
    DEV_ALLOC(ss_rowcnt, Ls + 1, s);
    DEV_ALLOC(ss_rowptr, Ls + 1, s);
    HIP_TRY(hipMemsetAsync(ss_rowcnt.p, 0, (Ls + 1) * 4, s));
    if (nss) {
        hipLaunchKernelGGL(k_reduce_runs, dim3(grid1(nent)), dim3(T256), 0, s, nent, ek2.p, ev2.p,
                           head.p, rank.p, lbits, ss_row.p, ss_col.p, ss_val.p);
        hipLaunchKernelGGL(k_ss_rowcnt, dim3(grid1(nss)), dim3(T256), 0, s, (int64_t)nss, ss_row.p,
                           d_sparse_idx.p, ss_rowcnt.p);
    }
    HIP_TRY(exclusive_sum(P.tmp, ss_rowcnt.p, ss_rowptr.p, Ls + 1, s));
    ek1.reset();
    ek2.reset();
    ev1.reset();
    ev2.reset();
    P.n_sparse_ss = nss;

    // blocks with a dense dimension
    DevBuf<double> PD, DD;
    DevBuf<uint8_t> PP, DP;
    DevBuf<int> colpres, cpos, colbuf;
    DEV_ALLOC(PD, std::max<int64_t>(Ls * Ld, 1), s);
    DEV_ALLOC(PP, std::max<int64_t>(Ls * Ld, 1), s);
    DEV_ALLOC(DD, std::max(Ld * Ld, 1), s);
    DEV_ALLOC(DP, std::max(Ld * Ld, 1), s);
    DEV_ALLOC(colpres, std::max(Ld, 1), s);
    DEV_ALLOC(cpos, std::max<int64_t>((int64_t)Ld * (Ls + 1), 1), s);
    if (Ld) {
        // the one dense dimension, if there is exactly one (k_sd_inc's short path)
        int dense1 = -1, ndense = 0;
        for (int t = 0; t < D; ++t)
            if (!sparse_dim[t]) {
                dense1 = t;
                ++ndense;
            }
        if (ndense != 1) dense1 = -1;
        {
            DevBuf<int> sflag, sidx;
            DEV_ALLOC(sflag, (int64_t)E + 1, s);
            DEV_ALLOC(sidx, (int64_t)E + 1, s);
            hipLaunchKernelGGL(k_sparse_flag, dim3(grid1((int64_t)E + 1)), dim3(T256), 0, s,
                               (int64_t)E, P.il.p, d_sparse_lv.p, sflag.p);
            HIP_TRY(exclusive_sum(P.tmp, sflag.p, sidx.p, (int64_t)E + 1, s));
            int Es = 0;
            HIP_TRY(scalar_to_host(&Es, sidx.p + E, s));
            DevBuf<double> Cq;
            DevBuf<uint64_t> Fq;
            DEV_ALLOC(Cq, std::max<int64_t>((int64_t)Es * Ld, 1), s);
            DEV_ALLOC(Fq, std::max(Es, 1), s);
            hipLaunchKernelGGL(k_sd_inc, dim3(grid1(E)), dim3(T256), 0, s, (int64_t)E, D, Ld,
                               P.ig.p, P.il.p, P.iu.p, sidx.p, d_sparse_lv.p, d_lvl_dim.p,
                               P.offs.p, d_sparse_dim.p, d_dense_idx.p, U.p, P.Ng.p,
                               P.igstart.p, P.icodes.p, M, P.in.p, Cq.p, Fq.p, dense1);
            hipLaunchKernelGGL(k_sd_sum, dim3(grid1(Ls * Ld)), dim3(T256), 0, s, Ls, Ld,
                               d_sparse_levels.p, P.lptr.p, P.lq.p, sidx.p, Cq.p, Fq.p, PD.p,
                               PP.p);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(s));
        }
        const int nbd = Gi > (int64_t)1 << 20 ? 4096 : 1024;
        DevBuf<double> ddp;
        DevBuf<int> ddpp;
        DEV_ALLOC(ddp, (int64_t)nbd * Ld * Ld, s);
        DEV_ALLOC(ddpp, (int64_t)nbd * Ld * Ld, s);
#define HDFE_DDR(LL)                                                                   \
    case LL:                                                                           \
        hipLaunchKernelGGL(k_dd_rows<LL>, dim3(nbd), dim3(T256), 0, s, Gi, U.p, P.Ng.p, \
                           ddp.p, ddpp.p);                                             \
        break;
        switch (dense1 >= 0 ? Ld : 0) {
            HDFE_DDR(2)
            HDFE_DDR(3)
            HDFE_DDR(4)
            HDFE_DDR(5)
            HDFE_DDR(6)
            HDFE_DDR(7)
            HDFE_DDR(8)
            default:
                hipLaunchKernelGGL(k_dd, dim3(nbd), dim3(T256), 0, s, Gi, Ld, d_dense_levels.p,
                                   d_lvl_dim.p, P.offs.p, U.p, P.Ng.p, P.igstart.p, P.icodes.p,
                                   M, P.in.p, ddp.p, ddpp.p);
        }
#undef HDFE_DDR
        hipLaunchKernelGGL(k_dd_finish, dim3(grid1((int64_t)Ld * Ld * 64)), dim3(T256), 0, s, nbd, Ld * Ld,
                           ddp.p, ddpp.p, DD.p, DP.p);
        // per dense level: positions of the present sparse levels
        DEV_ALLOC(colbuf, Ls + 1, s);
        std::vector<int> cp(Ld);
        for (int t = 0; t < Ld; ++t) {
            HIP_TRY(hipMemsetAsync(colbuf.p + Ls, 0, 4, s));
            hipLaunchKernelGGL(k_pp_col, dim3(grid1(Ls)), dim3(T256), 0, s, Ls, Ld, t, PP.p, colbuf.p);
            HIP_TRY(exclusive_sum(P.tmp, colbuf.p, cpos.p + (int64_t)t * (Ls + 1), Ls + 1, s));
        }
        for (int t = 0; t < Ld; ++t)
            HIP_TRY(hipMemcpyAsync(colpres.p + t, cpos.p + (int64_t)t * (Ls + 1) + Ls, 4,
                                   hipMemcpyDeviceToDevice, s));
        U.reset();
    }
    // CSR
    DevBuf<int> rnnz, indptr;
    DEV_ALLOC(rnnz, L + 1, s);
    DEV_ALLOC(indptr, L + 1, s);
    HIP_TRY(hipMemsetAsync(rnnz.p, 0, (L + 1) * 4, s));
    hipLaunchKernelGGL(k_row_nnz, dim3(grid1(L)), dim3(T256), 0, s, L, Ld, ss_rowcnt.p,
                       d_sparse_idx.p, d_dense_idx.p, PP.p, colpres.p, DP.p, rnnz.p);
    HIP_TRY(exclusive_sum(P.tmp, rnnz.p, indptr.p, L + 1, s));
    int nnz = 0;
    HIP_TRY(scalar_to_host(&nnz, indptr.p + L, s));
    P.nnz = nnz;
    DevBuf<int> indices;
    DevBuf<double> data, dinv;
    DEV_ALLOC(indices, std::max(nnz, 1), s);
    DEV_ALLOC(data, std::max(nnz, 1), s);
    DEV_ALLOC(dinv, L, s);
    if (Ld <= 16 && Ls > 0) {
        // W lanes per row, near the row's entries
        const int rw = run_width(nnz, Ls);
#define HDFE_L(W)                                                                              \
    hipLaunchKernelGGL(k_fill_sparse_rows_w<W>, dim3(grid1(Ls * W)), dim3(T256), 0, s, Ls, Ld, \
                       d_sparse_levels.p, d_dense_levels.p, ss_rowptr.p, ss_col.p, ss_val.p,    \
                       PP.p, PD.p, indptr.p, indices.p, data.p)
        HDFE_BY_WIDTH(rw, HDFE_L)
#undef HDFE_L
    } else {
        hipLaunchKernelGGL(k_fill_sparse_rows, dim3(grid1(Ls)), dim3(T256), 0, s, Ls, Ld,
                           d_sparse_levels.p, d_dense_levels.p, ss_rowptr.p, ss_col.p, ss_val.p,
                           PP.p, PD.p, indptr.p, indices.p, data.p);
    }
    if (Ld) {
        hipLaunchKernelGGL(k_fill_dense_sparse, dim3(grid1(Ls * Ld)), dim3(T256), 0, s, Ls, Ld,
                           d_sparse_levels.p, d_dense_levels.p, PP.p, PD.p, cpos.p, DP.p, indptr.p,
                           indices.p, data.p);
        hipLaunchKernelGGL(k_fill_dense_dense, dim3(grid1(Ld * Ld)), dim3(T256), 0, s, Ls, Ld,
                           d_sparse_levels.p, d_dense_levels.p, cpos.p, DP.p, DD.p, indptr.p,
                           indices.p, data.p);
    }
    hipLaunchKernelGGL(k_jacobi, dim3(grid1(L)), dim3(T256), 0, s, L, P.diag.p, dinv.p);
    HIP_TRY(hipGetLastError());
    std::vector<int> ip;
    HIP_TRY(to_host(ip, indptr.p, L + 1, s));
    if (P.matrix) {
        hdfe_destroy(P.matrix);
        P.matrix = nullptr;
    }
    void *mat = nullptr;
    // the matrix takes over the row pointers, indices and values (the copy
    // to the host above ordered every write to them)
    TRY(matrix_create(L, nnz, ip, indices.p, data.p, dinv.p, true, &mat, indptr.p, indices.p,
                      data.p));
    indptr.p = nullptr;
    indices.p = nullptr;
    data.p = nullptr;
    P.matrix = mat;
    if (matrix) *matrix = mat;
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return HDFE_STATUS_OK;
}

// Components of dimension 0's levels through the identifying groups, ranked
// by rows (comp: dimension-0 levels, owned by the panel).  Returns the count.
int hdfe_panel_components(void *h, int64_t *ncomp) {
    ARG_CHECK(h, "hdfe_panel_components: null handle");
    Panel &P = *(Panel *)h;
    hipStream_t s = P.s;
    ARG_CHECK(P.D >= 1, "hdfe_panel_components: no non-streamed dimension");
    const int64_t nl = P.nlev[0];
    DevBuf<int> parent, is_root, rpos;
    DEV_ALLOC(parent, nl, s);
    DEV_ALLOC(is_root, nl + 1, s);
    DEV_ALLOC(rpos, nl + 1, s);
    hipLaunchKernelGGL(k_uf_init, dim3(grid1(nl)), dim3(T256), 0, s, nl, parent.p);
    // one pass joins every group's levels (see k_uf_union)
    const int uw = run_width(P.M, P.Gi);
#define HDFE_L(W)                                                                          \
    hipLaunchKernelGGL(k_uf_union<W>, dim3(grid1(P.Gi * W)), dim3(T256), 0, s, P.Gi,        \
                       P.igstart.p, P.icodes.p, parent.p)
    HDFE_BY_WIDTH(uw, HDFE_L)
#undef HDFE_L
    hipLaunchKernelGGL(k_uf_flatten, dim3(grid1(nl)), dim3(T256), 0, s, nl, parent.p);
    DevBuf<unsigned long long> rows;
    DEV_ALLOC(rows, nl, s);
    HIP_TRY(hipMemsetAsync(rows.p, 0, nl * 8, s));
    HIP_TRY(hipMemsetAsync(is_root.p + nl, 0, 4, s));
    hipLaunchKernelGGL(k_comp_rows, dim3(grid1(nl)), dim3(T256), 0, s, nl, parent.p, P.cnt.p,
                       rows.p, is_root.p);
    HIP_TRY(exclusive_sum(P.tmp, is_root.p, rpos.p, nl + 1, s));
    int nc = 0;
    HIP_TRY(scalar_to_host(&nc, rpos.p + nl, s));
    DevBuf<uint64_t> k1, k2;
    DevBuf<int> v1, v2, rank_of_root;
    DEV_ALLOC(k1, nc, s);
    DEV_ALLOC(k2, nc, s);
    DEV_ALLOC(v1, nc, s);
    DEV_ALLOC(v2, nc, s);
    DEV_ALLOC(rank_of_root, nl, s);
    hipLaunchKernelGGL(k_comp_keys, dim3(grid1(nl)), dim3(T256), 0, s, nl, is_root.p, rows.p,
                       rpos.p, k1.p, v1.p);
    HIP_TRY(sort_pairs(P.tmp, k1.p, k2.p, v1.p, v2.p, nc, 0, 64, s));
    hipLaunchKernelGGL(k_comp_rank, dim3(grid1(nc)), dim3(T256), 0, s, (int64_t)nc, v2.p,
                       rank_of_root.p);
    DEV_ALLOC(P.comp, nl, s);
    hipLaunchKernelGGL(k_comp_label, dim3(grid1(nl)), dim3(T256), 0, s, nl, parent.p,
                       rank_of_root.p, P.comp.p);
    // levels grouped by component, for the normalisation
    DevBuf<uint64_t> lk1, lk2;
    DevBuf<int> lv1, nch;
    DEV_ALLOC(lk1, nl, s);
    DEV_ALLOC(lk2, nl, s);
    DEV_ALLOC(lv1, nl, s);
    DEV_ALLOC(P.clev, nl, s);
    DEV_ALLOC(P.cptr, nc + 1, s);
    const int lshift = std::max(1, bits_for((uint64_t)nl));
    const int kbits = lshift + std::max(1, bits_for((uint64_t)nc));
    hipLaunchKernelGGL(k_comp_level_keys, dim3(grid1(nl)), dim3(T256), 0, s, nl, lshift, P.comp.p,
                       lk1.p, lv1.p);
    HIP_TRY(sort_pairs(P.tmp, lk1.p, lk2.p, lv1.p, P.clev.p, nl, 0, std::min(kbits, 64), s));
    hipLaunchKernelGGL(k_comp_counts, dim3(grid1(nl)), dim3(T256), 0, s, nl, lshift, lk2.p,
                       P.cptr.p);
    // chunks of the sorted levels for the normalisation's sums
    DEV_ALLOC(nch, nc + 1, s);
    DEV_ALLOC(P.cchp, nc + 1, s);
    hipLaunchKernelGGL(k_comp_nchunk, dim3(grid1(nc + 1)), dim3(T256), 0, s, (int64_t)nc,
                       P.cptr.p, nch.p);
    HIP_TRY(exclusive_sum(P.tmp, nch.p, P.cchp.p, nc + 1, s));
    int nchunk = 0;
    HIP_TRY(scalar_to_host(&nchunk, P.cchp.p + nc, s));
    P.ncchunk = nchunk;
    DEV_ALLOC(P.cch_beg, std::max(nchunk, 1), s);
    DEV_ALLOC(P.cch_end, std::max(nchunk, 1), s);
    hipLaunchKernelGGL(k_comp_chunks, dim3(grid1(nc)), dim3(T256), 0, s, (int64_t)nc, P.cptr.p,
                       P.cchp.p, P.cch_beg.p);
    hipLaunchKernelGGL(k_comp_chunk_ends, dim3(grid1(nc)), dim3(T256), 0, s, (int64_t)nc,
                       P.cptr.p, P.cchp.p, P.cch_end.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    P.ncomp = nc;
    if (ncomp) *ncomp = nc;
    return HDFE_STATUS_OK;
}

// Connected components of the bipartite graph of two fixed-effect
// dimensions' levels linked by the rows (core/passes.py:964-979 for pairs
// with the streamed dimension, :525-527 for the others): a, b index the
// dimensions as the fit orders them, 0 the streamed one, t + 1 the
// non-streamed t, a < b.  Every level is in the sample.  With the streamed
// dimension every group has a row, so the count is that of b's levels
// joined through the groups (unions within each group's rows).
int hdfe_panel_pair_components(void *h, int a, int b, int64_t *count) {
    ARG_CHECK(h && count, "hdfe_panel_pair_components: null argument");
    Panel &P = *(Panel *)h;
    ARG_CHECK(a >= 0 && a < b && b <= P.D, "hdfe_panel_pair_components: bad pair");
    hipStream_t s = P.s;
    int ta = a - 1, tb = b - 1;
    // the side of at most 64 levels as b where there is one (a non-streamed
    // pair is symmetric)
    if (a > 0 && P.nlev[tb] > 64 && P.nlev[ta] <= 64) std::swap(ta, tb);
    const int64_t nb = P.nlev[tb];
    const int64_t na = a == 0 ? 0 : P.nlev[ta];
    const int64_t nv = na + nb;
    ARG_CHECK(nv < INT32_MAX, "hdfe_panel_pair_components: too many levels");
    if (nb <= 64) {
        DevBuf<unsigned long long> adj, masks;
        DEV_ALLOC(adj, 64, s);
        HIP_TRY(hipMemsetAsync(adj.p, 0, 64 * 8, s));
        const int *cbp = P.codes.p + (int64_t)tb * P.n;
        if (P.n > 0 && a == 0) {
#define HDFE_L(W)                                                                           \
    hipLaunchKernelGGL(k_group_masks<W>, dim3(grid1(P.G * W, 256, 2048)), dim3(T256), 0, s,  \
                       P.G, P.gstart.p, cbp, adj.p)
            HDFE_BY_WIDTH(run_width(P.n, P.G), HDFE_L)
#undef HDFE_L
        } else if (P.n > 0) {
            DEV_ALLOC(masks, na, s);
            HIP_TRY(hipMemsetAsync(masks.p, 0, na * 8, s));
            hipLaunchKernelGGL(k_row_masks, dim3(grid1(P.n)), dim3(T256), 0, s, P.n,
                               P.codes.p + (int64_t)ta * P.n, cbp, masks.p);
            hipLaunchKernelGGL(k_mask_adj, dim3(grid1(na, 256, 2048)), dim3(T256), 0, s, na,
                               masks.p, adj.p);
        }
        HIP_TRY(hipGetLastError());
        std::vector<unsigned long long> A;
        HIP_TRY(to_host(A, adj.p, 64, s));
        // components of the small side's nb levels (every level has rows)
        int par[64];
        for (int y = 0; y < 64; ++y) par[y] = y;
        auto root = [&](int y) {
            while (par[y] != y) y = par[y] = par[par[y]];
            return y;
        };
        for (int y = 0; y < nb; ++y)
            for (unsigned long long v = A[y]; v; v &= v - 1) {
                const int z = __builtin_ctzll(v), ry = root(y), rz = root(z);
                if (ry != rz) par[std::max(ry, rz)] = std::min(ry, rz);
            }
        int64_t c = 0;
        for (int y = 0; y < nb; ++y) c += root(y) == y;
        *count = c;
        return HDFE_STATUS_OK;
    }
    DevBuf<int> parent;
    DevBuf<unsigned long long> cnt;
    DEV_ALLOC(parent, std::max<int64_t>(nv, 1), s);
    DEV_ALLOC(cnt, 1, s);
    hipLaunchKernelGGL(k_uf_init, dim3(grid1(nv)), dim3(T256), 0, s, nv, parent.p);
    // the distinct edges first when their bitmap is modest (a few-level
    // dimension: its rows repeat few pairs), else a union per row
    const int64_t nx = a == 0 ? nb : na;
    const int64_t nbits = nx * nb;
    if (P.n > 0 && nbits <= ((int64_t)1 << 31)) {
        const int64_t nw = (nbits + 31) / 32;
        DevBuf<unsigned> bits;
        DEV_ALLOC(bits, nw, s);
        HIP_TRY(hipMemsetAsync(bits.p, 0, nw * 4, s));
        if (a == 0) {
#define HDFE_L(W)                                                                           \
    hipLaunchKernelGGL(k_group_pair_bits<W>, dim3(grid1(P.G * W)), dim3(T256), 0, s, P.G,    \
                       P.gstart.p, P.codes.p + (int64_t)tb * P.n, nb, bits.p)
            HDFE_BY_WIDTH(run_width(P.n, P.G), HDFE_L)
#undef HDFE_L
        } else {
            hipLaunchKernelGGL(k_pair_bits, dim3(grid1(P.n)), dim3(T256), 0, s, P.n,
                                   P.codes.p + (int64_t)ta * P.n,
                               P.codes.p + (int64_t)tb * P.n, nb, bits.p);
        }
        hipLaunchKernelGGL(k_uf_bits, dim3(grid1(nw)), dim3(T256), 0, s, nw, bits.p, nb,
                           a == 0 ? 0 : (int)na, parent.p);
    } else if (P.n > 0) {
        if (a == 0) {
            const int uw = run_width(P.n, P.G);
#define HDFE_L(W)                                                                          \
    hipLaunchKernelGGL(k_uf_union<W>, dim3(grid1(P.G * W)), dim3(T256), 0, s, P.G,          \
                       P.gstart.p, P.codes.p + (int64_t)tb * P.n, parent.p)
            HDFE_BY_WIDTH(uw, HDFE_L)
#undef HDFE_L
        } else {
            hipLaunchKernelGGL(k_uf_rows, dim3(grid1(P.n)), dim3(T256), 0, s, P.n,
                               P.codes.p + (int64_t)ta * P.n,
                               P.codes.p + (int64_t)tb * P.n, (int)na, parent.p);
        }
    }
    HIP_TRY(hipMemsetAsync(cnt.p, 0, 8, s));
    hipLaunchKernelGGL(k_uf_count_roots, dim3(grid1(nv)), dim3(T256), 0, s, nv, parent.p, cnt.p);
    HIP_TRY(hipGetLastError());
    unsigned long long c = 0;
    HIP_TRY(scalar_to_host(&c, cnt.p, s));
    *count = (int64_t)c;
    return HDFE_STATUS_OK;
}

// Normalise Gamma (L, k) on the device in place (core/solve.py:383-399).
int hdfe_panel_normalize(void *h, double *Gm, int k, void *stream) {
    ARG_CHECK(h && Gm, "hdfe_panel_normalize: null argument");
    Panel &P = *(Panel *)h;
    ARG_CHECK(P.comp.p, "hdfe_panel_normalize: components not computed");
    hipStream_t s = stream ? (hipStream_t)stream : P.s;
    const int64_t nl0 = P.nlev[0];
    DevBuf<double> mean;
    DEV_ALLOC(mean, P.ncomp * k, s);
    DevBuf<double> cpart;
    DEV_ALLOC(cpart, std::max<int64_t>(P.ncchunk, 1) * (k + 1), s);
    hipLaunchKernelGGL(k_comp_chunk_sums, dim3((int)std::min<int64_t>(P.ncchunk, 65536)),
                       dim3(T256), 0, s, P.ncchunk, k, P.cch_beg.p, P.cch_end.p, P.clev.p,
                       P.obs.p, Gm, cpart.p);
    hipLaunchKernelGGL(k_comp_means, dim3(grid1(P.ncomp * k)), dim3(T256), 0, s, P.ncomp, k,
                       P.cchp.p, cpart.p, mean.p);
    hipLaunchKernelGGL(k_sub_comp_means, dim3(grid1(nl0 * k)), dim3(T256), 0, s, nl0, k, P.comp.p,
                       mean.p, Gm);
    const int nb = 64;
    DevBuf<double> part;
    DEV_ALLOC(part, (int64_t)nb * (k + 1), s);
    for (int t = 1; t < P.D; ++t) {
        hipLaunchKernelGGL(k_wsum_block, dim3(nb), dim3(T256), 0, s, P.nlev[t], k, P.offs_h[t],
                           P.obs.p, Gm, part.p);
        hipLaunchKernelGGL(k_sub_global_mean, dim3(grid1(P.nlev[t] * k)), dim3(T256), 0, s,
                           P.nlev[t], k, P.offs_h[t], nb, part.p, Gm);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return HDFE_STATUS_OK;
}

// A = W_within + V' M_D V between cells (m x m, host out; more than eight
// variables by 8 x 8 tiles).
int hdfe_panel_assemble(void *h, const double *Gm, double *A) {
    ARG_CHECK(h && Gm && A, "hdfe_panel_assemble: null argument");
    Panel &P = *(Panel *)h;
    hipStream_t s = P.s;
    const int m = P.m, nb = 4096, cw = run_width(P.M, P.Gi);
    if (m > 8) {
        // by 8 x 8 tiles of the variables
        DevBuf<double> tpart;
        DEV_ALLOC(tpart, (int64_t)nb * 64, s);
        std::vector<double> th, Wh;
        HIP_TRY(to_host(Wh, P.W.p, (int64_t)m * m, s));
        const int nt = (m + 7) / 8;
        for (int ta = 0; ta < nt; ++ta)
            for (int tb = ta; tb < nt; ++tb) {
                const int ja = 8 * ta, jb = 8 * tb;
                const int na = std::min(8, m - ja), nbc = std::min(8, m - jb);
#define HDFE_L(W)                                                                               \
    hipLaunchKernelGGL(k_assemble_tile<W>, dim3(nb), dim3(T256), 0, s, P.Gi, P.D, m, ja, na, jb, \
                       nbc, P.igstart.p, P.icodes.p, P.M, P.offs.p, P.in.p, P.isums.p, Gm, tpart.p)
                HDFE_BY_WIDTH(cw, HDFE_L)
#undef HDFE_L
                HIP_TRY(hipGetLastError());
                HIP_TRY(to_host(th, tpart.p, (int64_t)nb * 64, s));
                for (int a = 0; a < na; ++a)
                    for (int b = 0; b < nbc; ++b) {
                        double v = Wh[(size_t)(ja + a) * m + jb + b];
                        for (int q = 0; q < nb; ++q) v += th[(size_t)q * 64 + a * 8 + b];
                        A[(ja + a) * m + jb + b] = v;
                        A[(jb + b) * m + ja + a] = v;
                    }
            }
        return HDFE_STATUS_OK;
    }
    DevBuf<double> part;
    DEV_ALLOC(part, (int64_t)nb * m * m, s);
    switch (m) {
#define HDFE_ASMW(MM, W)                                                                     \
    hipLaunchKernelGGL((k_assemble<MM, W>), dim3(nb), dim3(T256), 0, s, P.Gi, P.D, P.igstart.p, \
                       P.icodes.p, P.M, P.offs.p, P.in.p, P.isums.p, Gm, part.p)
#define HDFE_ASM(MM)                                          \
    case MM:                                                  \
        if (cw == 4) HDFE_ASMW(MM, 4);                        \
        else if (cw == 8) HDFE_ASMW(MM, 8);                   \
        else if (cw == 16) HDFE_ASMW(MM, 16);                 \
        else HDFE_ASMW(MM, 32);                               \
        break;
        HDFE_ASM(1) HDFE_ASM(2) HDFE_ASM(3) HDFE_ASM(4) HDFE_ASM(5) HDFE_ASM(6) HDFE_ASM(7)
        HDFE_ASM(8)
#undef HDFE_ASM
#undef HDFE_ASMW
    }
    HIP_TRY(hipGetLastError());
    std::vector<double> hp, Wh;
    HIP_TRY(to_host(hp, part.p, (int64_t)nb * m * m, s));
    HIP_TRY(to_host(Wh, P.W.p, m * m, s));
    for (int t = 0; t < m * m; ++t) {
        double v = Wh[t];
        for (int b = 0; b < nb; ++b) v += hp[(size_t)b * m * m + t];
        A[t] = v;
    }
    return HDFE_STATUS_OK;
}

// The y-equation's effects gam_y (L, device out), the streamed effects
// (G, device out) and the residuals (n, device out, sorted row order).
int hdfe_panel_pass2(void *h, const double *Gm, const double *beta, const int *xcols, int kx,
                     double *gam_y, double *geff, double *resid) {
    ARG_CHECK(h && Gm && gam_y && geff && resid, "hdfe_panel_pass2: null argument");
    Panel &P = *(Panel *)h;
    hipStream_t s = P.s;
    DevBuf<double> b;
    DevBuf<int> xc;
    DEV_ALLOC(b, std::max(kx, 1), s);
    DEV_ALLOC(xc, std::max(kx, 1), s);
    if (kx) {
        HIP_TRY(hipMemcpyAsync(b.p, beta, kx * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(xc.p, xcols, kx * 4, hipMemcpyHostToDevice, s));
    }
    hipLaunchKernelGGL(k_gam_y, dim3(grid1(P.L)), dim3(T256), 0, s, P.L, P.m, Gm, b.p, xc.p, kx, gam_y);
#define HDFE_L(W)                                                                            \
    hipLaunchKernelGGL(k_pass2<W>, dim3(grid1(P.G * W)), dim3(T256), 0, s, P.G, P.D, P.m,    \
                       P.gstart.p, P.codes.p, P.n, P.offs.p, P.V.p,                          \
                       P.weighted ? P.w.p : (const double *)nullptr, gam_y, b.p, xc.p, kx,   \
                       geff, resid)
    HDFE_BY_WIDTH(run_width(P.n, P.G), HDFE_L)
#undef HDFE_L
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return HDFE_STATUS_OK;
}

// New variables V (n, m) and weights w (n), device, in the panel's sorted row
// order, then pass 1 again; the rows, their order and the levels stay.  The
// caller rebuilds the system (hdfe_panel_system).  For IRLS, whose every step
// is a weighted fit with new weights and working response (core/irls.py:
// 92-123).
int hdfe_panel_reweight(void *h, const double *V, const double *w) {
    ARG_CHECK(h && V && w, "hdfe_panel_reweight: null argument");
    Panel &P = *(Panel *)h;
    ARG_CHECK(P.weighted, "hdfe_panel_reweight: the panel was made without weights");
    hipStream_t s = P.s;
    HIP_TRY(hipMemcpyAsync(P.V.p, V, P.n * P.m * 8, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(P.w.p, w, P.n * 8, hipMemcpyDeviceToDevice, s));
    TRY(run_pass1(P));
    HIP_TRY(hipStreamSynchronize(s));
    return HDFE_STATUS_OK;
}

// out (n, kc, device) = columns cols (host, kc) of the panel's variables
// partialled on the fixed effects, from the reduced solutions Gm (L, m).
int hdfe_panel_partial(void *h, const double *Gm, const int *cols, int kc, double *out) {
    ARG_CHECK(h && Gm && cols && out && kc >= 1, "hdfe_panel_partial: bad argument");
    Panel &P = *(Panel *)h;
    for (int c = 0; c < kc; ++c)
        ARG_CHECK(cols[c] >= 0 && cols[c] < P.m, "hdfe_panel_partial: column out of range");
    hipStream_t s = P.s;
    DevBuf<int> cc;
    DEV_ALLOC(cc, kc, s);
    HIP_TRY(hipMemcpyAsync(cc.p, cols, kc * 4, hipMemcpyHostToDevice, s));
#define HDFE_L(W)                                                                            \
    hipLaunchKernelGGL(k_partial<W>, dim3(grid1(P.G * W)), dim3(T256), 0, s, P.G, P.D, P.m,  \
                       P.gstart.p, P.codes.p, P.n, P.offs.p, P.V.p,                          \
                       P.weighted ? P.w.p : (const double *)nullptr, Gm, cc.p, kc, out)
    HDFE_BY_WIDTH(run_width(P.n, P.G), HDFE_L)
#undef HDFE_L
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return HDFE_STATUS_OK;
}

// The scores of the covariates cols (host, kc <= 8) with residuals e
// (sorted rows, device) and their sums (see k_scores): out (host) = [rss,
// sum_i u_i u_i' (kc x kc), sum_g s_g s_g' (kc x kc)]; U (n, kc, device)
// receives the scores when not null.
int hdfe_panel_scores(void *h, const double *Gm, const int *cols, int kc, const double *e,
                      double *U, double *out) {
    ARG_CHECK(h && Gm && cols && e && out && kc >= 1 && kc <= 8,
              "hdfe_panel_scores: bad argument (1 <= kc <= 8)");
    Panel &P = *(Panel *)h;
    for (int c = 0; c < kc; ++c)
        ARG_CHECK(cols[c] >= 0 && cols[c] < P.m, "hdfe_panel_scores: column out of range");
    hipStream_t s = P.s;
    ColSel cc{};
    for (int c = 0; c < kc; ++c) cc.c[c] = cols[c];
    const int W = run_width(P.n, P.G);
    const int nb = grid1(P.G * W, 256, 1024);
    const int K = kc <= 2 ? 2 : kc <= 4 ? 4 : 8;
    const int NV = K == 2 ? ScoreAcc<2>::NV : K == 4 ? ScoreAcc<4>::NV : ScoreAcc<8>::NV;
    DevBuf<double> part;
    DEV_ALLOC(part, (int64_t)nb * NV, s);
#define HDFE_L2(WW, KK)                                                                        \
    hipLaunchKernelGGL((k_scores<WW, KK>), dim3(nb), dim3(T256), 0, s, P.G, P.D, P.m,         \
                       P.gstart.p, P.codes.p, P.n, P.offs.p, P.V.p,                           \
                       P.weighted ? P.w.p : (const double *)nullptr, Gm, cc, kc, e, U, part.p)
#define HDFE_L(WW)                 \
    if (K == 2) HDFE_L2(WW, 2);    \
    else if (K == 4) HDFE_L2(WW, 4); \
    else HDFE_L2(WW, 8)
    HDFE_BY_WIDTH(W, HDFE_L)
#undef HDFE_L
#undef HDFE_L2
    HIP_TRY(hipGetLastError());
    std::vector<double> hp;
    HIP_TRY(to_host(hp, part.p, (int64_t)nb * NV, s));
    std::vector<double> tot(NV, 0.0);
    for (int b = 0; b < nb; ++b)
        for (int v = 0; v < NV; ++v) tot[v] += hp[(size_t)b * NV + v];
    const int T = K * (K + 1) / 2;
    out[0] = tot[0];
    int t = 0;
    for (int a = 0; a < K; ++a)
        for (int b = a; b < K; ++b, ++t)
            if (a < kc && b < kc) {
                out[1 + a * kc + b] = out[1 + b * kc + a] = tot[1 + t];
                out[1 + kc * kc + a * kc + b] = out[1 + kc * kc + b * kc + a] = tot[1 + T + t];
            }
    return HDFE_STATUS_OK;
}

void hdfe_panel_destroy(void *h) {
    if (!h) return;
    Panel *P = (Panel *)h;
    // the panel's arrays may have been read on other streams; its buffers
    // are freed on its stream, which must outlive them
    (void)hipDeviceSynchronize();
    if (P->matrix) hdfe_destroy(P->matrix);
    hipStream_t s = P->s, side = P->side;
    hipEvent_t side_ev = P->side_ev;
    const int dev = P->dev;
    delete P;
    stream_put(s, dev);
    stream_put(side, dev);
    if (side_ev) (void)hipEventDestroy(side_ev);
}

}  // extern "C"
