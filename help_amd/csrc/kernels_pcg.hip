// Kernels of the matrix-free PCG path (HELP PCG.cpp:38-395).
//
// * Row-split CSR products: A has rows of wildly different lengths (Zipf row
//   popularity gives rows with millions of nonzeros), so rows up to
//   kLightRow nonzeros take a thread, rows up to kHeavyRow a wave, and
//   longer rows are cut into chunks of kRowChunk, one workgroup each,
//   combined in chunk order: deterministic.
// * Greedy partial Cholesky (PCG.cpp:38-103): per pivot an argmax over the
//   Schur diagonal, the pivot row of A scattered into an n-vector, one
//   product with A for the pivot's column of ADA^T, and a left-looking
//   column update against the previous columns (column-major m x K factor,
//   so the thread-per-row update reads coalesced).
// * Preconditioner apply (PCG.cpp:183-196) with the factor in FP32: two
//   passes over the m x K factor (a product and a transposed product).
#include <algorithm>
#include <climits>
#include <cmath>

#include "device.h"
#include "launch.h"
#include "pcg.h"

namespace help {

namespace {

constexpr int kB = 256;
inline int grid_for(int64_t n) { return (int)std::max<int64_t>(1, (n + kB - 1) / kB); }

__device__ inline double block_sum(double v, double* red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
}

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// ------------------------------------------------------ row-split products
template <bool SQ>
__global__ __launch_bounds__(256) void k_csr_light(int32_t m, const int64_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ colind,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ x, double* __restrict__ y) {
    const int32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const int64_t b = rowptr[i], e = rowptr[i + 1];
    if (e - b > kLightRow) return;
    // 8 terms' loads in flight (indices past the row clamped, terms dropped)
    double acc = 0.0;
    for (int64_t p = b; p < e; p += 8) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int64_t q = p + u < e ? p + u : e - 1;
            const double v = val[q];
            t[u] = (SQ ? v * v : v) * x[colind[q]];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = (p + u < e) ? acc + t[u] : acc;
    }
    y[i] = acc;
}

// a wave per medium row: lane-strided partial sums with four terms in
// flight, combined by a fixed butterfly
template <bool SQ>
__global__ __launch_bounds__(256) void k_csr_medium(int32_t nmed, const int32_t* __restrict__ mrow,
                                                    const int64_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ colind,
                                                    const double* __restrict__ val,
                                                    const double* __restrict__ x, double* __restrict__ y) {
    const int32_t h = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (h >= nmed) return;
    const int32_t i = mrow[h];
    const int64_t b = rowptr[i], e = rowptr[i + 1];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t p0 = b + lane; p0 < e; p0 += 256) {
        double t[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t q = p0 + 64 * u < e ? p0 + 64 * u : e - 1;
            const double v = val[q];
            t[u] = (SQ ? v * v : v) * x[colind[q]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = (p0 + 64 * u < e) ? a[u] + t[u] : a[u];
    }
    const double s = wave_sum((a[0] + a[1]) + (a[2] + a[3]));
    if (lane == 0) y[i] = s;
}

template <bool SQ>
__global__ __launch_bounds__(256) void k_csr_chunk(const RowChunk* __restrict__ chunks,
                                                   const int32_t* __restrict__ colind,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ x, double* __restrict__ part) {
    __shared__ double red[kB];
    const RowChunk c = chunks[blockIdx.x];
    double acc = 0.0;
    for (int64_t p = c.b + threadIdx.x; p < c.e; p += kB) {
        const double v = val[p];
        acc += (SQ ? v * v : v) * x[colind[p]];
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// a wave per heavy row: its chunk partials summed lane-strided, fixed tree
__global__ __launch_bounds__(256) void k_csr_heavy(int32_t nheavy, const int32_t* __restrict__ hrow,
                                                   const int32_t* __restrict__ hptr,
                                                   const double* __restrict__ part, double* __restrict__ y) {
    const int32_t h = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (h >= nheavy) return;
    double acc = 0.0;
    for (int32_t c = hptr[h] + lane; c < hptr[h + 1]; c += 64) acc += part[c];
    const double s = wave_sum(acc);
    if (lane == 0) y[hrow[h]] = s;
}

__global__ void k_csc_ts(int32_t n, const int64_t* __restrict__ colptr, const int32_t* __restrict__ rowind,
                         const double* __restrict__ val, const double* __restrict__ dinv,
                         const double* __restrict__ x, double* __restrict__ t) {
    const int32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= n) return;
    const int64_t b = colptr[j], e = colptr[j + 1];
    double acc = 0.0;
    for (int64_t p = b; p < e; p += 4) {
        double tt[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t q = p + u < e ? p + u : e - 1;
            tt[u] = val[q] * x[rowind[q]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = (p + u < e) ? acc + tt[u] : acc;
    }
    t[j] = acc * dinv[j];
}

// ---------------------------------------------------- partial Cholesky build
__device__ inline bool better(double v, int32_t i, double bv, int32_t bi) {
    return v > bv || (v == bv && i < bi);
}

__global__ __launch_bounds__(256) void k_pc_argmax_part(int32_t m, int32_t k, const int32_t* __restrict__ perm,
                                                        const double* __restrict__ S, double* __restrict__ pval,
                                                        int32_t* __restrict__ pidx) {
    __shared__ double sv[kB];
    __shared__ int32_t si[kB];
    const int t = threadIdx.x;
    double bv = -HUGE_VAL;
    int32_t bi = INT_MAX;
    for (int64_t i = (int64_t)k + (int64_t)blockIdx.x * kB + t; i < m; i += (int64_t)gridDim.x * kB) {
        const double v = S[perm[i]];
        if (better(v, (int32_t)i, bv, bi)) { bv = v; bi = (int32_t)i; }
    }
    sv[t] = bv;
    si[t] = bi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w && better(sv[t + w], si[t + w], sv[t], si[t])) { sv[t] = sv[t + w]; si[t] = si[t + w]; }
        __syncthreads();
    }
    if (t == 0) { pval[blockIdx.x] = sv[0]; pidx[blockIdx.x] = si[0]; }
}

__global__ __launch_bounds__(256) void k_pc_pick(int32_t nb, int32_t k, const double* __restrict__ pval,
                                                 const int32_t* __restrict__ pidx, int32_t* __restrict__ perm,
                                                 int32_t* __restrict__ pos, int32_t* __restrict__ col_out) {
    __shared__ double sv[kB];
    __shared__ int32_t si[kB];
    const int t = threadIdx.x;
    double bv = -HUGE_VAL;
    int32_t bi = INT_MAX;
    for (int b = t; b < nb; b += kB)
        if (better(pval[b], pidx[b], bv, bi)) { bv = pval[b]; bi = pidx[b]; }
    sv[t] = bv;
    si[t] = bi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w && better(sv[t + w], si[t + w], sv[t], si[t])) { sv[t] = sv[t + w]; si[t] = si[t + w]; }
        __syncthreads();
    }
    if (t == 0) {
        int32_t cm = si[0] == INT_MAX ? k : si[0];
        const int32_t a = perm[k], b = perm[cm];
        perm[k] = b;
        perm[cm] = a;
        pos[a] = cm;
        pos[b] = k;
        *col_out = b;
    }
}

__global__ void k_pc_scatter(const int32_t* __restrict__ colp, const int64_t* __restrict__ rowptr,
                             const int32_t* __restrict__ colind, const double* __restrict__ val,
                             const double* __restrict__ dinv, double* __restrict__ sc, int clear) {
    const int32_t col = *colp;
    const int64_t b = rowptr[col], e = rowptr[col + 1];
    for (int64_t p = b + (int64_t)blockIdx.x * kB + threadIdx.x; p < e; p += (int64_t)gridDim.x * kB) {
        const int32_t j = colind[p];
        sc[j] = clear ? 0.0 : val[p] * dinv[j];
    }
}

// w_r = D_r L(col, r) for r < k and D_k = g(col) + delta - sum_r w_r L(col, r)
__global__ __launch_bounds__(256) void k_pc_pivot(int32_t m, int32_t k, const int32_t* __restrict__ colp,
                                                  const double* __restrict__ g, const double* __restrict__ L,
                                                  double delta, double* __restrict__ D, double* __restrict__ w) {
    __shared__ double red[kB];
    const int32_t col = *colp;
    double acc = 0.0;
    for (int32_t r = threadIdx.x; r < k; r += kB) {
        const double l = L[(int64_t)r * m + col];
        const double wr = D[r] * l;
        w[r] = wr;
        acc += wr * l;
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) D[k] = (g[col] + delta) - s;
}

// column k of L over all rows: 1 at the pivot, 0 at earlier pivots, else
// (g_j - sum_r L(j,r) w_r) / D_k; the Schur diagonal loses D_k L(j,k)^2
__global__ __launch_bounds__(256) void k_pc_column(int32_t m, int32_t k, const int32_t* __restrict__ colp,
                                                   const int32_t* __restrict__ pos, const double* __restrict__ g,
                                                   const double* __restrict__ wg, const double* __restrict__ D,
                                                   double* __restrict__ L, double* __restrict__ S) {
    __shared__ double w[kPcgMaxK];
    for (int r = threadIdx.x; r < k; r += kB) w[r] = wg[r];
    __syncthreads();
    const int32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= m) return;
    const int32_t col = *colp;
    double* Lk = L + (int64_t)k * m;
    if (j == col) { Lk[j] = 1.0; return; }
    if (pos[j] < k) { Lk[j] = 0.0; return; }
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int32_t r = 0;
    for (; r + 3 < k; r += 4) {
        a0 += L[(int64_t)r * m + j] * w[r];
        a1 += L[(int64_t)(r + 1) * m + j] * w[r + 1];
        a2 += L[(int64_t)(r + 2) * m + j] * w[r + 2];
        a3 += L[(int64_t)(r + 3) * m + j] * w[r + 3];
    }
    for (; r < k; ++r) a0 += L[(int64_t)r * m + j] * w[r];
    const double u = g[j] - ((a0 + a1) + (a2 + a3));
    const double dk = D[k];
    const double l = u / dk;
    Lk[j] = l;
    S[j] -= dk * l * l;
}

// L11 (row-major K x K, rows = pivots in order) and X = L11^-1 by columns:
// thread c solves L11 x = e_c forward (unit diagonal)
__global__ void k_pc_l11_gather(int32_t m, int32_t K, const int32_t* __restrict__ perm,
                                const double* __restrict__ L, double* __restrict__ L11) {
    const int64_t idx = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (idx >= (int64_t)K * K) return;
    const int32_t i = (int32_t)(idx / K), c = (int32_t)(idx % K);
    L11[idx] = (c <= i) ? L[(int64_t)c * m + perm[i]] : 0.0;
}

// Column c of X = L11^-1 (unit lower) by forward substitution, a wave per
// column: the column lives in LDS, row i's dot over t in [c, i) has the
// lanes over t (coalesced reads of row i of L11) and a fixed reduction.
__global__ __launch_bounds__(256) void k_pc_l11_inv(int32_t K, const double* __restrict__ L11,
                                                    double* __restrict__ X) {
    __shared__ double xs[4][kPcgMaxK];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int32_t c = blockIdx.x * 4 + wave;
    if (c >= K) return;
    double* x = xs[wave];
    for (int32_t i = lane; i < K; i += 64) x[i] = (i == c) ? 1.0 : 0.0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int32_t i = c + 1; i < K; ++i) {
        const double* Li = L11 + (int64_t)i * K;
        double a0 = 0.0, a1 = 0.0;
        int32_t t = c + lane;
        for (; t + 64 < i; t += 128) {
            a0 += Li[t] * x[t];
            a1 += Li[t + 64] * x[t + 64];
        }
        if (t < i) a0 += Li[t] * x[t];
        const double s = wave_sum(a0 + a1);
        if (lane == 0) x[i] = -s;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    for (int32_t i = lane; i < K; i += 64) X[(int64_t)i * K + c] = x[i];
}

__global__ void k_vdiv(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                       double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n) out[i] = a[i] / b[i];
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ x, float* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n) y[i] = (float)x[i];
}

// ------------------------------------------------------ preconditioner apply
// y1 = L11^-1 r(pivots)
__global__ __launch_bounds__(256) void k_pca_y1(int32_t K, const int32_t* __restrict__ perm,
                                                const double* __restrict__ X, const double* __restrict__ r,
                                                double* __restrict__ y1) {
    // a wave per row i of X (lanes over the columns, coalesced)
    const int32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= K) return;
    double a0 = 0.0, a1 = 0.0;
    int32_t c = lane;
    for (; c + 64 <= i; c += 128) {
        a0 += X[(int64_t)i * K + c] * r[perm[c]];
        a1 += X[(int64_t)i * K + c + 64] * r[perm[c + 64]];
    }
    if (c <= i) a0 += X[(int64_t)i * K + c] * r[perm[c]];
    const double s = wave_sum(a0 + a1);
    if (lane == 0) y1[i] = s;
}

// z_j = (r_j - sum_c L(j,c) y1_c) / S_j off the pivots, 0 on them; the
// factor in FP32 (the default) or FP64, the sums in FP64
template <typename T>
__global__ __launch_bounds__(256) void k_pca_t2(int32_t m, int32_t K, const int32_t* __restrict__ pos,
                                                const T* __restrict__ L32, const double* __restrict__ y1g,
                                                const double* __restrict__ S, const double* __restrict__ r,
                                                double* __restrict__ z) {
    __shared__ double y1[kPcgMaxK];
    for (int c = threadIdx.x; c < K; c += kB) y1[c] = y1g[c];
    __syncthreads();
    const int32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= m) return;
    if (pos[j] < K) { z[j] = 0.0; return; }
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int32_t c = 0;
    for (; c + 3 < K; c += 4) {
        a0 += (double)L32[(int64_t)c * m + j] * y1[c];
        a1 += (double)L32[(int64_t)(c + 1) * m + j] * y1[c + 1];
        a2 += (double)L32[(int64_t)(c + 2) * m + j] * y1[c + 2];
        a3 += (double)L32[(int64_t)(c + 3) * m + j] * y1[c + 3];
    }
    for (; c < K; ++c) a0 += (double)L32[(int64_t)c * m + j] * y1[c];
    z[j] = (r[j] - ((a0 + a1) + (a2 + a3))) / S[j];
}

// ltpart[c * nb + b] = sum over the kLtRows rows of block b of L(j,c) z_j
// (z is 0 on the pivots); a wave per column, 32 rows per lane in flight
template <typename T>
__global__ __launch_bounds__(256) void k_pca_lt(int32_t m, int32_t K, const T* __restrict__ L32,
                                                const double* __restrict__ z, double* __restrict__ ltpart) {
    __shared__ double zs[kLtRows];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * kLtRows;
    for (int t = threadIdx.x; t < kLtRows; t += kB) zs[t] = (r0 + t < m) ? z[r0 + t] : 0.0;
    __syncthreads();
    const int64_t last = (int64_t)m - 1;
    for (int c = wave; c < K; c += 4) {
        const T* Lc = L32 + (int64_t)c * m;
        double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int u = 0; u < kLtRows / 64; ++u) {
            const int64_t j = r0 + lane + 64 * u;
            a[u & 3] += (double)Lc[j < m ? j : last] * zs[lane + 64 * u];
        }
        const double s = wave_sum((a[0] + a[1]) + (a[2] + a[3]));
        if (lane == 0) ltpart[(int64_t)c * gridDim.x + blockIdx.x] = s;
    }
}

// w_c = y1_c / S(pivot c) - sum_b ltpart[c][b]; a wave per column c
__global__ __launch_bounds__(256) void k_pca_w(int32_t K, int32_t nb, const int32_t* __restrict__ perm,
                                               const double* __restrict__ y1, const double* __restrict__ S,
                                               const double* __restrict__ ltpart, double* __restrict__ w) {
    const int32_t c = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= K) return;
    double a0 = 0.0, a1 = 0.0;
    int32_t b = lane;
    for (; b + 64 < nb; b += 128) {
        a0 += ltpart[(int64_t)c * nb + b];
        a1 += ltpart[(int64_t)c * nb + b + 64];
    }
    if (b < nb) a0 += ltpart[(int64_t)c * nb + b];
    const double s = wave_sum(a0 + a1);
    if (lane == 0) w[c] = y1[c] / S[perm[c]] - s;
}

// z(pivot i) = sum_{c >= i} X[c][i] w_c   (L11^-T w): a thread per i, the
// reads of a row c coalesced across the threads, 8 terms in flight
__global__ void k_pca_final(int32_t K, const int32_t* __restrict__ perm, const double* __restrict__ X,
                            const double* __restrict__ w, double* __restrict__ z) {
    const int32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= K) return;
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int32_t c = i;
    for (; c + 7 < K; c += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] += X[(int64_t)(c + u) * K + i] * w[c + u];
    }
    for (; c < K; ++c) a[0] += X[(int64_t)c * K + i] * w[c];
    z[perm[i]] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

// ---- conjugate gradients with the scalars on the device ----
// s[kCgRR] = r.r, s[kCgPQ] = p.q, s[kCgRZ + 0/1] = r.z of alternate
// iterations, s[kCgState] = 0 running / 1 converged / 2 breakdown,
// s[kCgIters] = iterations started.  Every kernel of an iteration returns at
// once when the state is not 0, so iterations the host enqueued after the
// stop change nothing.  The block partition of the reductions is that of
// k_dot_partial (kDotPartials blocks, grid-stride), and products and sums are
// not contracted into FMAs, as in those kernels, so every value is the one
// the separate dot and axpy kernels form.
__device__ inline bool cg_stopped(const double* s) { return s[kCgState] != 0.0; }

__device__ inline void block_sum_to(double acc, double* __restrict__ dst) {
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *dst = red[0];
}

__global__ __launch_bounds__(256) void k_cg_dot_partial(const double* __restrict__ a, const double* __restrict__ b,
                                                        int64_t n, double* __restrict__ part,
                                                        const double* __restrict__ s) {
#pragma clang fp contract(off)
    if (cg_stopped(s)) return;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        acc = acc + a[i] * b[i];
    block_sum_to(acc, part + blockIdx.x);
}

// s[slot] = sum of the partials; kind 1 (p.q): the iteration has started,
// and it breaks down unless p.q > 0 and the current r.z > 0; kind 2 (r.r):
// converged when sqrt(r.r) <= target
__global__ __launch_bounds__(256) void k_cg_final(const double* __restrict__ part, int np, double* __restrict__ s,
                                                  int slot, int kind, int rz_cur, double target) {
#pragma clang fp contract(off)
    if (cg_stopped(s)) return;
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) acc = acc + part[i];
    __shared__ double tot;
    block_sum_to(acc, &tot);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double v = tot;
    s[slot] = v;
    if (kind == 1) {
        s[kCgIters] += 1.0;
        if (!(v > 0) || !(s[rz_cur] > 0)) s[kCgState] = 2.0;
    } else if (kind == 2) {
        if (sqrt(v) <= target) s[kCgState] = 1.0;
    }
}

// alpha = r.z / p.q;  x += alpha p;  r -= alpha q;  partials of r.r
__global__ __launch_bounds__(256) void k_cg_update(int64_t n, const double* __restrict__ s, int rz_cur,
                                                   const double* __restrict__ p, const double* __restrict__ q,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ part) {
#pragma clang fp contract(off)
    if (cg_stopped(s)) return;
    const double alpha = s[rz_cur] / s[kCgPQ];
    const double nalpha = -alpha;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        x[i] = x[i] + alpha * p[i];
        const double ri = r[i] + nalpha * q[i];
        r[i] = ri;
        acc = acc + ri * ri;
    }
    block_sum_to(acc, part + blockIdx.x);
}

// beta = r.z(new) / r.z(old);  p = z + beta p
__global__ void k_cg_xpby(int64_t n, const double* __restrict__ s, int rz_cur, int rz_next,
                          const double* __restrict__ z, double* __restrict__ p) {
#pragma clang fp contract(off)
    if (cg_stopped(s)) return;
    const double beta = s[rz_next] / s[rz_cur];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

__global__ void k_cg_init(double* __restrict__ s, double rr, double rz) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        s[kCgRR] = rr; s[kCgPQ] = 0.0; s[kCgRZ] = rz; s[kCgRZ + 1] = 0.0; s[kCgState] = 0.0; s[kCgIters] = 0.0;
    }
}

}  // namespace

void build_row_split(int32_t m, const std::vector<int64_t>& rowptr, std::vector<RowChunk>& chunks,
                     std::vector<int32_t>& hrow, std::vector<int32_t>& hptr, std::vector<int32_t>& mrow) {
    chunks.clear();
    hrow.clear();
    mrow.clear();
    hptr.assign(1, 0);
    for (int32_t i = 0; i < m; ++i) {
        const int64_t b = rowptr[i], e = rowptr[i + 1];
        if (e - b <= kLightRow) continue;
        if (e - b <= kHeavyRow) { mrow.push_back(i); continue; }
        hrow.push_back(i);
        for (int64_t p = b; p < e; p += kRowChunk) chunks.push_back({i, p, std::min(e, p + kRowChunk)});
        hptr.push_back((int32_t)chunks.size());
    }
}

void launch_csr_split(const RowSplitDev& rs, const int64_t* rowptr, const int32_t* colind, const double* val,
                      const double* x, double* y, bool square, hipStream_t st) {
    if (rs.m <= 0) return;
    if (rs.nmed) {
        const int nbm = (rs.nmed + 3) / 4;
        if (square) help::launch_kernel(k_csr_medium<true>, nbm, kB, 0, st, rs.nmed, rs.mrow, rowptr, colind, val, x, y);
        else help::launch_kernel(k_csr_medium<false>, nbm, kB, 0, st, rs.nmed, rs.mrow, rowptr, colind, val, x, y);
    }
    if (square) {
        help::launch_kernel(k_csr_light<true>, grid_for(rs.m), kB, 0, st, rs.m, rowptr, colind, val, x, y);
        if (rs.nchunk) help::launch_kernel(k_csr_chunk<true>, rs.nchunk, kB, 0, st, rs.chunks, colind, val, x, rs.part);
    } else {
        help::launch_kernel(k_csr_light<false>, grid_for(rs.m), kB, 0, st, rs.m, rowptr, colind, val, x, y);
        if (rs.nchunk) help::launch_kernel(k_csr_chunk<false>, rs.nchunk, kB, 0, st, rs.chunks, colind, val, x, rs.part);
    }
    if (rs.nheavy) help::launch_kernel(k_csr_heavy, (rs.nheavy + 3) / 4, kB, 0, st, rs.nheavy, rs.hrow, rs.hptr, rs.part, y);
}

void launch_csc_t_scaled(int32_t n, const int64_t* colptr, const int32_t* rowind, const double* val,
                         const double* dinv, const double* x, double* t, hipStream_t st) {
    if (n > 0) help::launch_kernel(k_csc_ts, grid_for(n), kB, 0, st, n, colptr, rowind, val, dinv, x, t);
}

void launch_pc_argmax(int32_t m, int32_t k, const double* S, double* pval, int32_t* pidx, int32_t* perm,
                      int32_t* pos, int32_t* col_out, hipStream_t st) {
    const int nb = (int)std::min<int64_t>(kArgmaxBlocks, std::max<int64_t>(1, ((int64_t)m - k + kB - 1) / kB));
    help::launch_kernel(k_pc_argmax_part, nb, kB, 0, st, m, k, perm, S, pval, pidx);
    help::launch_kernel(k_pc_pick, 1, kB, 0, st, nb, k, pval, pidx, perm, pos, col_out);
}

void launch_pc_scatter(const int32_t* col, int32_t max_row_nnz, const int64_t* rowptr, const int32_t* colind,
                       const double* val, const double* dinv, double* sc, bool clear, hipStream_t st) {
    const int nb = (int)std::min<int64_t>(2048, std::max<int64_t>(1, ((int64_t)max_row_nnz + kB - 1) / kB));
    help::launch_kernel(k_pc_scatter, nb, kB, 0, st, col, rowptr, colind, val, dinv, sc, clear ? 1 : 0);
}

void launch_pc_pivot(int32_t m, int32_t k, const int32_t* col, const double* g, const double* L, double delta,
                     double* D, double* w, hipStream_t st) {
    help::launch_kernel(k_pc_pivot, 1, kB, 0, st, m, k, col, g, L, delta, D, w);
}

void launch_pc_column(int32_t m, int32_t k, const int32_t* col, const int32_t* pos, const double* g,
                      const double* w, const double* D, double* L, double* S, hipStream_t st) {
    help::launch_kernel(k_pc_column, grid_for(m), kB, 0, st, m, k, col, pos, g, w, D, L, S);
}

void launch_pc_l11(int32_t m, int32_t K, const int32_t* perm, const double* L, double* L11, double* X,
                   hipStream_t st) {
    if (K <= 0) return;
    help::launch_kernel(k_pc_l11_gather, grid_for((int64_t)K * K), kB, 0, st, m, K, perm, L, L11);
    help::launch_kernel(k_pc_l11_inv, (K + 3) / 4, kB, 0, st, K, L11, X);
}

void launch_vdiv(int64_t n, const double* a, const double* b, double* out, hipStream_t st) {
    if (n > 0) help::launch_kernel(k_vdiv, grid_for(n), kB, 0, st, n, a, b, out);
}

void launch_to_f32(int64_t n, const double* x, float* y, hipStream_t st) {
    if (n > 0) help::launch_kernel(k_to_f32, grid_for(n), kB, 0, st, n, x, y);
}

template <typename T>
void launch_pc_apply(int32_t m, int32_t K, const int32_t* perm, const int32_t* pos, const T* L32,
                     const double* X, const double* S, const double* r, double* y1, double* ltpart, double* z,
                     hipStream_t st) {
    // K = 0: z = r / S
    const int nb = (int)(((int64_t)m + kLtRows - 1) / kLtRows);
    if (K > 0) help::launch_kernel(k_pca_y1, (K + 3) / 4, kB, 0, st, K, perm, X, r, y1);
    help::launch_kernel(k_pca_t2<T>, grid_for(m), kB, 0, st, m, K, pos, L32, y1, S, r, z);
    if (K > 0) {
        help::launch_kernel(k_pca_lt<T>, nb, kB, 0, st, m, K, L32, z, ltpart);
        double* w = y1 + kPcgMaxK;
        help::launch_kernel(k_pca_w, (K + 3) / 4, kB, 0, st, K, nb, perm, y1, S, ltpart, w);
        help::launch_kernel(k_pca_final, grid_for(K), kB, 0, st, K, perm, X, w, z);
    }
}

template void launch_pc_apply<float>(int32_t, int32_t, const int32_t*, const int32_t*, const float*,
                                      const double*, const double*, const double*, double*, double*, double*,
                                      hipStream_t);
template void launch_pc_apply<double>(int32_t, int32_t, const int32_t*, const int32_t*, const double*,
                                       const double*, const double*, const double*, double*, double*, double*,
                                       hipStream_t);

void launch_cg_init(double* s, double rr, double rz, hipStream_t st) {
    help::launch_kernel(k_cg_init, 1, 64, 0, st, s, rr, rz);
}

// one CG iteration after the matrix product q = M p and before the
// preconditioner: p.q (breakdown test), the x/r update with r.r
// (convergence test); then, after z = P^-1 r, r.z and the new direction
void launch_cg_step1(int64_t n, double* s, int rz_cur, double target, const double* p, const double* q, double* x,
                     double* r, double* part, hipStream_t st) {
    const int nb = (int)std::min<int64_t>(kDotPartials, std::max<int64_t>(1, (n + 255) / 256));
    help::launch_kernel(k_cg_dot_partial, nb, 256, 0, st, p, q, n, part, (const double*)s);
    help::launch_kernel(k_cg_final, 1, 256, 0, st, (const double*)part, nb, s, (int)kCgPQ, 1, rz_cur, target);
    help::launch_kernel(k_cg_update, nb, 256, 0, st, n, (const double*)s, rz_cur, p, q, x, r, part);
    help::launch_kernel(k_cg_final, 1, 256, 0, st, (const double*)part, nb, s, (int)kCgRR, 2, rz_cur, target);
}
void launch_cg_step2(int64_t n, double* s, int rz_cur, int rz_next, const double* r, const double* z, double* p,
                     double* part, hipStream_t st) {
    const int nb = (int)std::min<int64_t>(kDotPartials, std::max<int64_t>(1, (n + 255) / 256));
    help::launch_kernel(k_cg_dot_partial, nb, 256, 0, st, r, z, n, part, (const double*)s);
    help::launch_kernel(k_cg_final, 1, 256, 0, st, (const double*)part, nb, s, rz_next, 0, rz_cur, 0.0);
    help::launch_kernel(k_cg_xpby, (unsigned)((n + 255) / 256), 256, 0, st, n, (const double*)s, rz_cur, rz_next, z, p);
}

}  // namespace help
