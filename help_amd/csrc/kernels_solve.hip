// Supernodal triangular solves L y = b and L^T x = y, level by level.
//
// Forward (leaves to root): each supernode gathers its children's update
// vectors (the multifrontal analogue of extend-add, children in fixed
// order), solves with its diagonal block L11 and forms its own update vector
// W_s = L21 y_s + (children's rows beyond s).  Backward (root to leaves):
// each supernode subtracts L21^T x(rows) and solves with L11^T; the rows are
// ancestors, already final.  Within L11 the solve walks 64-column blocks and
// applies the stored inverse of each diagonal block as a parallel product,
// so a block costs a few barriers instead of 64 dependent steps.  Vectors
// are FP64 whatever the factor type.  Big supernodes hand the L21 products
// to separate many-workgroup kernels.
//
// The solves are bound by load latency, not bandwidth: a supernode is one
// dependent chain, so every inner loop keeps several independent loads in
// flight (multiple accumulators, summed in a fixed order).  Supernodes with
// at most 64 rows (the large majority, near the leaves) take one wave each,
// four to a workgroup, with the rows on the lanes and no workgroup barrier.
#include "device.h"

namespace help {

namespace {

constexpr int kMaxLdsCols = kMaxSnCols;   // x segment of a supernode held in LDS
constexpr int kMaxLdsRows = 2048;         // gathered x(rows) held in LDS by k_bwd
constexpr int kMaxGemvRows = 4096;        // gathered x(rows) held in LDS by k_bwd_gemv

__device__ inline const void* dinv_block(const DevSym& sym, const void* base, int32_t s, int32_t jb,
                                         int esz) {
    return (const char*)base + (sym.dinvoff[s] + (int64_t)(jb / 64) * 64 * 64) * esz;
}

// orders this wave's LDS accesses before and after (lanes exchange data
// through LDS without a workgroup barrier)
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// sum_{k<cnt} a[k*stride] * v[k] with two accumulators
template <typename T>
__device__ inline double dot_strided(const T* __restrict__ a, int64_t stride, const double* v, int32_t cnt) {
    double a0 = 0.0, a1 = 0.0;
    int32_t k = 0;
#pragma unroll 4
    for (; k + 1 < cnt; k += 2) {
        a0 += (double)a[(int64_t)k * stride] * v[k];
        a1 += (double)a[(int64_t)(k + 1) * stride] * v[k + 1];
    }
    if (k < cnt) a0 += (double)a[(int64_t)k * stride] * v[k];
    return a0 + a1;
}

// dst[i] = src[idx[i]] for i < cnt, four per thread in flight (the index
// loads first, then the value loads)
__device__ inline void gather_idx(double* dst, const double* __restrict__ src, const int32_t* __restrict__ idx,
                                  int32_t cnt, int tid, int nt) {
    for (int32_t i0 = tid; i0 < cnt; i0 += 4 * nt) {
        int32_t ii[4];
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) ii[q] = idx[min(i0 + q * nt, cnt - 1)];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = src[ii[q]];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (i0 + q * nt < cnt) dst[i0 + q * nt] = v[q];
    }
}

// One child's update vector into the supernode's gathered right-hand side
// (entries rr < n: g[rr] -= v for rr < lim) and its own update vector
// (rr >= n: Ws[rr - n] += v, when `upd`), four entries per thread in flight.
// The relative map of a child is injective, so the entries are independent.
__device__ inline void gather_child(double* g, int32_t lim, double* __restrict__ Ws, bool upd, int32_t n,
                                    int32_t cu, const int32_t* __restrict__ rel, const double* __restrict__ Wc,
                                    int32_t chc, int tid, int nt) {
    for (int32_t u0 = tid; u0 < chc; u0 += 4 * nt) {
        int32_t rr[4];
        double v[4], wv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t u = min(u0 + q * nt, chc - 1);
            rr[q] = rel[u];
            v[q] = Wc[u];
        }
        if (upd && cu > 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) wv[q] = Ws[min(max(rr[q] - n, 0), cu - 1)];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (u0 + q * nt >= chc) continue;
            if (rr[q] < n) {
                if (rr[q] < lim) g[rr[q]] -= v[q];
            } else if (upd) {
                Ws[rr[q] - n] = wv[q] + v[q];
            }
        }
    }
}

// ------------------------------------------------------------- tiny (c <= 64)
// The forward solve reads the right-hand side b and writes y; the backward
// solve reads y and writes x (x may be b's storage).  Keeping input and
// output apart lets the blocks of one wide supernode run in parallel
// without reading entries another block is writing.
template <typename T, int G>
__device__ void fwd_tiny(int32_t s, const DevSym& sym, const T* __restrict__ L, const T* __restrict__ Dinv,
                         const double* __restrict__ b, double* __restrict__ y, double* __restrict__ W,
                         double* xw) {
    const int lane = threadIdx.x & (G - 1);   // G lanes per supernode (c <= G)
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const T* Lp = L + sym.loff[s];
    xw[lane] = (lane < n) ? b[f + lane] : 0.0;
    wave_sync();
    for (int32_t t = sym.childptr[s]; t < sym.childptr[s + 1]; ++t) {
        const int32_t ch = sym.child[t];
        const int32_t chc = sym.snrow[ch] - (sym.sfirst[ch + 1] - sym.sfirst[ch]);
        if (lane < chc) {
            const int32_t rr = sym.relmap[sym.reloff[ch] + lane];
            const double v = W[sym.reloff[ch] + lane];
            if (rr < n) xw[rr] -= v;
            else xw[rr] += v;
        }
        wave_sync();
    }
    // one diagonal block: x1 = Dinv x1
    const T* Db = (const T*)dinv_block(sym, Dinv, s, 0, sizeof(T));
    const double v = (lane < n) ? dot_strided<T>(Db + lane, n, xw, n) : 0.0;
    wave_sync();
    if (lane < n) xw[lane] = v;
    wave_sync();
    if (lane < n) {
        y[f + lane] = v;
    } else if (lane < c) {
        const double a = dot_strided<T>(Lp + lane, c, xw, n);
        W[sym.reloff[s] + lane - n] = xw[lane] + a;
    }
}

template <typename T, int G>
__device__ void bwd_tiny(int32_t s, const DevSym& sym, const T* __restrict__ L, const T* __restrict__ Dinv,
                         const double* __restrict__ y, double* __restrict__ x, double* xw) {
    const int lane = threadIdx.x & (G - 1);   // G lanes per supernode (c <= G)
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const T* Lp = L + sym.loff[s];
    const int32_t* rows = sym.rowidx + sym.rowoff[s];
    xw[lane] = (lane >= n && lane < c) ? x[rows[lane]] : 0.0;
    wave_sync();
    double v = 0.0;
    if (lane < n) {
        const T* col = Lp + (int64_t)lane * c;
        v = y[f + lane] - dot_strided<T>(col + n, 1, xw + n, c - n);
    }
    wave_sync();
    if (lane < n) xw[lane] = v;
    wave_sync();
    if (lane < n) {
        const T* Db = (const T*)dinv_block(sym, Dinv, s, 0, sizeof(T));
        x[f + lane] = dot_strided<T>(Db + (int64_t)lane * n, 1, xw, n);
    }
}

// ------------------------------------------------- one workgroup per supernode
// x1 = Dinv x1 for the 64-column block at jb: thread (r, part) sums the
// columns k = part (mod 4), the four parts combine in a fixed order.
template <typename T, bool Trans>
__device__ inline void apply_dinv(const T* __restrict__ Db, int32_t nb, double* xs, double (*red)[64]) {
    const int r = threadIdx.x & 63, part = threadIdx.x >> 6;
    double a0 = 0.0, a1 = 0.0;
    if (nb == 64) {   // all loads issued together
#pragma unroll
        for (int u = 0; u < 16; u += 2) {
            const int k0 = part + 4 * u, k1 = k0 + 4;
            a0 += (double)(Trans ? Db[r * 64 + k0] : Db[k0 * 64 + r]) * xs[k0];
            a1 += (double)(Trans ? Db[r * 64 + k1] : Db[k1 * 64 + r]) * xs[k1];
        }
    } else if (r < nb) {
        for (int k = part; k < nb; k += 4)
            a0 += (double)(Trans ? Db[r * nb + k] : Db[k * nb + r]) * xs[k];
    }
    red[part][r] = a0 + a1;
    __syncthreads();
    if (part == 0 && r < nb) xs[r] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
    __syncthreads();
}

// Forward for a supernode that is not wide: gather the children's update
// vectors, walk the 64-column blocks with the diagonal-block inverses and
// (big[s] == 0) apply L21 here; big supernodes leave L21 to k_fwd_gemv.
template <typename T>
__device__ void fwd_sn(int32_t s, const DevSym& sym, const int8_t* __restrict__ big, const T* __restrict__ L,
                       const T* __restrict__ Dinv, const double* __restrict__ b, double* __restrict__ y,
                       double* __restrict__ W, double* xs, double (*red)[64]) {
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const int32_t cu = c - n;
    const T* Lp = L + sym.loff[s];
    double* Ws = W + sym.reloff[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int32_t t = tid; t < cu; t += nt) Ws[t] = 0.0;
    for (int32_t t = tid; t < n; t += nt) xs[t] = b[f + t];
    __syncthreads();
    for (int32_t t = sym.childptr[s]; t < sym.childptr[s + 1]; ++t) {
        const int32_t ch = sym.child[t];
        const int32_t chc = sym.snrow[ch] - (sym.sfirst[ch + 1] - sym.sfirst[ch]);
        const int32_t* rel = sym.relmap + sym.reloff[ch];
        const double* Wc = W + sym.reloff[ch];
        gather_child(xs, n, Ws, true, n, cu, rel, Wc, chc, tid, nt);
        __syncthreads();
    }
    const int32_t rend = big[s] ? n : c;
    for (int32_t jb = 0; jb < n; jb += 64) {
        const int32_t nb = min(64, n - jb);
        apply_dinv<T, false>((const T*)dinv_block(sym, Dinv, s, jb, sizeof(T)), nb, xs + jb, red);
        // rows below the block: one row per thread, 64 loads in flight
        const double* xv = xs + jb;
        for (int32_t i = jb + nb + tid; i < rend; i += nt) {
            const T* lp = Lp + (int64_t)jb * c + i;
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
            if (nb == 64) {
#pragma unroll
                for (int j = 0; j < 64; j += 4) {
                    a0 += (double)lp[(int64_t)j * c] * xv[j];
                    a1 += (double)lp[(int64_t)(j + 1) * c] * xv[j + 1];
                    a2 += (double)lp[(int64_t)(j + 2) * c] * xv[j + 2];
                    a3 += (double)lp[(int64_t)(j + 3) * c] * xv[j + 3];
                }
            } else {
                a0 = dot_strided<T>(lp, c, xv, nb);
            }
            const double a = (a0 + a1) + (a2 + a3);
            if (i < n) xs[i] -= a;
            else Ws[i - n] += a;
        }
        __syncthreads();
    }
    for (int32_t t = tid; t < n; t += nt) y[f + t] = xs[t];
}

// Forward, wide supernode, 64-row block i with the explicit inverse:
// y[f+64i+r] = sum_{j<=i} sum_q X_ij[r][q] g[64j+q] (X_ii = Dinv_i), where g
// is the supernode's gathered right-hand side.  Every block gathers the g
// entries it needs itself (children in order, the same operations as the
// gather of fwd_sn); block 0 also starts the update vector W_s.  Lanes over
// r, wave w takes q = 16w..16w+15 of every block, the four partial sums
// combine in a fixed order.
template <typename T>
__device__ void fwd_xinv(const SolveTask tk, const DevSym& sym, const T* __restrict__ Dinv,
                         const T* __restrict__ X, const double* __restrict__ b, double* __restrict__ y,
                         double* __restrict__ W, double* ys, double (*red)[64]) {
    const int32_t s = tk.s, i = tk.blk;
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t cu = sym.snrow[s] - n;
    double* Ws = W + sym.reloff[s];
    const int tid = threadIdx.x, r = tid & 63, w = tid >> 6;
    const int32_t hi = min(n, 64 * (i + 1));
    if (i == 0)
        for (int32_t t = tid; t < cu; t += 256) Ws[t] = 0.0;
    for (int32_t t = tid; t < hi; t += 256) ys[t] = b[f + t];
    __syncthreads();
    for (int32_t t = sym.childptr[s]; t < sym.childptr[s + 1]; ++t) {
        const int32_t ch = sym.child[t];
        const int32_t chc = sym.snrow[ch] - (sym.sfirst[ch + 1] - sym.sfirst[ch]);
        const int32_t* rel = sym.relmap + sym.reloff[ch];
        const double* Wc = W + sym.reloff[ch];
        gather_child(ys, hi, Ws, i == 0, n, cu, rel, Wc, chc, tid, 256);
        __syncthreads();
    }
    // the 16 loads of a block are staged in registers before the products
    // (one load in flight at a time otherwise), and masked terms multiply
    // by zero instead of branching around their loads
    double a0 = 0.0, a1 = 0.0;
    int32_t j = 0;
    for (; j + 1 < i; j += 2) {   // two blocks' loads in flight, summed in block order
        const T* Xb0 = X + xblock_offset(sym.xoff, s, i, j) + r;
        const T* Xb1 = X + xblock_offset(sym.xoff, s, i, j + 1) + r;
        T xv0[16], xv1[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) xv0[u] = Xb0[64 * (16 * w + u)];
#pragma unroll
        for (int u = 0; u < 16; ++u) xv1[u] = Xb1[64 * (16 * w + u)];
        const double* yj0 = ys + 64 * j + 16 * w;
        const double* yj1 = yj0 + 64;
#pragma unroll
        for (int u = 0; u < 16; u += 2) {
            a0 += (double)xv0[u] * yj0[u];
            a1 += (double)xv0[u + 1] * yj0[u + 1];
        }
#pragma unroll
        for (int u = 0; u < 16; u += 2) {
            a0 += (double)xv1[u] * yj1[u];
            a1 += (double)xv1[u + 1] * yj1[u + 1];
        }
    }
    if (j < i) {
        const T* Xb = X + xblock_offset(sym.xoff, s, i, j) + r;
        const double* yj = ys + 64 * j + 16 * w;
        T xv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) xv[u] = Xb[64 * (16 * w + u)];
#pragma unroll
        for (int u = 0; u < 16; u += 2) {
            a0 += (double)xv[u] * yj[u];
            a1 += (double)xv[u + 1] * yj[u + 1];
        }
    }
    {
        const int32_t nb = min(64, n - 64 * i);
        const T* Db = Dinv + sym.dinvoff[s] + (int64_t)i * 4096;
        const int rr = min(r, nb - 1);
        const double* yi = ys + 64 * i;
        T dv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) dv[u] = Db[min(16 * w + u, nb - 1) * nb + rr];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = 16 * w + u;
            const double yv = (r < nb && q < nb) ? yi[min(q, nb - 1)] : 0.0;
            a0 += (double)dv[u] * yv;
        }
    }
    red[w][r] = a0 + a1;
    __syncthreads();
    if (w == 0 && 64 * i + r < n) y[f + 64 * i + r] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
}

// One launch per level: blocks [0, ntb) solve four tiny supernodes each
// (list[0, ntiny)), the next nrest blocks one supernode each
// (list[ntiny, ntiny + nrest)), the last nx blocks one 64-row block of a
// wide supernode each.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_level(const int32_t* __restrict__ list, int32_t ntiny, int32_t nrest,
                                                   const SolveTask* __restrict__ xt, DevSym sym,
                                                   const int8_t* __restrict__ big, const T* __restrict__ L,
                                                   const T* __restrict__ Dinv, const T* __restrict__ X,
                                                   const double* __restrict__ b, double* __restrict__ y,
                                                   double* __restrict__ W, RhsStride rs) {
    __shared__ double xs[kMaxLdsCols];
    __shared__ double red[4][64];
    b += blockIdx.y * rs.v; y += blockIdx.y * rs.v; W += blockIdx.y * rs.w;
    const int32_t ntb = (ntiny + 3) / 4;
    const int32_t bid = blockIdx.x;
    if (bid < ntb) {
        const int wave = threadIdx.x >> 6;
        const int32_t idx = bid * 4 + wave;
        if (idx < ntiny) fwd_tiny<T, 64>(list[idx], sym, L, Dinv, b, y, W, red[wave]);
    } else if (bid < ntb + nrest) {
        fwd_sn<T>(list[ntiny + bid - ntb], sym, big, L, Dinv, b, y, W, xs, red);
    } else {
        fwd_xinv<T>(xt[bid - ntb - nrest], sym, Dinv, X, b, y, W, xs, red);
    }
}

// levels made of tiny supernodes only (the leaves: many of them, and a
// kernel with a small register and LDS footprint)
template <typename T, int G>
__global__ __launch_bounds__(256) void k_fwd_tiny(const int32_t* __restrict__ list, int32_t count, DevSym sym,
                                                  const T* __restrict__ L, const T* __restrict__ Dinv,
                                                  const double* __restrict__ b, double* __restrict__ y,
                                                  double* __restrict__ W, RhsStride rs) {
    __shared__ double xw[256 / G][G];
    b += blockIdx.y * rs.v; y += blockIdx.y * rs.v; W += blockIdx.y * rs.w;
    const int g = threadIdx.x / G;
    const int32_t idx = blockIdx.x * (256 / G) + g;
    if (idx < count) fwd_tiny<T, G>(list[idx], sym, L, Dinv, b, y, W, xw[g]);
}

// W(rows of L21) += L21 y_s for a 64-row block of a big supernode's L21.
// Every row sums its columns as four parts (columns j = part mod 4), each
// part as four interleaved chains a_q (columns part + 4q + 16t while the
// group start part + 16t has its 16 columns; the group tail's columns go to
// a_0 in order), a part as (a_0 + a_1) + (a_2 + a_3) and the row as
// (part 0 + part 1) + (part 2 + part 3).  One wave per chain (16 waves, the
// lanes on the rows), so a block keeps 16 chains' loads in flight.
template <typename T>
__global__ __launch_bounds__(1024) void k_fwd_gemv(const SolveTask* __restrict__ tasks, DevSym sym,
                                                   const T* __restrict__ L,
                                                   const double* __restrict__ y,
                                                   double* __restrict__ W, RhsStride rs) {
    __shared__ double xs[kMaxLdsCols];
    __shared__ double red[16][64];
    y += blockIdx.y * rs.v; W += blockIdx.y * rs.w;
    const SolveTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s;
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const T* Lp = L + sym.loff[s];
    double* Ws = W + sym.reloff[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int32_t t = tid; t < n; t += nt) xs[t] = y[f + t];
    __syncthreads();
    const int32_t r = tid & 63, ch = tid >> 6;   // chain ch = part + 4 q
    const int32_t part = ch & 3, q = ch >> 2;
    const int32_t i = n + tk.blk * 64 + r;
    double a = 0.0;
    if (i < c) {
        const T* lp = Lp + i;
        int32_t jb = part;
#pragma unroll 8
        for (; jb + 12 < n; jb += 16) {
            const int32_t j = jb + 4 * q;
            a += (double)lp[(int64_t)j * c] * xs[j];
        }
        if (q == 0)
            for (; jb < n; jb += 4) a += (double)lp[(int64_t)jb * c] * xs[jb];
    }
    red[ch][r] = a;
    __syncthreads();
    if (ch == 0 && i < c) {
        double pv[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) pv[p] = (red[p][r] + red[p + 4][r]) + (red[p + 8][r] + red[p + 12][r]);
        Ws[i - n] += (pv[0] + pv[1]) + (pv[2] + pv[3]);
    }
}

// t[f+j] = sum_{i>=n} L21(i,j) x[row_i] for 16 columns j of a big supernode
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_gemv(const SolveTask* __restrict__ tasks, DevSym sym,
                                                  const T* __restrict__ L,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ tout, RhsStride rs) {
    __shared__ double xr[kMaxGemvRows];
    x += blockIdx.y * rs.v; tout += blockIdx.y * rs.v;
    const SolveTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s;
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const int32_t cu = c - n;
    const T* Lp = L + sym.loff[s];
    const int32_t* rows = sym.rowidx + sym.rowoff[s];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const bool lds = cu <= kMaxGemvRows;
    if (lds)
        gather_idx(xr, x, rows + n, cu, tid, 256);
    __syncthreads();
    // each wave: 4 columns at once, lanes over the rows
    const int32_t j0 = tk.blk * 16 + wave * 4;
    if (j0 >= n) return;
    const int32_t nq = min(4, n - j0);
    const T* col = Lp + (int64_t)j0 * c + n;
    // columns past the supernode's last repeat its last one (loads stay
    // unconditional); their sums are not stored
    int64_t coff[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) coff[q] = (int64_t)min(q, nq - 1) * c;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
    for (int32_t i = lane; i < cu; i += 64) {
        const double xv = lds ? xr[i] : x[rows[n + i]];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] += (double)col[coff[q] + i] * xv;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double v = wave_sum(acc[q]);
        if (lane == q && q < nq) tout[f + j0 + q] = v;
    }
}

// Backward for a supernode that is not wide: x_s = L11^-T (y_s - L21^T x(rows));
// big supernodes take the L21 product t from k_bwd_gemv.
template <typename T>
__device__ void bwd_sn(int32_t s, const DevSym& sym, const int8_t* __restrict__ big, const T* __restrict__ L,
                       const T* __restrict__ Dinv, const double* __restrict__ y, const double* __restrict__ tin,
                       double* __restrict__ x, double* xs, double* xr, double (*red)[64]) {
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const int32_t cu = c - n;
    const T* Lp = L + sym.loff[s];
    const int32_t* rows = sym.rowidx + sym.rowoff[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    if (big[s]) {
        for (int32_t t = tid; t < n; t += nt) xs[t] = y[f + t] - tin[f + t];
    } else {
        const bool lds = cu <= kMaxLdsRows;
        if (lds)
            gather_idx(xr, x, rows + n, cu, tid, nt);
        __syncthreads();
        // few rows below the diagonal block: a thread per column
        if (cu <= 64) {
            for (int32_t j = tid; j < n; j += nt)
                xs[j] = y[f + j] - dot_strided<T>(Lp + (int64_t)j * c + n, 1, xr, cu);
        } else if (lds) {
            // four columns per wave at a time, their loads in flight together;
            // each column keeps its own four accumulators and combine order
            for (int32_t j0 = 4 * wave; j0 < n; j0 += 16) {
                const int32_t nq = min(4, n - j0);
                const T* col[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) col[q] = Lp + (int64_t)(j0 + min(q, nq - 1)) * c + n;
                double a0[4], a1[4], a2[4], a3[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) a0[q] = a1[q] = a2[q] = a3[q] = 0.0;
                int32_t i = lane;
                for (; i + 192 < cu; i += 256) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        a0[q] += (double)col[q][i] * xr[i];
                        a1[q] += (double)col[q][i + 64] * xr[i + 64];
                        a2[q] += (double)col[q][i + 128] * xr[i + 128];
                        a3[q] += (double)col[q][i + 192] * xr[i + 192];
                    }
                }
                for (; i < cu; i += 64) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) a0[q] += (double)col[q][i] * xr[i];
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double acc = wave_sum((a0[q] + a1[q]) + (a2[q] + a3[q]));
                    if (lane == 0 && q < nq) xs[j0 + q] = y[f + j0 + q] - acc;
                }
            }
        } else {
            for (int32_t j = wave; j < n; j += 4) {
                const T* col = Lp + (int64_t)j * c + n;
                double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
                int32_t i = lane;
                for (; i + 192 < cu; i += 256) {
                    a0 += (double)col[i] * x[rows[n + i]];
                    a1 += (double)col[i + 64] * x[rows[n + i + 64]];
                    a2 += (double)col[i + 128] * x[rows[n + i + 128]];
                    a3 += (double)col[i + 192] * x[rows[n + i + 192]];
                }
                for (; i < cu; i += 64) a0 += (double)col[i] * x[rows[n + i]];
                const double acc = wave_sum((a0 + a1) + (a2 + a3));
                if (lane == 0) xs[j] = y[f + j] - acc;
            }
        }
    }
    __syncthreads();
    const int32_t nblk = (n + 63) / 64;
    for (int32_t bi = nblk - 1; bi >= 0; --bi) {
        const int32_t jb = bi * 64;
        const int32_t nb = min(64, n - jb);
        const int32_t below = jb + nb;
        if (below < n) {
            // xs[jb+q] -= sum_{below <= i < n} L(i, jb+q) xs[i]; wave w takes
            // the 16 columns 16w.., lanes over the rows, 16 loads in flight
            // (only the last block can be partial, and nothing lies below it)
            const int32_t q0 = wave * 16;
            const T* base = Lp + (int64_t)(jb + q0) * c;
            double acc[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) acc[u] = 0.0;
#pragma unroll 2
            for (int32_t i = below + lane; i < n; i += 64) {
                const double xv = xs[i];
#pragma unroll
                for (int u = 0; u < 16; ++u) acc[u] += (double)base[(int64_t)u * c + i] * xv;
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const double v = wave_sum(acc[u]);
                if (lane == u) xs[jb + q0 + u] -= v;
            }
            __syncthreads();
        }
        apply_dinv<T, true>((const T*)dinv_block(sym, Dinv, s, jb, sizeof(T)), nb, xs + jb, red);
    }
    for (int32_t t = tid; t < n; t += nt) x[f + t] = xs[t];
}

// Backward, wide supernode, 64-column block j with the explicit inverse:
// x[f+64j+q] = sum_{i>=j} sum_r X_ij[r][q] (y - t)[f+64i+r].  Wave w takes the
// 16 columns q = 16w.., lanes over r; the sum over blocks and lanes has a
// fixed order.
template <typename T>
__device__ void bwd_xinv(const SolveTask tk, const DevSym& sym, const T* __restrict__ Dinv,
                         const T* __restrict__ X, const double* __restrict__ y, const double* __restrict__ tin,
                         double* __restrict__ x, double* ys) {
    const int32_t s = tk.s, j = tk.blk;
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t nblk = (n + 63) / 64;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int32_t t = 64 * j + tid; t < n; t += 256) ys[t] = y[f + t] - tin[f + t];
    __syncthreads();
    double acc[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) acc[u] = 0.0;
    {
        const int32_t nb = min(64, n - 64 * j);
        const T* Db = Dinv + sym.dinvoff[s] + (int64_t)j * 4096;
        const int rr = min(lane, nb - 1);
        const double yv = lane < nb ? ys[64 * j + rr] : 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = min(16 * w + u, nb - 1);
            acc[u] += (double)Db[q * nb + rr] * yv;
        }
    }
    int32_t i = j + 1;
    for (; i + 1 < nblk; i += 2) {   // two blocks' loads in flight, summed in block order
        const T* Xb0 = X + xblock_offset(sym.xoff, s, i, j) + lane + 64 * 16 * w;
        const T* Xb1 = X + xblock_offset(sym.xoff, s, i + 1, j) + lane + 64 * 16 * w;
        const int32_t nb1 = min(64, n - 64 * (i + 1));
        const double yv0 = ys[64 * i + lane];       // block i is full (i + 1 < nblk)
        const double yv1 = lane < nb1 ? ys[64 * (i + 1) + min(lane, nb1 - 1)] : 0.0;
        T xv0[16], xv1[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) xv0[u] = Xb0[64 * u];
#pragma unroll
        for (int u = 0; u < 16; ++u) xv1[u] = Xb1[64 * u];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc[u] += (double)xv0[u] * yv0;
#pragma unroll
        for (int u = 0; u < 16; ++u) acc[u] += (double)xv1[u] * yv1;
    }
    if (i < nblk) {
        const T* Xb = X + xblock_offset(sym.xoff, s, i, j) + lane + 64 * 16 * w;
        const int32_t nbi = min(64, n - 64 * i);
        const double yv = lane < nbi ? ys[64 * i + min(lane, nbi - 1)] : 0.0;
        T xv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) xv[u] = Xb[64 * u];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc[u] += (double)xv[u] * yv;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const double v = wave_sum(acc[u]);
        const int32_t q = 64 * j + 16 * w + u;
        if (lane == u && q < n) x[f + q] = v;
    }
}

// One launch per level, block kinds as in k_fwd_level.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_level(const int32_t* __restrict__ list, int32_t ntiny, int32_t nrest,
                                                   const SolveTask* __restrict__ xt, DevSym sym,
                                                   const int8_t* __restrict__ big, const T* __restrict__ L,
                                                   const T* __restrict__ Dinv, const T* __restrict__ X,
                                                   const double* __restrict__ y, const double* __restrict__ tin,
                                                   double* __restrict__ x, RhsStride rs) {
    __shared__ double xs[kMaxLdsCols];
    __shared__ double xr[kMaxLdsRows];
    __shared__ double red[4][64];
    y += blockIdx.y * rs.v; tin += blockIdx.y * rs.v; x += blockIdx.y * rs.v;
    const int32_t ntb = (ntiny + 3) / 4;
    const int32_t bid = blockIdx.x;
    if (bid < ntb) {
        const int wave = threadIdx.x >> 6;
        const int32_t idx = bid * 4 + wave;
        if (idx < ntiny) bwd_tiny<T, 64>(list[idx], sym, L, Dinv, y, x, red[wave]);
    } else if (bid < ntb + nrest) {
        bwd_sn<T>(list[ntiny + bid - ntb], sym, big, L, Dinv, y, tin, x, xs, xr, red);
    } else {
        bwd_xinv<T>(xt[bid - ntb - nrest], sym, Dinv, X, y, tin, x, xs);
    }
}

template <typename T, int G>
__global__ __launch_bounds__(256) void k_bwd_tiny(const int32_t* __restrict__ list, int32_t count, DevSym sym,
                                                  const T* __restrict__ L, const T* __restrict__ Dinv,
                                                  const double* __restrict__ y, double* __restrict__ x,
                                                  RhsStride rs) {
    __shared__ double xw[256 / G][G];
    y += blockIdx.y * rs.v; x += blockIdx.y * rs.v;
    const int g = threadIdx.x / G;
    const int32_t idx = blockIdx.x * (256 / G) + g;
    if (idx < count) bwd_tiny<T, G>(list[idx], sym, L, Dinv, y, x, xw[g]);
}

}  // namespace

template <typename T>
void launch_fwd_level(const int32_t* list, int32_t count, int32_t ntiny, const SolveTask* xt, int32_t nx,
                      DevSym sym, const int8_t* big, const T* L, const T* Dinv, const T* X, const double* b,
                      double* y, double* W, int nrhs, RhsStride rs, hipStream_t st, int tiny_lanes) {
    const int32_t nrest = count - ntiny;
    if (nrest == 0 && nx == 0) {
        if (ntiny > 0 && tiny_lanes == 32)
            help::launch_kernel(k_fwd_tiny<T, 32>, dim3((ntiny + 7) / 8, nrhs), 256, 0, st, list, ntiny, sym, L,
                                Dinv, b, y, W, rs);
        else if (ntiny > 0)
            help::launch_kernel(k_fwd_tiny<T, 64>, dim3((ntiny + 3) / 4, nrhs), 256, 0, st, list, ntiny, sym, L,
                                Dinv, b, y, W, rs);
        return;
    }
    const int32_t grid = (ntiny + 3) / 4 + nrest + nx;
    help::launch_kernel(k_fwd_level<T>, dim3(grid, nrhs), 256, 0, st, list, ntiny, nrest, xt, sym, big, L, Dinv, X,
                        b, y, W, rs);
}
template <typename T>
void launch_bwd_level(const int32_t* list, int32_t count, int32_t ntiny, const SolveTask* xt, int32_t nx,
                      DevSym sym, const int8_t* big, const T* L, const T* Dinv, const T* X, const double* y,
                      const double* t, double* x, int nrhs, RhsStride rs, hipStream_t st, int tiny_lanes) {
    const int32_t nrest = count - ntiny;
    if (nrest == 0 && nx == 0) {
        if (ntiny > 0 && tiny_lanes == 32)
            help::launch_kernel(k_bwd_tiny<T, 32>, dim3((ntiny + 7) / 8, nrhs), 256, 0, st, list, ntiny, sym, L,
                                Dinv, y, x, rs);
        else if (ntiny > 0)
            help::launch_kernel(k_bwd_tiny<T, 64>, dim3((ntiny + 3) / 4, nrhs), 256, 0, st, list, ntiny, sym, L,
                                Dinv, y, x, rs);
        return;
    }
    const int32_t grid = (ntiny + 3) / 4 + nrest + nx;
    help::launch_kernel(k_bwd_level<T>, dim3(grid, nrhs), 256, 0, st, list, ntiny, nrest, xt, sym, big, L, Dinv, X,
                        y, t, x, rs);
}
template <typename T>
void launch_fwd_gemv(const SolveTask* tasks, int32_t count, DevSym sym, const T* L, const double* y,
                     double* W, int nrhs, RhsStride rs, hipStream_t st) {
    if (count > 0) help::launch_kernel(k_fwd_gemv<T>, dim3(count, nrhs), 1024, 0, st, tasks, sym, L, y, W, rs);
}
template <typename T>
void launch_bwd_gemv(const SolveTask* tasks, int32_t count, DevSym sym, const T* L, const double* x,
                     double* t, int nrhs, RhsStride rs, hipStream_t st) {
    if (count > 0) help::launch_kernel(k_bwd_gemv<T>, dim3(count, nrhs), 256, 0, st, tasks, sym, L, x, t, rs);
}

#define HELP_INST(T)                                                                                   \
    template void launch_fwd_level<T>(const int32_t*, int32_t, int32_t, const SolveTask*, int32_t,     \
                                      DevSym, const int8_t*, const T*, const T*, const T*,              \
                                      const double*, double*, double*, int, RhsStride, hipStream_t,     \
                                      int);                                                             \
    template void launch_bwd_level<T>(const int32_t*, int32_t, int32_t, const SolveTask*, int32_t,     \
                                      DevSym, const int8_t*, const T*, const T*, const T*,              \
                                      const double*, const double*, double*, int, RhsStride,            \
                                      hipStream_t, int);                                                \
    template void launch_fwd_gemv<T>(const SolveTask*, int32_t, DevSym, const T*, const double*,       \
                                     double*, int, RhsStride, hipStream_t);                             \
    template void launch_bwd_gemv<T>(const SolveTask*, int32_t, DevSym, const T*, const double*,       \
                                     double*, int, RhsStride, hipStream_t);
HELP_INST(double)
HELP_INST(float)
#undef HELP_INST

}  // namespace help
