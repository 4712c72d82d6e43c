// Supernodal multifrontal Cholesky on CDNA4.
//
// Storage per supernode s (ncol n, nrow c, structure rows sorted):
//   L panel  c x n, column-major (lda = c), rows = the structure of s;
//   U        (c-n) x (c-n), column-major lower, the update ("contribution")
//            matrix passed to the parent by extend-add;
//   Dinv     the inverse of every 64-column diagonal block of L11 (nb x nb,
//            column-major), used by the panel solve and the triangular solves.
// The symbolic schedule processes the elimination tree level by level;
// within a level every front is independent.  A parent pulls its children's
// update matrices one child after the other, so the summation order is fixed
// and no atomics touch the factor.
//
// Small fronts (c <= 256, c*n <= 7680): one workgroup per front does the
// extend-add, the dense partial factorization in LDS, the diagonal-block
// inverses and the update SYRK.
// Large fronts: extend-add by column chunks, then for each 64-column panel a
// one-wave diagonal-block Cholesky + inverse, the panel rows times the
// inverse on MFMA, an MFMA GEMM on the trailing panel columns, and finally
// one MFMA SYRK for U with the full inner dimension n.
#include "device.h"

namespace help {

namespace {

template <typename T> struct Huge;
template <> struct Huge<double> { static constexpr double v = 1e128; };
template <> struct Huge<float> { static constexpr float v = 1e30f; };

// pivot policy: a non-positive (or NaN) pivot is replaced by a huge value, so
// the column of L becomes ~0 (that row is dropped from the solve), and counted.
template <typename T>
__device__ inline T safe_sqrt_pivot(T d, int32_t* err, bool count) {
    if (!(d > T(0))) {
        if (count) atomicAdd(err, 1);
        d = Huge<T>::v;
    }
    return sqrt(d);
}

__device__ inline int64_t dinv_block_offset(const DevSym& sym, int32_t s, int32_t b) {
    return sym.dinvoff[s] + (int64_t)(b / kTile) * kTile * kTile;
}

// X = L^-1 of the factored lower-triangular nb x nb block at Lsrc (LDS,
// leading dimension lds) by row operations on the identity (256 threads),
// written to Db (column-major nb x nb).  X is LDS scratch of nb*(nb+1).
template <typename T>
__device__ void invert_lower_block(const T* Lsrc, int32_t lds, int32_t nb, T* __restrict__ Db, T* X) {
    const int tid = threadIdx.x;
    const int ldx = nb + 1;
    for (int idx = tid; idx < nb * nb; idx += 256) {
        int j = idx / nb, i = idx - j * nb;
        X[j * ldx + i] = (i == j) ? T(1) : T(0);
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        const T djj = Lsrc[j * lds + j];
        if (tid <= j) X[tid * ldx + j] /= djj;      // row j, columns k <= j
        __syncthreads();
        // rows i > j, columns k <= j:  X[i][k] -= L[i][j] X[j][k]
        const int cnt = (nb - j - 1) * (j + 1);
        for (int idx = tid; idx < cnt; idx += 256) {
            int k = idx / (nb - j - 1), i = j + 1 + idx - k * (nb - j - 1);
            X[k * ldx + i] -= Lsrc[j * lds + i] * X[k * ldx + j];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < nb * nb; idx += 256) {
        int j = idx / nb, i = idx - j * nb;
        Db[j * nb + i] = (i >= j) ? X[j * ldx + i] : T(0);
    }
}

// ------------------------------------------------------------------ small
// Dynamic LDS: the c x n panel followed by the (nb x nb+1) inverse scratch.
template <typename T>
__global__ __launch_bounds__(256) void k_small_fronts(const int32_t* __restrict__ list, DevSym sym,
                                                      T* __restrict__ L, T* __restrict__ U,
                                                      T* __restrict__ Dinv,
                                                      int32_t* __restrict__ err) {
    extern __shared__ unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);
    __shared__ T piv;
    const int32_t s = list[blockIdx.x];
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const int32_t cu = c - n;
    T* Lp = L + sym.loff[s];
    T* Up = U + sym.uoff[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int32_t cn = c * n;
    for (int32_t idx = tid; idx < cn; idx += nt) F[idx] = Lp[idx];
    __syncthreads();
    // extend-add the children, one after the other
    for (int32_t t = sym.childptr[s]; t < sym.childptr[s + 1]; ++t) {
        const int32_t ch = sym.child[t];
        const int32_t chn = sym.sfirst[ch + 1] - sym.sfirst[ch];
        const int32_t chc = sym.snrow[ch] - chn;
        const int32_t* rel = sym.relmap + sym.reloff[ch];
        const T* Uc = U + sym.uoff[ch];
        const int64_t tot = (int64_t)chc * chc;
        for (int64_t idx = tid; idx < tot; idx += nt) {
            int32_t jj = (int32_t)(idx / chc), ii = (int32_t)(idx - (int64_t)jj * chc);
            if (ii < jj) continue;
            int32_t r = rel[ii], cc = rel[jj];
            T v = Uc[idx];
            if (cc < n) F[cc * c + r] += v;
            else Up[(int64_t)(cc - n) * cu + (r - n)] += v;
        }
        __syncthreads();
    }
    // dense partial Cholesky of the panel, right-looking by columns
    for (int32_t j = 0; j < n; ++j) {
        if (tid == 0) {
            T d = safe_sqrt_pivot<T>(F[j * c + j], err, true);
            F[j * c + j] = d;
            piv = d;
        }
        __syncthreads();
        const T p = piv;
        for (int32_t i = j + 1 + tid; i < c; i += nt) F[j * c + i] /= p;
        __syncthreads();
        const int32_t rem = n - j - 1;
        const int32_t tot = rem * c;
        for (int32_t idx = tid; idx < tot; idx += nt) {
            int32_t k = j + 1 + idx / c, i = idx % c;
            if (i >= k) F[k * c + i] -= F[j * c + i] * F[j * c + k];
        }
        __syncthreads();
    }
    for (int32_t idx = tid; idx < cn; idx += nt) Lp[idx] = F[idx];
    // inverses of the diagonal blocks
    T* X = F + cn;
    for (int32_t b = 0; b < n; b += kTile) {
        const int32_t nb = min(kTile, n - b);
        invert_lower_block<T>(F + (int64_t)b * c + b, c, nb, Dinv + dinv_block_offset(sym, s, b), X);
    }
    // U -= L21 L21^T
    const int32_t tot = cu * cu;
    for (int32_t idx = tid; idx < tot; idx += nt) {
        int32_t jj = idx / cu, ii = idx - jj * cu;
        if (ii < jj) continue;
        T acc = 0;
        for (int32_t j = 0; j < n; ++j) acc += F[j * c + n + ii] * F[j * c + n + jj];
        Up[idx] -= acc;
    }
}

// ------------------------------------------------------------ extend-add
// Task = 32 columns [c0, c1) of a front.  For each child in order, the
// child columns that land in the range are spread over the 4 waves (a
// column per wave), the lanes take the rows, 4 rows per lane in flight.
template <typename T>
__global__ __launch_bounds__(256) void k_extend_add(const EaTask* __restrict__ tasks, DevSym sym,
                                                    T* __restrict__ L, T* __restrict__ U) {
    const EaTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s;
    const int32_t n = sym.sfirst[s + 1] - sym.sfirst[s];
    const int32_t c = sym.snrow[s];
    const int32_t cu = c - n;
    T* Lp = L + sym.loff[s];
    T* Up = U + sym.uoff[s];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int32_t t = sym.childptr[s]; t < sym.childptr[s + 1]; ++t) {
        const int32_t ch = sym.child[t];
        const int32_t chn = sym.sfirst[ch + 1] - sym.sfirst[ch];
        const int32_t chc = sym.snrow[ch] - chn;
        const int32_t* rel = sym.relmap + sym.reloff[ch];
        const T* Uc = U + sym.uoff[ch];
        // rel is ascending: the child columns that land in [c0, c1)
        int32_t lo = 0, hi = chc;
        while (lo < hi) { int32_t mid = (lo + hi) >> 1; if (rel[mid] < tk.c0) lo = mid + 1; else hi = mid; }
        for (int32_t jj = lo + wave; jj < chc; jj += 4) {
            const int32_t cc = rel[jj];
            if (cc >= tk.c1) break;
            const T* col = Uc + (int64_t)jj * chc;
            T* dst = (cc < n) ? Lp + (int64_t)cc * c : Up + (int64_t)(cc - n) * cu - n;
            for (int32_t i0 = jj + lane; i0 < chc; i0 += 256) {
                int32_t rr[4];
                T v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int32_t ii = i0 + 64 * u;
                    rr[u] = (ii < chc) ? rel[ii] : -1;
                    v[u] = (ii < chc) ? col[ii] : T(0);
                }
                T d[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) d[u] = (rr[u] >= 0) ? dst[rr[u]] : T(0);
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (rr[u] >= 0) dst[rr[u]] = d[u] + v[u];
            }
        }
        __syncthreads();
    }
}

// ------------------------------------ diagonal block Cholesky + inverse
// Elimination with unit multipliers on the augmented block [A | I]: step j
// stores m_ij = a_ij / d_j in place, updates the trailing lower part
// (a_ik -= d_j m_ij m_kj) and applies the same row operation to the identity
// (x_ik -= m_ij x_jk).  At the end A holds L_u (unit) and D, X holds L_u^-1,
// so L = L_u D^1/2 and L^-1 = D^-1/2 L_u^-1.  The block is padded with the
// identity to 64, which leaves the leading nb x nb results unchanged.
template <typename T>
__device__ void factor_invert_block(const T* __restrict__ src, int64_t lds, int32_t nb,
                                    T* __restrict__ Ldst, T* __restrict__ Db, T* A, T* X,
                                    int32_t* err) {
    constexpr int ld = kTile + 1;
    const int tid = threadIdx.x;
    const int row = tid & 63, part = tid >> 6;
    for (int idx = tid; idx < kTile * kTile; idx += 256) {
        int j = idx >> 6, i = idx & 63;
        T v = (i == j) ? T(1) : T(0);
        if (i < nb && j < nb && i >= j) v = src[(int64_t)j * lds + i];
        A[j * ld + i] = v;
        X[j * ld + i] = (i == j) ? T(1) : T(0);
    }
    __syncthreads();
    for (int j = 0; j < kTile; ++j) {
        T piv = A[j * ld + j];
        if (!(piv > T(0))) {
            if (tid == 0 && j < nb) atomicAdd(err, 1);
            piv = Huge<T>::v;
        }
        __syncthreads();
        if (part == 0) {
            if (row > j) A[j * ld + row] /= piv;
            else if (row == j) A[j * ld + j] = piv;
        }
        __syncthreads();
        if (row > j) {
            const T mi = A[j * ld + row];
            for (int k = part; k <= row; k += 4) {
                if (k > j) A[k * ld + row] -= piv * mi * A[j * ld + k];
                else X[k * ld + row] -= mi * X[k * ld + j];
            }
        }
        __syncthreads();
    }
    for (int idx = tid; idx < kTile * kTile; idx += 256) {
        int j = idx >> 6, i = idx & 63;
        if (i >= nb || j >= nb) continue;
        const T sdj = sqrt(A[j * ld + j]);
        if (i >= j) Ldst[(int64_t)j * lds + i] = (i == j) ? sdj : A[j * ld + i] * sdj;
        Db[j * nb + i] = (i >= j) ? X[j * ld + i] / sqrt(A[i * ld + i]) : T(0);
    }
}

// Register-resident form of factor_invert_block with 512 threads: thread
// (r, p) holds A[r][k] and X[r][k] for the 8 columns k = 8p + q.  Step j
// broadcasts column j of A and row j of X through double-buffered LDS (one
// barrier per step); every thread then applies the same predicated update
// to its 8 columns, so the step is a short branch-free sequence.  Entries
// above the diagonal of A are updated too but never read.
template <typename T>
__global__ __launch_bounds__(512) void k_potrf_diag(const PanelTask* __restrict__ tasks, DevSym sym,
                                                    T* __restrict__ L, T* __restrict__ Dinv,
                                                    int32_t* __restrict__ err) {
    constexpr int Q = 8;
    __shared__ T colb[2][kTile];
    __shared__ T xrow[2][kTile];
    __shared__ T dvec[kTile];
    const PanelTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s, b = tk.b;
    const int32_t n = sym.sfirst[s + 1] - sym.sfirst[s];
    const int32_t c = sym.snrow[s];
    const int32_t nb = min(kTile, n - b);
    T* Lp = L + sym.loff[s] + (int64_t)b * c + b;
    T* Db = Dinv + dinv_block_offset(sym, s, b);
    const int tid = threadIdx.x;
    const int r = tid & 63, p = tid >> 6;
    const int k0 = Q * p;
    T a[Q], x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int k = k0 + q;
        T v = (r == k) ? T(1) : T(0);
        if (r < nb && k < nb && r >= k) v = Lp[(int64_t)k * c + r];
        a[q] = v;
        x[q] = (r == k) ? T(1) : T(0);
    }
    int nbad = 0;
    for (int j = 0; j < kTile; ++j) {
        const int buf = j & 1;
        const int jq = j - k0;               // register of column j in this wave (0..Q-1 if owned)
        if (jq >= 0 && jq < Q) {
            T v = a[0];
#pragma unroll
            for (int q = 1; q < Q; ++q) v = (q == jq) ? a[q] : v;
            colb[buf][r] = v;
        }
        if (r == j) {
#pragma unroll
            for (int q = 0; q < Q; ++q) xrow[buf][k0 + q] = x[q];
        }
        __syncthreads();
        T piv = colb[buf][j];
        if (!(piv > T(0))) { if (j < nb) ++nbad; piv = Huge<T>::v; }
        const T rpiv = T(1) / piv;
        const T cr = colb[buf][r];
        const T m = (r > j) ? cr * rpiv : T(0);
        T ca[Q], cx[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) { ca[q] = colb[buf][k0 + q]; cx[q] = xrow[buf][k0 + q]; }
        // columns right of j update A, the others X (both operands finite)
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const bool right = k0 + q > j;
            const T ma = right ? m : T(0), mx = right ? T(0) : m;
            a[q] -= ma * ca[q];
            x[q] -= mx * cx[q];
        }
        if (jq >= 0 && jq < Q) {
            const T nv = (r > j) ? m : piv;
#pragma unroll
            for (int q = 0; q < Q; ++q)
                if (q == jq && r >= j) a[q] = nv;
        }
    }
    if (tid == 0 && nbad) atomicAdd(err, nbad);
    // diagonal d_k lives in thread (k, k/Q), register k%Q
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (r == k0 + q) dvec[r] = a[q];
    __syncthreads();
    const T sdr = sqrt(dvec[r]);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int k = k0 + q;
        if (r >= nb || k >= nb) continue;
        if (r >= k) {
            Lp[(int64_t)k * c + r] = (r == k) ? sdr : a[q] * sqrt(dvec[k]);
            Db[k * nb + r] = x[q] / sdr;
        } else {
            Db[k * nb + r] = T(0);
        }
    }
}

// ------------------------------------------------------------- MFMA GEMM
// 4 waves, each a 32x32 quarter of a 64x64 tile built from 2x2 MFMA 16x16x4
// blocks.  Lane layout (measured on gfx950 by help_selftest_mfma): A[i][k] in
// lane i + 16k, B[k][j] in lane j + 16k; the result D[i][j] sits in lane
// j + 16(i%4), element i/4 for V_MFMA_F64_16X16X4_F64 and in lane
// j + 16(i/4), element i%4 for V_MFMA_F32_16X16X4_F32.
typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef double4_t acc_t;
    static __device__ inline acc_t fma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <> struct Mfma<float> {
    typedef float4_t acc_t;
    static __device__ inline acc_t fma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// acc += As[k][rows] * Bs[k][cols]^T over k in [0, kc), As/Bs k-major (ld 64)
template <typename T>
__device__ inline void mfma_tile(const T* As, const T* Bs, int kc, int lane, int wr, int wc,
                                 typename Mfma<T>::acc_t (&acc)[2][2]) {
    for (int kk = 0; kk < kc; kk += 4) {
        const int kr = kk + (lane >> 4);
        T a0 = As[kr * kTile + wr + (lane & 15)];
        T a1 = As[kr * kTile + wr + 16 + (lane & 15)];
        T b0 = Bs[kr * kTile + wc + (lane & 15)];
        T b1 = Bs[kr * kTile + wc + 16 + (lane & 15)];
        acc[0][0] = Mfma<T>::fma(a0, b0, acc[0][0]);
        acc[0][1] = Mfma<T>::fma(a0, b1, acc[0][1]);
        acc[1][0] = Mfma<T>::fma(a1, b0, acc[1][0]);
        acc[1][1] = Mfma<T>::fma(a1, b1, acc[1][1]);
    }
}

constexpr int kKc = 16;   // inner-dimension chunk staged in LDS

// C[row0:+64, col0:+64] -= P[rows, k0:k1] * P[cols, k0:k1]^T, lower part only,
// with P the L panel of the front.  The k chunks are double-buffered in LDS
// and the next chunk's global loads are in flight while the MFMAs consume
// the current one (one barrier per chunk).
template <typename T>
__global__ __launch_bounds__(256) void k_gemm(const GemmTask* __restrict__ tasks, DevSym sym,
                                              T* __restrict__ L, T* __restrict__ U) {
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int kPer = kKc / 4;
    __shared__ T As[2][kKc * kTile];
    __shared__ T Bs[2][kKc * kTile];
    const GemmTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s;
    const int32_t n = sym.sfirst[s + 1] - sym.sfirst[s];
    const int32_t c = sym.snrow[s];
    const int32_t cu = c - n;
    const T* Lp = L + sym.loff[s];
    T* Lw = L + sym.loff[s];
    T* Up = U + sym.uoff[s];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    acc_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    const int li = tid & 63, lk = tid >> 6;
    const int32_t ia = tk.row0 + li, ib = tk.col0 + li;
    const bool ain = ia < c, bin = ib < c;
    const T* pa = Lp + min(ia, c - 1);
    const T* pb = Lp + min(ib, c - 1);
    T ra[kPer], rb[kPer];
    auto load = [&](int32_t kb) {
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const int32_t kg = min(kb + lk + 4 * r, tk.k1 - 1);
            ra[r] = pa[(int64_t)kg * c];
            rb[r] = pb[(int64_t)kg * c];
        }
    };
    load(tk.k0);
    int buf = 0;
    for (int32_t kb = tk.k0; kb < tk.k1; kb += kKc, buf ^= 1) {
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const int k = lk + 4 * r;
            const bool kin = kb + k < tk.k1;
            As[buf][k * kTile + li] = (kin && ain) ? ra[r] : T(0);
            Bs[buf][k * kTile + li] = (kin && bin) ? rb[r] : T(0);
        }
        __syncthreads();
        if (kb + kKc < tk.k1) load(kb + kKc);
        mfma_tile<T>(As[buf], Bs[buf], kKc, lane, wr, wc, acc);
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int r = 0; r < 4; ++r) {
                int32_t i = tk.row0 + wr + a * 16 + Mfma<T>::row(lane, r);
                int32_t j = tk.col0 + wc + b * 16 + (lane & 15);
                if (i >= c || j >= tk.col_end || i < j) continue;
                T v = acc[a][b][r];
                if (j < n) Lw[(int64_t)j * c + i] -= v;
                else Up[(int64_t)(j - n) * cu + (i - n)] -= v;
            }
}

// X := X L_bb^-T for the 64 rows of block rb-1 below the diagonal block,
// computed as X * (L_bb^-1)^T on MFMA in place.
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_gemm(const PanelTask* __restrict__ tasks, DevSym sym,
                                                   T* __restrict__ L, const T* __restrict__ Dinv) {
    typedef typename Mfma<T>::acc_t acc_t;
    __shared__ T As[kTile * kTile];
    __shared__ T Bs[kTile * kTile];
    const PanelTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s, b = tk.b;
    const int32_t n = sym.sfirst[s + 1] - sym.sfirst[s];
    const int32_t c = sym.snrow[s];
    const int32_t nb = min(kTile, n - b);
    const int32_t r0 = b + nb + (tk.rb - 1) * kTile;
    T* Lp = L + sym.loff[s];
    const T* Db = Dinv + dinv_block_offset(sym, s, b);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    // As[k][i] = X[i][k] = P[r0+i][b+k];  Bs[k][j] = Dinv[j][k] (Dinv column-major:
    // Db[k*nb + j] = Dinv[j][k]).  Every load is issued before the first LDS
    // store: out-of-range elements read a clamped in-range address and are
    // zeroed afterwards, so no branch separates the loads.
    constexpr int kPer = kTile * kTile / 256;
    T va[kPer], vb[kPer];
    {
        const int i = tid & 63;
        const int32_t ri = min(r0 + i, c - 1);
        const int ib = min(i, nb - 1);
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            const int k = (tid >> 6) + 4 * t;
            const int kc = min(k, nb - 1);
            va[t] = Lp[(int64_t)(b + kc) * c + ri];
            vb[t] = Db[kc * nb + ib];
        }
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            const int k = (tid >> 6) + 4 * t;
            As[k * kTile + i] = (k < nb && r0 + i < c) ? va[t] : T(0);
            Bs[k * kTile + i] = (k < nb && i < nb) ? vb[t] : T(0);
        }
    }
    __syncthreads();
    acc_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int q = 0; q < 2; ++q) acc[a][q] = acc_t{0, 0, 0, 0};
    mfma_tile<T>(As, Bs, kTile, lane, wr, wc, acc);
    for (int a = 0; a < 2; ++a)
        for (int q = 0; q < 2; ++q)
            for (int r = 0; r < 4; ++r) {
                int32_t i = wr + a * 16 + Mfma<T>::row(lane, r);
                int32_t j = wc + q * 16 + (lane & 15);
                if (r0 + i >= c || j >= nb) continue;
                Lp[(int64_t)(b + j) * c + r0 + i] = acc[a][q][r];
            }
}

// one wave: lane l feeds a[l], b[l]; writes its 4 result elements
template <typename T>
__global__ void k_mfma_probe(const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out) {
    typedef typename Mfma<T>::acc_t acc_t;
    const int l = threadIdx.x;
    acc_t acc = acc_t{0, 0, 0, 0};
    acc = Mfma<T>::fma(a[l], b[l], acc);
    for (int r = 0; r < 4; ++r) out[4 * l + r] = acc[r];
}

}  // namespace

template <typename T>
void launch_mfma_probe(const T* a, const T* b, T* out, hipStream_t st) {
    k_mfma_probe<T><<<1, 64, 0, st>>>(a, b, out);
}
template void launch_mfma_probe<double>(const double*, const double*, double*, hipStream_t);
template void launch_mfma_probe<float>(const float*, const float*, float*, hipStream_t);

template <typename T>
void launch_small_fronts(const int32_t* list, int32_t count, DevSym sym, T* L, T* U, T* Dinv,
                         int32_t* err, int32_t lds_elems, hipStream_t st) {
    if (count > 0)
        k_small_fronts<T><<<count, 256, (size_t)lds_elems * sizeof(T), st>>>(list, sym, L, U, Dinv, err);
}
template <typename T>
void launch_extend_add(const EaTask* tasks, int32_t count, DevSym sym, T* L, T* U, hipStream_t st) {
    if (count > 0) k_extend_add<T><<<count, 256, 0, st>>>(tasks, sym, L, U);
}
template <typename T>
void launch_gemm(const GemmTask* tasks, int32_t count, DevSym sym, T* L, T* U, hipStream_t st) {
    if (count > 0) k_gemm<T><<<count, 256, 0, st>>>(tasks, sym, L, U);
}
template <typename T>
void launch_potrf_diag(const PanelTask* tasks, int32_t count, DevSym sym, T* L, T* Dinv, int32_t* err,
                       hipStream_t st) {
    if (count > 0) k_potrf_diag<T><<<count, 512, 0, st>>>(tasks, sym, L, Dinv, err);
}
template <typename T>
void launch_trsm_rows(const PanelTask* tasks, int32_t count, DevSym sym, T* L, const T* Dinv,
                      hipStream_t st) {
    if (count > 0) k_trsm_gemm<T><<<count, 256, 0, st>>>(tasks, sym, L, Dinv);
}

#define HELP_INST(T)                                                                                \
    template void launch_small_fronts<T>(const int32_t*, int32_t, DevSym, T*, T*, T*, int32_t*,      \
                                         int32_t, hipStream_t);                                     \
    template void launch_extend_add<T>(const EaTask*, int32_t, DevSym, T*, T*, hipStream_t);       \
    template void launch_gemm<T>(const GemmTask*, int32_t, DevSym, T*, T*, hipStream_t);           \
    template void launch_potrf_diag<T>(const PanelTask*, int32_t, DevSym, T*, T*, int32_t*,        \
                                       hipStream_t);                                                \
    template void launch_trsm_rows<T>(const PanelTask*, int32_t, DevSym, T*, const T*, hipStream_t);
HELP_INST(double)
HELP_INST(float)
#undef HELP_INST

}  // namespace help
