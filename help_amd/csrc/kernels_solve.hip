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
#include "device.h"

namespace help {

namespace {

constexpr int kMaxLdsCols = kMaxSnCols;   // x segment of a supernode held in LDS
constexpr int kMaxLdsRows = 4096;         // gathered x(rows) held in LDS

__device__ inline const void* dinv_block(const DevSym& sym, const void* base, int32_t s, int32_t jb,
                                         int esz) {
    return (const char*)base + (sym.dinvoff[s] + (int64_t)(jb / 64) * 64 * 64) * esz;
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwd(const int32_t* __restrict__ list, DevSym sym,
                                             const int8_t* __restrict__ big,
                                             const T* __restrict__ L, const T* __restrict__ Dinv,
                                             double* __restrict__ x, double* __restrict__ W) {
    __shared__ double xs[kMaxLdsCols];
    __shared__ double red[4][64];
    const int32_t s = list[blockIdx.x];
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const int32_t cu = c - n;
    const T* Lp = L + sym.loff[s];
    double* Ws = W + sym.reloff[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int r = tid & 63, part = tid >> 6;
    for (int32_t t = tid; t < cu; t += nt) Ws[t] = 0.0;
    for (int32_t t = tid; t < n; t += nt) xs[t] = x[f + t];
    __syncthreads();
    for (int32_t t = sym.childptr[s]; t < sym.childptr[s + 1]; ++t) {
        const int32_t ch = sym.child[t];
        const int32_t chc = sym.snrow[ch] - (sym.sfirst[ch + 1] - sym.sfirst[ch]);
        const int32_t* rel = sym.relmap + sym.reloff[ch];
        const double* Wc = W + sym.reloff[ch];
        for (int32_t u = tid; u < chc; u += nt) {
            int32_t rr = rel[u];
            double v = Wc[u];
            if (rr < n) xs[rr] -= v;
            else Ws[rr - n] += v;
        }
        __syncthreads();
    }
    const int32_t rend = big[s] ? n : c;
    for (int32_t jb = 0; jb < n; jb += 64) {
        const int32_t nb = min(64, n - jb);
        const T* Db = (const T*)dinv_block(sym, Dinv, s, jb, sizeof(T));
        double acc = 0.0;
        if (r < nb)
            for (int32_t k = part; k < nb; k += 4) acc += (double)Db[k * nb + r] * xs[jb + k];
        red[part][r] = acc;
        __syncthreads();
        if (part == 0 && r < nb) xs[jb + r] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
        __syncthreads();
        for (int32_t i = jb + nb + tid; i < rend; i += nt) {
            double a = 0.0;
            for (int32_t j = 0; j < nb; ++j) a += (double)Lp[(int64_t)(jb + j) * c + i] * xs[jb + j];
            if (i < n) xs[i] -= a;
            else Ws[i - n] += a;
        }
        __syncthreads();
    }
    for (int32_t t = tid; t < n; t += nt) x[f + t] = xs[t];
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwd_gemv(const SolveTask* __restrict__ tasks, DevSym sym,
                                                  const T* __restrict__ L,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ W) {
    __shared__ double xs[kMaxLdsCols];
    __shared__ double red[4][64];
    const SolveTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s;
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const T* Lp = L + sym.loff[s];
    double* Ws = W + sym.reloff[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int32_t t = tid; t < n; t += nt) xs[t] = x[f + t];
    __syncthreads();
    // 64 rows, 4 threads per row splitting the columns, fixed combine order
    const int32_t r = tid & 63, part = tid >> 6;
    const int32_t i = n + tk.blk * 64 + r;
    double acc = 0.0;
    if (i < c)
        for (int32_t j = part; j < n; j += 4) acc += (double)Lp[(int64_t)j * c + i] * xs[j];
    red[part][r] = acc;
    __syncthreads();
    if (part == 0 && i < c) Ws[i - n] += (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
}

// t[f+j] = sum_{i>=n} L21(i,j) x[row_i] for 16 columns j of a big supernode
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_gemv(const SolveTask* __restrict__ tasks, DevSym sym,
                                                  const T* __restrict__ L,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ tout) {
    __shared__ double xr[kMaxLdsRows];
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
    const bool lds = cu <= kMaxLdsRows;
    if (lds)
        for (int32_t i = tid; i < cu; i += 256) xr[i] = x[rows[n + i]];
    __syncthreads();
    for (int32_t q = 0; q < 4; ++q) {
        const int32_t j = tk.blk * 16 + wave * 4 + q;
        if (j >= n) break;
        const T* col = Lp + (int64_t)j * c + n;
        double acc = 0.0;
        if (lds)
            for (int32_t i = lane; i < cu; i += 64) acc += (double)col[i] * xr[i];
        else
            for (int32_t i = lane; i < cu; i += 64) acc += (double)col[i] * x[rows[n + i]];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (lane == 0) tout[f + j] = acc;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd(const int32_t* __restrict__ list, DevSym sym,
                                             const int8_t* __restrict__ big,
                                             const T* __restrict__ L, const T* __restrict__ Dinv,
                                             double* __restrict__ x, const double* __restrict__ tin) {
    __shared__ double xs[kMaxLdsCols];
    __shared__ double red[4][64];
    const int32_t s = list[blockIdx.x];
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const T* Lp = L + sym.loff[s];
    const int32_t* rows = sym.rowidx + sym.rowoff[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int r = tid & 63, part = tid >> 6;
    if (big[s]) {
        for (int32_t t = tid; t < n; t += nt) xs[t] = x[f + t] - tin[f + t];
    } else {
        for (int32_t j = wave; j < n; j += 4) {
            const T* col = Lp + (int64_t)j * c;
            double acc = 0.0;
            for (int32_t i = n + lane; i < c; i += 64) acc += (double)col[i] * x[rows[i]];
            for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
            if (lane == 0) xs[j] = x[f + j] - acc;
        }
    }
    __syncthreads();
    const int32_t nblk = (n + 63) / 64;
    for (int32_t bi = nblk - 1; bi >= 0; --bi) {
        const int32_t jb = bi * 64;
        const int32_t nb = min(64, n - jb);
        const int32_t below = jb + nb;
        if (below < n) {
            for (int32_t q = wave; q < nb; q += 4) {
                const T* col = Lp + (int64_t)(jb + q) * c;
                double acc = 0.0;
                for (int32_t i = below + lane; i < n; i += 64) acc += (double)col[i] * xs[i];
                for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
                if (lane == 0) xs[jb + q] -= acc;
            }
            __syncthreads();
        }
        const T* Db = (const T*)dinv_block(sym, Dinv, s, jb, sizeof(T));
        double acc = 0.0;
        if (r < nb)
            for (int32_t k = part; k < nb; k += 4) acc += (double)Db[r * nb + k] * xs[jb + k];
        red[part][r] = acc;
        __syncthreads();
        if (part == 0 && r < nb) xs[jb + r] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
        __syncthreads();
    }
    for (int32_t t = tid; t < n; t += nt) x[f + t] = xs[t];
}

}  // namespace

template <typename T>
void launch_fwd(const int32_t* list, int32_t count, DevSym sym, const int8_t* big, const T* L,
                const T* Dinv, double* x, double* W, hipStream_t st) {
    if (count > 0) k_fwd<T><<<count, 256, 0, st>>>(list, sym, big, L, Dinv, x, W);
}
template <typename T>
void launch_fwd_gemv(const SolveTask* tasks, int32_t count, DevSym sym, const T* L, const double* x,
                     double* W, hipStream_t st) {
    if (count > 0) k_fwd_gemv<T><<<count, 256, 0, st>>>(tasks, sym, L, x, W);
}
template <typename T>
void launch_bwd_gemv(const SolveTask* tasks, int32_t count, DevSym sym, const T* L, const double* x,
                     double* t, hipStream_t st) {
    if (count > 0) k_bwd_gemv<T><<<count, 256, 0, st>>>(tasks, sym, L, x, t);
}
template <typename T>
void launch_bwd(const int32_t* list, int32_t count, DevSym sym, const int8_t* big, const T* L,
                const T* Dinv, double* x, const double* t, hipStream_t st) {
    if (count > 0) k_bwd<T><<<count, 256, 0, st>>>(list, sym, big, L, Dinv, x, t);
}

#define HELP_INST(T)                                                                                  \
    template void launch_fwd<T>(const int32_t*, int32_t, DevSym, const int8_t*, const T*, const T*,   \
                                double*, double*, hipStream_t);                                       \
    template void launch_fwd_gemv<T>(const SolveTask*, int32_t, DevSym, const T*, const double*,      \
                                     double*, hipStream_t);                                           \
    template void launch_bwd_gemv<T>(const SolveTask*, int32_t, DevSym, const T*, const double*,      \
                                     double*, hipStream_t);                                           \
    template void launch_bwd<T>(const int32_t*, int32_t, DevSym, const int8_t*, const T*, const T*,   \
                                double*, const double*, hipStream_t);
HELP_INST(double)
HELP_INST(float)
#undef HELP_INST

}  // namespace help
