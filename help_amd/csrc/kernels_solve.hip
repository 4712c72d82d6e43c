// Supernodal triangular solves L y = b and L^T x = y, level by level.
//
// Forward (leaves to root): each supernode gathers its children's update
// vectors (the multifrontal analogue of extend-add, children in fixed
// order), solves with its diagonal block L11 and forms its own update vector
// W_s = L21 y_s + (children's rows beyond s).  Backward (root to leaves):
// each supernode subtracts L21^T x(rows) and solves with L11^T; the rows are
// ancestors, already final.  Vectors are FP64 whatever the factor type.
// Big supernodes hand the L21 products to separate many-workgroup kernels.
#include "device.h"

namespace help {

namespace {

constexpr int kMaxLdsCols = kMaxSnCols;   // x segment of a supernode held in LDS

template <typename T>
__global__ __launch_bounds__(256) void k_fwd(const int32_t* __restrict__ list, DevSym sym,
                                             const int8_t* __restrict__ big,
                                             const T* __restrict__ L, double* __restrict__ x,
                                             double* __restrict__ W) {
    __shared__ double xs[kMaxLdsCols];
    const int32_t s = list[blockIdx.x];
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const int32_t cu = c - n;
    const T* Lp = L + sym.loff[s];
    double* Ws = W + sym.reloff[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    for (int32_t t = tid; t < cu; t += nt) Ws[t] = 0.0;
    for (int32_t t = tid; t < n; t += nt) xs[t] = x[f + t];
    __syncthreads();
    for (int32_t t = sym.childptr[s]; t < sym.childptr[s + 1]; ++t) {
        const int32_t ch = sym.child[t];
        const int32_t chc = sym.snrow[ch] - (sym.sfirst[ch + 1] - sym.sfirst[ch]);
        const int32_t* rel = sym.relmap + sym.reloff[ch];
        const double* Wc = W + sym.reloff[ch];
        for (int32_t u = tid; u < chc; u += nt) {
            int32_t r = rel[u];
            double v = Wc[u];
            if (r < n) xs[r] -= v;
            else Ws[r - n] += v;
        }
        __syncthreads();
    }
    const int32_t rend = big[s] ? n : c;
    for (int32_t jb = 0; jb < n; jb += 64) {
        const int32_t nb = min(64, n - jb);
        if (wave == 0) {
            double xi = lane < nb ? xs[jb + lane] : 0.0;
            for (int32_t j = 0; j < nb; ++j) {
                const T* col = Lp + (int64_t)(jb + j) * c + jb;
                double lij = (lane > j && lane < nb) ? (double)col[lane] : 0.0;
                double xj = __shfl(xi, j) / (double)col[j];
                if (lane == j) xi = xj;
                if (lane > j) xi -= lij * xj;
            }
            if (lane < nb) xs[jb + lane] = xi;
        }
        __syncthreads();
        for (int32_t i = jb + nb + tid; i < rend; i += nt) {
            double acc = 0.0;
            for (int32_t j = 0; j < nb; ++j) acc += (double)Lp[(int64_t)(jb + j) * c + i] * xs[jb + j];
            if (i < n) xs[i] -= acc;
            else Ws[i - n] += acc;
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
    __shared__ double red[4][64];
    double acc = 0.0;
    if (i < c)
        for (int32_t j = part; j < n; j += 4) acc += (double)Lp[(int64_t)j * c + i] * xs[j];
    red[part][r] = acc;
    __syncthreads();
    if (part == 0 && i < c) Ws[i - n] += (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd_gemv(const SolveTask* __restrict__ tasks, DevSym sym,
                                                  const T* __restrict__ L,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ tout) {
    const SolveTask tk = tasks[blockIdx.x];
    const int32_t s = tk.s;
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const T* Lp = L + sym.loff[s];
    const int32_t* rows = sym.rowidx + sym.rowoff[s];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    // each wave: 16 columns of the block; lanes stride the rows
    for (int32_t q = 0; q < 16; ++q) {
        const int32_t j = tk.blk * 64 + wave * 16 + q;
        if (j >= n) break;
        const T* col = Lp + (int64_t)j * c;
        double acc = 0.0;
        for (int32_t i = n + lane; i < c; i += 64) acc += (double)col[i] * x[rows[i]];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (lane == 0) tout[f + j] = acc;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd(const int32_t* __restrict__ list, DevSym sym,
                                             const int8_t* __restrict__ big,
                                             const T* __restrict__ L, double* __restrict__ x,
                                             const double* __restrict__ tin) {
    __shared__ double xs[kMaxLdsCols];
    const int32_t s = list[blockIdx.x];
    const int32_t f = sym.sfirst[s];
    const int32_t n = sym.sfirst[s + 1] - f;
    const int32_t c = sym.snrow[s];
    const T* Lp = L + sym.loff[s];
    const int32_t* rows = sym.rowidx + sym.rowoff[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
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
        // subtract the rows of L11 below this block (already solved)
        const int32_t below = jb + nb;
        if (below < n) {
            for (int32_t q = wave; q < nb; q += 4) {
                const T* col = Lp + (int64_t)(jb + q) * c;
                double acc = 0.0;
                for (int32_t i = below + lane; i < n; i += 64) acc += (double)col[i] * xs[i];
                for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
                if (lane == 0) xs[jb + q] -= acc;
            }
        }
        __syncthreads();
        if (wave == 0) {
            double xj = lane < nb ? xs[jb + lane] : 0.0;
            for (int32_t i = nb - 1; i >= 0; --i) {
                double lii = (double)Lp[(int64_t)(jb + i) * c + jb + i];
                double zi = __shfl(xj, i) / lii;
                if (lane == i) xj = zi;
                double lij = (lane < i) ? (double)Lp[(int64_t)(jb + lane) * c + jb + i] : 0.0;
                if (lane < i) xj -= lij * zi;
            }
            if (lane < nb) xs[jb + lane] = xj;
        }
        __syncthreads();
    }
    for (int32_t t = tid; t < n; t += nt) x[f + t] = xs[t];
}

}  // namespace

template <typename T>
void launch_fwd(const int32_t* list, int32_t count, DevSym sym, const int8_t* big, const T* L,
                double* x, double* W, hipStream_t st) {
    if (count > 0) k_fwd<T><<<count, 256, 0, st>>>(list, sym, big, L, x, W);
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
                double* x, const double* t, hipStream_t st) {
    if (count > 0) k_bwd<T><<<count, 256, 0, st>>>(list, sym, big, L, x, t);
}

#define HELP_INST(T)                                                                              \
    template void launch_fwd<T>(const int32_t*, int32_t, DevSym, const int8_t*, const T*, double*, \
                                double*, hipStream_t);                                            \
    template void launch_fwd_gemv<T>(const SolveTask*, int32_t, DevSym, const T*, const double*,  \
                                     double*, hipStream_t);                                       \
    template void launch_bwd_gemv<T>(const SolveTask*, int32_t, DevSym, const T*, const double*,  \
                                     double*, hipStream_t);                                       \
    template void launch_bwd<T>(const int32_t*, int32_t, DevSym, const int8_t*, const T*, double*, \
                                const double*, hipStream_t);
HELP_INST(double)
HELP_INST(float)
#undef HELP_INST

}  // namespace help
