// HBM-bound sparse kernels: ADA^T numeric assembly, SpMV gathers, the
// right-hand side and back-substitution of the KKT solve.
//
// Summation order is fixed and sequential per output, and multiply/add are
// never contracted into an FMA, so every kernel here is bit-identical to the
// CPU restatement of the same HELP loop.
#pragma clang fp contract(off)

#include <algorithm>

#include "device.h"

namespace help {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// acc <- acc (+/-) term(t) for t = b..e-1, in that order.  The loads of U
// consecutive terms are issued together: indices past the end are clamped
// to e-1 (an in-range duplicate) and their terms dropped by a select, so no
// branch separates the loads and the rounding is that of the plain loop.
template <int U, bool SUB, typename F>
__device__ inline double seq_acc(int64_t b, int64_t e, double acc, F term) {
    for (int64_t t = b; t < e; t += U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = term(t + u < e ? t + u : e - 1);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double nx = SUB ? acc - v[u] : acc + v[u];
            acc = (t + u < e) ? nx : acc;
        }
    }
    return acc;
}

__global__ void k_dinv(const double* __restrict__ d, double* __restrict__ dinv, int32_t n) {
    int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) dinv[j] = 1.0 / d[j];   // UsePardiso.cpp:172-173
}

// w_p = A_ij * dinv_j for every nonzero p = (i,j) of A (CSC order: the
// column index is nondecreasing in p, so the dinv reads stream): the factor
// (A_ij * dinv_j) of Helper.cpp:360, computed once per nonzero instead of
// once per product it enters.
__global__ void k_scale_nz(int64_t nnz, const int32_t* __restrict__ colof, const double* __restrict__ val,
                           const double* __restrict__ dinv, double* __restrict__ w) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) w[p] = val[p] * dinv[colof[p]];
}

// One slot (i,k) of the upper ADA^T per thread: sum over the columns j it
// shares with, in ascending j, of (A_ij * dinv_j) * A_kj starting from zero
// (Helper.cpp:358-373), with w_p = A_ij * dinv_j from k_scale_nz.  The
// result goes to dest[s] (a position in the supernodal factor storage) or
// to slot s when dest is null.
__global__ void k_adat(const int64_t* __restrict__ pairptr, const int32_t* __restrict__ pp,
                       const int32_t* __restrict__ pq, const double* __restrict__ w,
                       const double* __restrict__ val, int64_t nslots, const int64_t* __restrict__ dest,
                       double* __restrict__ outd, float* __restrict__ outf) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const double acc = seq_acc<4, false>(pairptr[s], pairptr[s + 1], 0.0,
                                         [&](int64_t t) { return w[pp[t]] * val[pq[t]]; });
    int64_t o = dest ? dest[s] : s;
    if (outd) outd[o] = acc;
    else outf[o] = (float)acc;
}

__global__ void k_add_beta(const int64_t* __restrict__ diagpos, int32_t m, double beta,
                           double* __restrict__ Ld, float* __restrict__ Lf) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int64_t o = diagpos[i];
    if (Ld) Ld[o] = Ld[o] + beta;
    else Lf[o] = (float)((double)Lf[o] + beta);
}

// FP32 factor input: the slot value of the symmetrically scaled matrix
// S (A D^-1 A^T + beta I) S, S = diag(1/sqrt(diag)), so every entry is at
// most 1 in magnitude and the FP32 range cannot overflow
__global__ void k_adat_scaled_f32(const int64_t* __restrict__ pairptr, const int32_t* __restrict__ pp,
                                  const int32_t* __restrict__ pq, const double* __restrict__ w,
                                  const double* __restrict__ val, int64_t nslots,
                                  const int32_t* __restrict__ slot_row, const int32_t* __restrict__ slot_col,
                                  const double* __restrict__ scale, const int64_t* __restrict__ dest,
                                  float* __restrict__ outf) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const double acc = seq_acc<4, false>(pairptr[s], pairptr[s + 1], 0.0,
                                         [&](int64_t t) { return w[pp[t]] * val[pq[t]]; });
    outf[dest[s]] = (float)(acc * scale[slot_row[s]] * scale[slot_col[s]]);
}

__global__ void k_add_beta_scaled_f32(const int64_t* __restrict__ diagpos, int32_t m, double beta,
                                      const double* __restrict__ scale, float* __restrict__ Lf) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int64_t o = diagpos[i];
    Lf[o] = (float)((double)Lf[o] + beta * scale[i] * scale[i]);
}

// s_i = 1 / sqrt(diag_i + beta)
__global__ void k_diag_scale(int32_t m, const double* __restrict__ diag, double beta, double* __restrict__ s) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) s[i] = 1.0 / sqrt(diag[i] + beta);
}

// v[iperm[i]] *= s[i]  (a vector in the factor's numbering scaled by S)
__global__ void k_scale_perm(int32_t m, double* __restrict__ v, const double* __restrict__ s,
                             const int32_t* __restrict__ iperm) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) v[iperm[i]] *= s[i];
}

__global__ void k_spmv_csr(int32_t m, const int64_t* __restrict__ rowptr,
                           const int32_t* __restrict__ colind, const double* __restrict__ val,
                           const double* __restrict__ x, double* __restrict__ y) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    y[i] = seq_acc<8, false>(rowptr[i], rowptr[i + 1], 0.0,
                             [&](int64_t t) { return val[t] * x[colind[t]]; });
}

__global__ void k_spmv_csc_t(int32_t n, const int64_t* __restrict__ colptr,
                             const int32_t* __restrict__ rowind, const double* __restrict__ val,
                             const double* __restrict__ x, double* __restrict__ y) {
    int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    y[j] = seq_acc<4, false>(colptr[j], colptr[j + 1], 0.0,
                             [&](int64_t p) { return val[p] * x[rowind[p]]; });
}

// r_i = -b2_i + sum_j A_ij (b1_j / d_j), columns ascending (cholmod_sdmult
// with alpha = beta = 1 on Y = -b2, UseCholmod.cpp:173-179), written to the
// permuted position iperm[i] (UseCholmod.cpp:180-182).
// t_j = b1_j / d_j (UseCholmod.cpp:176-177), then per row the gather of t
__global__ void k_div(int32_t n, const double* __restrict__ b1, const double* __restrict__ d,
                      double* __restrict__ t) {
    int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) t[j] = b1[j] / d[j];
}

// rp[iperm[i]] = -b2_i + sum_j A_ij t_j, over the row's columns in
// ascending order (UseCholmod.cpp:173-182: A (b1 ./ d) - b2, permuted)
__global__ void k_rhs(int32_t m, const int64_t* __restrict__ rowptr,
                      const int32_t* __restrict__ colind, const double* __restrict__ csrval,
                      const double* __restrict__ t, const double* __restrict__ b2,
                      const int32_t* __restrict__ iperm, double* __restrict__ rp) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double acc = seq_acc<8, false>(rowptr[i], rowptr[i + 1], -b2[i],
                                         [&](int64_t q) { return csrval[q] * t[colind[q]]; });
    rp[iperm[i]] = acc;
}

// x1_j = (b1_j - sum_i A_ij x2_i) / d_j, subtracting each term from b1_j over
// ascending rows (Helper.cpp:145-154 order; UseCholmod.cpp:226-231).
__global__ void k_backsub(int32_t n, const int64_t* __restrict__ colptr,
                          const int32_t* __restrict__ rowind, const double* __restrict__ val,
                          const double* __restrict__ b1, const double* __restrict__ d,
                          const double* __restrict__ x2, double* __restrict__ x1) {
    int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double acc = seq_acc<4, true>(colptr[j], colptr[j + 1], b1[j],
                                        [&](int64_t p) { return val[p] * x2[rowind[p]]; });
    x1[j] = acc / d[j];
}

__global__ void k_gather_perm(int32_t m, const double* __restrict__ xp,
                              const int32_t* __restrict__ iperm, double* __restrict__ x) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) x[i] = xp[iperm[i]];
}

__global__ void k_scatter_perm(int32_t m, const double* __restrict__ x,
                               const int32_t* __restrict__ iperm, double* __restrict__ xp) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) xp[iperm[i]] = x[i];
}

__global__ void k_csc_t_scaled(int32_t n, const int64_t* __restrict__ colptr,
                               const int32_t* __restrict__ rowind, const double* __restrict__ val,
                               const double* __restrict__ dinv, const double* __restrict__ x,
                               double* __restrict__ y) {
    int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double acc = seq_acc<4, false>(colptr[j], colptr[j + 1], 0.0,
                                         [&](int64_t p) { return val[p] * x[rowind[p]]; });
    y[j] = acc * dinv[j];
}

__global__ void k_csr_plus_shift(int32_t m, const int64_t* __restrict__ rowptr,
                                 const int32_t* __restrict__ colind, const double* __restrict__ val,
                                 const double* __restrict__ t, double beta,
                                 const double* __restrict__ x, double* __restrict__ y) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double acc = seq_acc<8, false>(rowptr[i], rowptr[i + 1], 0.0,
                                         [&](int64_t k) { return val[k] * t[colind[k]]; });
    y[i] = acc + beta * x[i];
}

// deterministic dot product: fixed-size grid of block partials, then one block
constexpr int kDotBlocks = 512;

__global__ __launch_bounds__(256) void k_dot_partial(const double* __restrict__ a,
                                                     const double* __restrict__ b, int64_t n,
                                                     double* __restrict__ part) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        acc = acc + a[i] * b[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_dot_final(const double* __restrict__ part, int np,
                                                   double* __restrict__ out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) acc = acc + part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// y += alpha x
__global__ void k_axpy(int64_t n, double alpha, const double* __restrict__ x, double* __restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = y[i] + alpha * x[i];
}
// y = x + beta y
__global__ void k_xpby(int64_t n, const double* __restrict__ x, double beta, double* __restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] + beta * y[i];
}
// z = x - y
__global__ void k_sub(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                      double* __restrict__ z) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = x[i] - y[i];
}

}  // namespace

void launch_dot(const double* a, const double* b, int64_t n, double* part, double* out,
                hipStream_t st) {
    int nb = (int)std::min<int64_t>(kDotBlocks, std::max<int64_t>(1, (n + 255) / 256));
    help::launch_kernel(k_dot_partial, nb, 256, 0, st, a, b, n, part);
    help::launch_kernel(k_dot_final, 1, 256, 0, st, part, nb, out);
}
void launch_axpy(int64_t n, double alpha, const double* x, double* y, hipStream_t st) {
    if (n > 0) help::launch_kernel(k_axpy, grid_for(n), kBlock, 0, st, n, alpha, x, y);
}
void launch_xpby(int64_t n, const double* x, double beta, double* y, hipStream_t st) {
    if (n > 0) help::launch_kernel(k_xpby, grid_for(n), kBlock, 0, st, n, x, beta, y);
}
void launch_sub(int64_t n, const double* x, const double* y, double* z, hipStream_t st) {
    if (n > 0) help::launch_kernel(k_sub, grid_for(n), kBlock, 0, st, n, x, y, z);
}

void launch_dinv(const double* d, double* dinv, int32_t n, hipStream_t st) {
    if (n > 0) help::launch_kernel(k_dinv, grid_for(n), kBlock, 0, st, d, dinv, n);
}

void launch_adat_assemble(const int64_t* pairptr, const int32_t* pair_p, const int32_t* pair_q,
                          const int32_t* colof, const double* val, const double* dinv, int64_t nnz,
                          double* w, int64_t nslots, const int64_t* dest, double* out_d, float* out_f,
                          hipStream_t st) {
    if (nslots <= 0) return;
    if (nnz > 0) help::launch_kernel(k_scale_nz, grid_for(nnz), kBlock, 0, st, nnz, colof, val, dinv, w);
    help::launch_kernel(k_adat, grid_for(nslots), kBlock, 0, st, pairptr, pair_p, pair_q, (const double*)w, val,
                            nslots, dest, out_d, out_f);
}

void launch_adat_scaled_f32(const int64_t* pairptr, const int32_t* pair_p, const int32_t* pair_q,
                            const int32_t* colof, const double* val, const double* dinv, int64_t nnz,
                            double* w, int64_t nslots, const int32_t* slot_row, const int32_t* slot_col,
                            const double* scale, const int64_t* dest, float* out_f, hipStream_t st) {
    if (nslots <= 0) return;
    if (nnz > 0) help::launch_kernel(k_scale_nz, grid_for(nnz), kBlock, 0, st, nnz, colof, val, dinv, w);
    help::launch_kernel(k_adat_scaled_f32, grid_for(nslots), kBlock, 0, st, pairptr, pair_p, pair_q,
                            (const double*)w, val, nslots, slot_row, slot_col, scale, dest, out_f);
}

void launch_add_beta_scaled_f32(const int64_t* diagpos, int32_t m, double beta, const double* scale,
                                float* Lf, hipStream_t st) {
    if (m > 0) help::launch_kernel(k_add_beta_scaled_f32, grid_for(m), kBlock, 0, st, diagpos, m, beta, scale, Lf);
}

void launch_diag_scale(int32_t m, const double* diag, double beta, double* s, hipStream_t st) {
    if (m > 0) help::launch_kernel(k_diag_scale, grid_for(m), kBlock, 0, st, m, diag, beta, s);
}

void launch_scale_perm(int32_t m, double* v, const double* s, const int32_t* iperm, hipStream_t st) {
    if (m > 0) help::launch_kernel(k_scale_perm, grid_for(m), kBlock, 0, st, m, v, s, iperm);
}

void launch_add_beta(const int64_t* diagpos, int32_t m, double beta, double* Ld, float* Lf,
                     hipStream_t st) {
    if (m > 0) help::launch_kernel(k_add_beta, grid_for(m), kBlock, 0, st, diagpos, m, beta, Ld, Lf);
}

void launch_spmv_csr(int32_t m, const int64_t* rowptr, const int32_t* colind, const double* val,
                     const double* x, double* y, hipStream_t st) {
    if (m > 0) help::launch_kernel(k_spmv_csr, grid_for(m), kBlock, 0, st, m, rowptr, colind, val, x, y);
}

void launch_spmv_csc_t(int32_t n, const int64_t* colptr, const int32_t* rowind, const double* val,
                       const double* x, double* y, hipStream_t st) {
    if (n > 0) help::launch_kernel(k_spmv_csc_t, grid_for(n), kBlock, 0, st, n, colptr, rowind, val, x, y);
}

void launch_rhs(int32_t m, int32_t n, const int64_t* rowptr, const int32_t* colind, const double* csrval,
                const double* b1, const double* d, const double* b2, const int32_t* iperm, double* t,
                double* rp, hipStream_t st) {
    if (n > 0) help::launch_kernel(k_div, grid_for(n), kBlock, 0, st, n, b1, d, t);
    if (m > 0) help::launch_kernel(k_rhs, grid_for(m), kBlock, 0, st, m, rowptr, colind, csrval, (const double*)t, b2,
                                   iperm, rp);
}

void launch_backsub(int32_t n, const int64_t* colptr, const int32_t* rowind, const double* val,
                    const double* b1, const double* d, const double* x2, double* x1,
                    hipStream_t st) {
    if (n > 0) help::launch_kernel(k_backsub, grid_for(n), kBlock, 0, st, n, colptr, rowind, val, b1, d, x2, x1);
}

void launch_gather_perm(int32_t m, const double* xp, const int32_t* iperm, double* x,
                        hipStream_t st) {
    if (m > 0) help::launch_kernel(k_gather_perm, grid_for(m), kBlock, 0, st, m, xp, iperm, x);
}

void launch_scatter_perm(int32_t m, const double* x, const int32_t* iperm, double* xp,
                         hipStream_t st) {
    if (m > 0) help::launch_kernel(k_scatter_perm, grid_for(m), kBlock, 0, st, m, x, iperm, xp);
}

void launch_normal_matvec(int32_t m, int32_t n, const int64_t* rowptr, const int32_t* colind,
                          const double* csrval, const int64_t* colptr, const int32_t* rowind,
                          const double* cscval, const double* dinv, double beta, const double* x,
                          double* tmp_n, double* y, hipStream_t st) {
    if (n > 0) help::launch_kernel(k_csc_t_scaled, grid_for(n), kBlock, 0, st, n, colptr, rowind, cscval, dinv, x, tmp_n);
    if (m > 0) help::launch_kernel(k_csr_plus_shift, grid_for(m), kBlock, 0, st, m, rowptr, colind, csrval, tmp_n, beta, x, y);
}

}  // namespace help
