// Matrix-free PCG on A (D + gamma I)^-1 A^T + delta I with the greedy
// partial-Cholesky preconditioner (HELP PCG.cpp:38-395), private to the
// backend.  Vectors FP64; the K factor columns are built in FP64 and kept in
// FP32 for the preconditioner apply.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace help {

constexpr int kPcgMaxK = 1024;      // partial-Cholesky columns held in LDS by the kernels
constexpr int kLightRow = 64;       // CSR rows up to this many nonzeros take one thread
constexpr int kHeavyRow = 1024;     // CSR rows longer than this are split into chunks
constexpr int kRowChunk = 4096;     // nonzeros per chunk of a heavy row
constexpr int kLtRows = 2048;       // rows per workgroup of the transposed factor product

struct RowChunk {
    int32_t row;
    int64_t b, e;                   // nonzero range [b, e) of the row
};

// Row partition of a CSR matrix for deterministic products with heavy rows:
// rows up to kLightRow nonzeros take one thread (sequential sum), rows up to
// kHeavyRow one wave (lane-strided partial sums, fixed reduction tree), the
// others are summed in chunks of kRowChunk by one workgroup each and
// combined in chunk order.
struct RowSplitDev {
    int32_t m = 0;
    int32_t nchunk = 0, nheavy = 0, nmed = 0;
    const int32_t* mrow = nullptr;   // medium rows
    const RowChunk* chunks = nullptr;
    const int32_t* hrow = nullptr;   // heavy rows
    const int32_t* hptr = nullptr;   // chunk range of heavy row h: [hptr[h], hptr[h+1])
    double* part = nullptr;          // one partial per chunk
};

void build_row_split(int32_t m, const std::vector<int64_t>& rowptr, std::vector<RowChunk>& chunks,
                     std::vector<int32_t>& hrow, std::vector<int32_t>& hptr, std::vector<int32_t>& mrow);

// y = A x (square = false) or y = (A .* A) x (square = true)
void launch_csr_split(const RowSplitDev& rs, const int64_t* rowptr, const int32_t* colind,
                      const double* val, const double* x, double* y, bool square, hipStream_t st);
// out = a ./ b
void launch_vdiv(int64_t n, const double* a, const double* b, double* out, hipStream_t st);
// t_j = dinv_j * (A^T x)_j over the CSC of A (the scaled transposed product)
void launch_csc_t_scaled(int32_t n, const int64_t* colptr, const int32_t* rowind, const double* val,
                         const double* dinv, const double* x, double* t, hipStream_t st);

// ---- partial Cholesky build (one pivot per call sequence) ----
constexpr int kArgmaxBlocks = 512;
// picks the pivot of step k: the largest S over positions [k, m) (ties to
// the lowest position), swaps it to position k, writes its row to col_out
void launch_pc_argmax(int32_t m, int32_t k, const double* S, double* pval, int32_t* pidx,
                      int32_t* perm, int32_t* pos, int32_t* col_out, hipStream_t st);
void launch_pc_scatter(const int32_t* col, int32_t max_row_nnz, const int64_t* rowptr,
                       const int32_t* colind, const double* val, const double* dinv, double* sc,
                       bool clear, hipStream_t st);
void launch_pc_pivot(int32_t m, int32_t k, const int32_t* col, const double* g, const double* L,
                     double delta, double* D, double* w, hipStream_t st);
void launch_pc_column(int32_t m, int32_t k, const int32_t* col, const int32_t* pos, const double* g,
                      const double* w, const double* D, double* L, double* S, hipStream_t st);
void launch_pc_l11(int32_t m, int32_t K, const int32_t* perm, const double* L, double* L11,
                   double* X, hipStream_t st);
void launch_to_f32(int64_t n, const double* x, float* y, hipStream_t st);

// ---- preconditioner apply z = M^-1 r (PCG.cpp:183-196) ----
// y1: scratch of 2 * kPcgMaxK doubles; ltpart: K * ceil(m / kLtRows) doubles;
// the m x K factor L32 in FP32 or FP64 (T = float / double)
template <typename T>
void launch_pc_apply(int32_t m, int32_t K, const int32_t* perm, const int32_t* pos, const T* L32,
                     const double* X, const double* S, const double* r, double* y1, double* ltpart,
                     double* z, hipStream_t st);

// ---- conjugate gradients with device-resident scalars ----
// slots of the scalar array: r.r, p.q, r.z of alternate iterations, the
// state (0 running, 1 converged, 2 breakdown) and the iterations started
constexpr int kCgRR = 0, kCgPQ = 1, kCgRZ = 2, kCgState = 4, kCgIters = 5, kCgSlots = 6;
void launch_cg_init(double* s, double rr, double rz, hipStream_t st);
// after q = M p: p.q with the breakdown test, x += alpha p and r -= alpha q
// with r.r and the convergence test sqrt(r.r) <= target
void launch_cg_step1(int64_t n, double* s, int rz_cur, double target, const double* p, const double* q, double* x,
                     double* r, double* part, hipStream_t st);
// after z = P^-1 r: r.z into slot rz_next, p = z + (r.z new / r.z old) p
void launch_cg_step2(int64_t n, double* s, int rz_cur, int rz_next, const double* r, const double* z, double* p,
                     double* part, hipStream_t st);

struct PcgState;
// host side (pcg.hip): allocation at create, the partial-Cholesky build at
// renew, the conjugate-gradient solve (rhs in, x2 out, device vectors)
void pcg_free(PcgState* p);

}  // namespace help
