// Internal declarations shared by the backend's HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "launch.h"

namespace help {

// Supernodal structure as uploaded to the device (all arrays device-resident).
struct DevSym {
    int32_t nsuper;
    const int32_t* sfirst;    // nsuper+1
    const int32_t* snrow;     // nsuper
    const int64_t* loff;      // nsuper+1
    const int64_t* uoff;      // nsuper+1
    const int64_t* rowoff;    // nsuper+1
    const int32_t* rowidx;
    const int64_t* reloff;    // nsuper+1
    const int32_t* relmap;
    const int32_t* childptr;  // nsuper+1
    const int32_t* child;
    const int64_t* dinvoff;   // nsuper+1 offsets into the diagonal-block inverses
    const int64_t* xoff;      // nsuper+1 offsets into the off-diagonal inverse blocks
};

// Off-diagonal 64x64 block X_ij (i > j) of inv(L11) of a supernode wider than
// one block, at xoff[s] + (i(i-1)/2 + j) * 4096, column-major, zero outside
// the supernode's columns.  The diagonal blocks are the Dinv blocks.
struct XTask {
    int32_t s, i, j, k;   // block (i, j) at the step that finalises X_kj (j <= k <= i)
};
__host__ __device__ inline int64_t xblock_offset(const int64_t* xoff, int32_t s, int32_t i, int32_t j) {
    return xoff[s] + ((int64_t)i * (i - 1) / 2 + j) * 4096;
}

// C -= A B^T task on one 64x64 tile of a front.  Front coordinates: rows and
// columns 0..nrow-1 of supernode s; columns < ncol live in the L panel, the
// rest in the update matrix U.
struct GemmTask {
    int32_t s;
    int32_t row0, col0;   // tile origin (front coordinates)
    int32_t k0, k1;       // inner range (panel columns)
    int32_t col_end;      // columns >= col_end are not updated
    // update-matrix tile built over several panels (the accumulator is parked
    // in Uacc, same layout as U, between panels): kGemmDirect C -= A B^T;
    // kGemmFirst Uacc = A B^T; kGemmNext Uacc += A B^T; kGemmLast
    // C -= Uacc + A B^T.  The MFMA accumulation runs in the same k order as
    // one product over the whole panel, so the result is the same.
    int32_t mode = 0;
    // the front's sizes and storage offsets, copied from the symbolic
    // structure when the schedule is built: the task's first load carries
    // everything its addresses depend on
    int32_t n = 0, c = 0;
    int64_t loff = 0, uoff = 0;
};
enum : int32_t { kGemmDirect = 0, kGemmFirst = 1, kGemmNext = 2, kGemmLast = 3 };

// Panel task: factor the diagonal block [b, b+nb) of front s (every task of
// the step does it redundantly in LDS) and solve the rows of block rb below it.
struct PanelTask {
    int32_t s;
    int32_t b;            // first panel column
    int32_t rb;           // 0: write back the diagonal block; >=1: row block rb-1
};

// Extend-add task: add every child's update matrix into the columns
// [c0, c1) of the front of supernode s (kEaCols columns per task: the
// launches of the large fronts need many workgroups to cover the load
// latency of the scattered read-modify-writes).
constexpr int kEaCols = 8;
struct EaTask {
    int32_t s;
    int32_t c0, c1;
    int32_t r0, nref;     // refs[r0 .. r0+nref): the children with columns in [c0, c1), in child order
    int32_t n, c;         // the front's columns and rows
    int64_t loff, uoff;   // the front's storage offsets
};
// One child of an extend-add task: where its update matrix and relative map
// live, its update rows, and its columns [lo, hi) that land in the task's
// range (the relative map is ascending).
struct EaRef {
    int64_t uoff;
    int64_t reloff;
    int32_t chc;
    int32_t lo, hi;
    int32_t pad = 0;
};

constexpr int kTile = 64;      // GEMM tile and panel width
// most workgroups of a fused FP64 panel launch (one per CU: 104 KB of LDS)
constexpr int kFusedPanelMax = 240;
constexpr int kSmallRows = 256;
constexpr int kSmallElems = 4032;  // panel elements of a small front; with the 64x65
                                   // inverse scratch at most 8192 (64 KB FP64) in LDS

// ---- sparse kernels (kernels_sparse.hip) ----
void launch_dinv(const double* d, double* dinv, int32_t n, hipStream_t st);
// w (nnz doubles) receives A_ij * dinv_j per nonzero on the way
void launch_adat_assemble(const int64_t* pairptr, const int32_t* pair_p, const int32_t* pair_q,
                          const int32_t* colof, const double* val, const double* dinv, int64_t nnz,
                          double* w, int64_t nslots, const int64_t* dest, double* out_d, float* out_f,
                          hipStream_t st);
void launch_add_beta(const int64_t* diagpos, int32_t m, double beta, double* Ld, float* Lf,
                     hipStream_t st);
// FP32 factor of S (A D^-1 A^T + beta I) S with S = diag(1/sqrt(diag(A D^-1 A^T) + beta))
void launch_adat_scaled_f32(const int64_t* pairptr, const int32_t* pair_p, const int32_t* pair_q,
                            const int32_t* colof, const double* val, const double* dinv, int64_t nnz,
                            double* w, int64_t nslots, const int32_t* slot_row, const int32_t* slot_col,
                            const double* scale, const int64_t* dest, float* out_f, hipStream_t st);
void launch_add_beta_scaled_f32(const int64_t* diagpos, int32_t m, double beta, const double* scale,
                                float* Lf, hipStream_t st);
void launch_diag_scale(int32_t m, const double* diag, double beta, double* s, hipStream_t st);
void launch_scale_perm(int32_t m, double* v, const double* s, const int32_t* iperm, hipStream_t st);
void launch_spmv_csr(int32_t m, const int64_t* rowptr, const int32_t* colind, const double* val,
                     const double* x, double* y, hipStream_t st);
void launch_spmv_csc_t(int32_t n, const int64_t* colptr, const int32_t* rowind, const double* val,
                       const double* x, double* y, hipStream_t st);
// t (n doubles) receives b1 ./ d on the way
void launch_rhs(int32_t m, int32_t n, const int64_t* rowptr, const int32_t* colind, const double* csrval,
                const double* b1, const double* d, const double* b2, const int32_t* iperm, double* t,
                double* rp, hipStream_t st);
void launch_backsub(int32_t n, const int64_t* colptr, const int32_t* rowind, const double* val,
                    const double* b1, const double* d, const double* x2, double* x1,
                    hipStream_t st);
void launch_gather_perm(int32_t m, const double* xp, const int32_t* iperm, double* x,
                        hipStream_t st);
void launch_scatter_perm(int32_t m, const double* x, const int32_t* iperm, double* xp,
                         hipStream_t st);
// FP64 vector kernels (deterministic dot: fixed grid of partials + one block)
constexpr int kDotPartials = 512;
void launch_dot(const double* a, const double* b, int64_t n, double* part /* kDotPartials */,
                double* out, hipStream_t st);
void launch_axpy(int64_t n, double alpha, const double* x, double* y, hipStream_t st);   // y += a x
void launch_xpby(int64_t n, const double* x, double beta, double* y, hipStream_t st);    // y = x + b y
void launch_sub(int64_t n, const double* x, const double* y, double* z, hipStream_t st); // z = x - y
// y = (A D^-1 A^T + beta I) x in original numbering (refinement operator)
void launch_normal_matvec(int32_t m, int32_t n, const int64_t* rowptr, const int32_t* colind,
                          const double* csrval, const int64_t* colptr, const int32_t* rowind,
                          const double* cscval, const double* dinv, double beta, const double* x,
                          double* tmp_n, double* y, hipStream_t st);

// ---- supernodal factorization (kernels_factor.hip) ----
// lds_elems: dynamic LDS per workgroup in elements of T, at least
// max over the fronts of c*n + min(64,n)*(min(64,n)+1)
template <typename T>
void launch_small_fronts(const int32_t* list, int32_t count, DevSym sym, T* L, T* U, T* Dinv,
                         int32_t* err, int32_t lds_elems, hipStream_t st);
// the same for fronts of at most 64 rows, one wave each (lds_elems per front);
// lanes = 32: fronts of at most 32 rows, two per wave
template <typename T>
void launch_small_wave(const int32_t* list, int32_t count, DevSym sym, T* L, T* U, T* Dinv,
                       int32_t* err, int32_t lds_elems, int32_t lanes, hipStream_t st);
constexpr int kSmallWaveLds = 2048;   // LDS elements per front of the one-wave form
template <typename T>
void launch_extend_add(const EaTask* tasks, int32_t count, const EaRef* refs, DevSym sym, T* L, T* U,
                       int32_t depth, hipStream_t st);
constexpr int kEaDeepRows = 2048;   // launches with a front this tall take 8 rows per lane in flight
// factor the 64x64 diagonal block at column b and store its inverse
template <typename T>
void launch_potrf_diag(const PanelTask* tasks, int32_t count, DevSym sym, T* L, T* Dinv,
                       int32_t* err, hipStream_t st);
// FP64: the diagonal block and one row block per task (PanelTask.rb >= 1;
// rb = 0 for a panel with no rows below), each task factoring the block
// redundantly; the task with rb <= 1 stores Dinv and L11, the latter into
// Lstage (Dinv layout), because the other tasks of the launch may still be
// reading the unfactored block from L
void launch_panel_fp64(const PanelTask* tasks, int32_t count, DevSym sym, double* L, double* Dinv,
                       double* Lstage, int32_t* err, hipStream_t st);
// L11 blocks (s, b) from Lstage into L
void launch_copy_diag_fp64(const PanelTask* tasks, int32_t count, DevSym sym, const double* Lstage,
                           double* L, hipStream_t st);
// panel rows below the diagonal block times the stored inverse (MFMA)
template <typename T>
void launch_trsm_rows(const PanelTask* tasks, int32_t count, DevSym sym, T* L, const T* Dinv,
                      hipStream_t st);
template <typename T>
void launch_gemm(const GemmTask* tasks, int32_t count, DevSym sym, T* L, T* U, T* Uacc, hipStream_t st);
template <typename T>
void launch_mfma_probe(const T* a, const T* b, T* out, hipStream_t st);
// off-diagonal inverse blocks X_ij = -Dinv_i S_ij, S_ij = sum_{k=j}^{i-1} L_ik X_kj
// (X_jj = Dinv_j), right-looking: the launch for step t finalises X_kj,
// k = j + t, and adds L_ik X_kj to the partial sums S_ij of every i > k (S is
// written at t = 0 and accumulated in k order afterwards)
template <typename T>
void launch_xinv(const XTask* tasks, int32_t count, DevSym sym, const T* L, const T* Dinv, T* X, T* S,
                 hipStream_t st);

// ---- supernodal triangular solves (kernels_solve.hip) ----
constexpr int kMaxSnCols = 2048;   // supernodes are split to at most this many columns

// Row/column block task of a big supernode in the triangular solves.
struct SolveTask {
    int32_t s;
    int32_t blk;          // 64-row block of L21 (forward) / 64-column block (backward)
};
constexpr int kTinyRows = 64;      // supernodes with at most this many rows take one wave
// The forward solve L y = b reads b and writes y, the backward solve
// L^T x = y reads y and writes x (x may share b's storage).  Per level:
// launch_fwd_level, then launch_fwd_gemv for the L21 products of big
// supernodes; launch_bwd_gemv, then launch_bwd_level.
// launch_fwd_level / launch_bwd_level: `list` holds the level's supernodes
// that are not wide, its first `ntiny` with snrow <= kTinyRows (one wave
// each, four per workgroup), the rest one workgroup each; `xt` holds one task
// per 64-row (forward) or 64-column (backward) block of the level's wide
// supernodes, solved with the explicit inverse of L11:
//   forward  y_s(block i) = sum_{j<=i} X_ij g_s(block j), g_s = b_s minus the
//            children's update vectors;
//   backward x_s(block j) = sum_{i>=j} X_ij^T (y_s - t_s)(block i).
// Supernodes with big[s] != 0 leave their L21 products to the gemv kernels.
// tiny_lanes = 32 (a level of tiny supernodes only, all with at most 32
// rows): two supernodes per wave, 32 lanes each, same per-lane arithmetic.
// Every launch solves nrhs right-hand sides at once (blockIdx.y): right-hand
// side r uses the vectors at offset r * rs.v and the update vectors at
// offset r * rs.w, with the same operations as a single solve.
struct RhsStride {
    int64_t v = 0;   // b, y, x, t
    int64_t w = 0;   // W
};
template <typename T>
void launch_fwd_level(const int32_t* list, int32_t count, int32_t ntiny, const SolveTask* xt, int32_t nx,
                      DevSym sym, const int8_t* big, const T* L, const T* Dinv, const T* X, const double* b,
                      double* y, double* W, int nrhs, RhsStride rs, hipStream_t st, int tiny_lanes = 64);
template <typename T>
void launch_bwd_level(const int32_t* list, int32_t count, int32_t ntiny, const SolveTask* xt, int32_t nx,
                      DevSym sym, const int8_t* big, const T* L, const T* Dinv, const T* X, const double* y,
                      const double* t, double* x, int nrhs, RhsStride rs, hipStream_t st, int tiny_lanes = 64);
// Forward, big supernodes: W_s += L21 y_s by 64-row blocks.
template <typename T>
void launch_fwd_gemv(const SolveTask* tasks, int32_t count, DevSym sym, const T* L,
                     const double* y, double* W, int nrhs, RhsStride rs, hipStream_t st);
// Backward, big supernodes: t[f+j] = sum_{i>=n} L21(i,j) x[row_i] by 16-column blocks.
template <typename T>
void launch_bwd_gemv(const SolveTask* tasks, int32_t count, DevSym sym, const T* L,
                     const double* x, double* t, int nrhs, RhsStride rs, hipStream_t st);

}  // namespace help
