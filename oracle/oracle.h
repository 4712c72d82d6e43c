/*
 * oracle.h — CPU FP64 restatement of HELP's normal-equation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * backend: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it, and never as the thing measured or shipped.
 *
 * Every function cites the HELP source it follows (file:line).  Sparse
 * matrices are 0-based; A is m x n in CSC (colptr int64, rowind int32,
 * rows ascending within a column, as HELP's presolve leaves them,
 * Presolve.cpp:60-75,111-112).
 */
#ifndef HELP_ORACLE_H
#define HELP_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- CSR of A (row lists sorted by column, Presolve.cpp:43-58) ---- */
void or_csc_to_csr(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                   int64_t* rowptr, int32_t* colind, int64_t* cscpos);

/* ---- ADA^T symbolic, Helper.cpp:273-329 (linker form) ---- */
int64_t or_adat_count(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                      const int64_t* rowptr, const int32_t* colind, const int64_t* cscpos,
                      int64_t* adat_rowptr);
void or_adat_fill(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                  const int64_t* rowptr, const int32_t* colind, const int64_t* cscpos,
                  const int64_t* adat_rowptr, int32_t* adat_colind);
/* ---- ADA^T numeric, Helper.cpp:350-377 with dinv = 1/d (UsePardiso.cpp:172-174) ---- */
void or_adat_values(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                    const double* val, const int64_t* rowptr, const int32_t* colind,
                    const int64_t* cscpos, const double* d, const int64_t* adat_rowptr,
                    const int32_t* adat_colind, double* adat_val);

/* ---- SpMV, Helper.cpp:57-81 (CSRMV_N / CSRMV_T) ---- */
void or_csrmv_n(int32_t nrow, double alpha, const double* val, const int32_t* colind,
                const int64_t* rowptr, const double* x, double beta, double* y);
void or_csrmv_t(int32_t nrow, int32_t ncol, double alpha, const double* val,
                const int32_t* colind, const int64_t* rowptr, const double* x, double beta,
                double* y);
/* y_i = y_i + sign * sum_j A_ij x_j accumulated over columns in ascending
 * order (Helper.cpp:110-132; cholmod_sdmult with beta=1, UseCholmod.cpp:179) */
void or_a_times(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                const double* val, int32_t sign, const double* x, double* y);
/* y_j = y_j + sign * sum_i A_ij x_i, each column from y_j downward over
 * ascending rows (Helper.cpp:133-155) */
void or_at_times(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                 const double* val, int32_t sign, const double* x, double* y);

/* ---- symbolic factorization by column-structure merging (independent of
 *      the backend's row-subtree algorithm): pattern of L for P M P^T. ---- */
/* returns nnz(L); parent[m], colcount[m] out; Lp[m+1], Li[nnz] filled when non-NULL */
int64_t or_symbolic(int32_t m, const int64_t* adat_rowptr, const int32_t* adat_colind,
                    const int32_t* perm, int32_t* parent, int32_t* colcount,
                    int64_t* Lp, int32_t* Li);

/* ---- numeric left-looking Cholesky of P (M + beta I) P^T on the pattern
 *      from or_symbolic (semantics of cholmod_l_factorize_p with
 *      beta = {1e-5, 0}, UseCholmod.cpp:136-137).  Returns 0, or k+1 when
 *      pivot k is not positive. ---- */
int or_cholesky(int32_t m, const int64_t* adat_rowptr, const int32_t* adat_colind,
                const double* adat_val, double beta, const int32_t* perm,
                const int64_t* Lp, const int32_t* Li, double* Lx);

/* ---- x = P^T L^-T L^-1 P b  (cholmod_l_solve, UseCholmod.cpp:180-219) ---- */
void or_chol_solve(int32_t m, const int32_t* perm, const int64_t* Lp, const int32_t* Li,
                   const double* Lx, const double* b, double* x, double* work);

/* ---- SolveLinearEquation composition (UseCholmod.cpp:157-238) ---- */
void or_solve_kkt(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                  const double* val, const int32_t* perm, const int64_t* Lp,
                  const int32_t* Li, const double* Lx, const double* d, const double* b1,
                  const double* b2, double* x1, double* x2, double* work /* m+n */);

/* ---- Partial Cholesky preconditioner, PCG.cpp:38-177 ---- */
typedef struct or_pchol {
    int32_t m, k;
    int32_t* perm;     /* PCG_PERM: pivots first                              */
    double* L;         /* m x k, L[row*k + col] (PARTIAL_CHOL_L)               */
    double* D;         /* k    (PARTIAL_CHOL_D)                                */
    double* diag_s;    /* m    (DIAG_S, Schur-complement diagonal)             */
    double* l11inv;    /* k x k (L11Inv)                                       */
} or_pchol;
int  or_pchol_build(int32_t m, int32_t n, const int64_t* rowptr, const int32_t* colind,
                    const double* csrval, const double* d, double gamma, double delta,
                    int32_t k, or_pchol* out);
void or_pchol_free(or_pchol* pc);
/* ret = M^-1 r (PCG.cpp:183-196); r and ret in the PCG_PERM-permuted order */
void or_pchol_apply(const or_pchol* pc, const double* r, double* ret, double* work /* 2m */);

/* ---- PCG, PCG.cpp:296-395 (x0 given instead of zero, PCG.cpp:309-310;
 *      stop on ||r|| < tol ||r0|| or max_iter).  b, x in original order.
 *      Returns iterations. ---- */
int or_pcg(int32_t m, int32_t n, const int64_t* rowptr, const int32_t* colind,
           const double* csrval, const double* d, double gamma, double delta,
           const or_pchol* pc, const double* b, const double* x0, double* x, double tol,
           int32_t max_iter, double* final_rel_res);

/* ---- a CPU FP64 backend with CHOLMOD semantics (help_backend-compatible) ---- */
typedef struct or_backend or_backend;
or_backend* or_backend_new(double beta, const int32_t* perm /* may be NULL: natural */,
                           int32_t perm_len);
void or_backend_free(or_backend* b);
int or_be_construct(void* ctx, const void* lp /* help_lp */);
int or_be_destruct(void* ctx);
int or_be_renew(void* ctx, const double* d);
int or_be_solve(void* ctx, const double* d, const double* b1, const double* b2, double* x1,
                double* x2);
double or_be_seconds_renew(const or_backend* b);
double or_be_seconds_solve(const or_backend* b);

#ifdef __cplusplus
}
#endif
#endif
