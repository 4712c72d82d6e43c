/*
 * help_mi355x.h — C ABI of the MI355X normal-equation backend for HELP.
 *
 * The backend solves the per-iteration KKT system of HELP's homogeneous
 * self-dual interior-point driver,
 *
 *     [ D  A^T ] [x1]   [b1]        (A D^-1 A^T + beta I) x2 = A D^-1 b1 - b2
 *     [ A   0  ] [x2] = [b2]   <=>  x1 = D^-1 (b1 - A^T x2)
 *
 * with beta = 1e-5 on the factor only.  Every entry point is a blocking,
 * host-synchronous call on one HIP stream owned by the handle; all status
 * codes are ints (0 = success) and the library never calls exit().
 *
 * Reference interface replaced (HELP, LP.h:157-160, one backend linked per
 * binary, Makefile:18-27):
 *   int  LinearEquation_Construct();                       -> help_gpu_create
 *   int  LinearEquation_Destruct();                        -> help_gpu_destroy
 *   void RenewLinearEquation(double* d);                   -> help_gpu_renew
 *   int  SolveLinearEquation(d, b_1, b_2, x_1, x_2);       -> help_gpu_solve
 * The adapter translation unit help_amd/csrc/UseMI355X.cpp implements those
 * four C++ symbols over this ABI.
 */
#ifndef HELP_MI355X_H
#define HELP_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* status codes                                                        */
/* ------------------------------------------------------------------ */
#define HELP_OK                 0
#define HELP_ERR_ARG            1   /* bad argument / size               */
#define HELP_ERR_HIP            2   /* HIP runtime error                 */
#define HELP_ERR_NOT_PD         3   /* non-positive pivot in the factor  */
#define HELP_ERR_STATE          4   /* call out of order (e.g. solve before renew) */
#define HELP_ERR_NOMEM          5
#define HELP_ERR_NOCONV         6   /* refinement / PCG did not converge */
#define HELP_ERR_NOGPU          7   /* no HIP device visible              */

/* ------------------------------------------------------------------ */
/* solve modes                                                         */
/* ------------------------------------------------------------------ */
#define HELP_MODE_DIRECT_FP64    0  /* FP64 supernodal factor, CHOLMOD semantics  */
#define HELP_MODE_DIRECT_FP32    1  /* FP32 factor + FP64 iterative refinement    */
#define HELP_MODE_PCG            2  /* matrix-free PCG, partial-Cholesky precond. */
#define HELP_MODE_PCG_LEARNED    3  /* PCG with message-passing network x0        */

/* fault-injection flags (host-forced; every kernel still exits normally) */
#define HELP_FAULT_FP32_BREAKDOWN  1  /* treat the FP32 factor as broken down     */
#define HELP_FAULT_REFINE_NOCONV   2  /* treat refinement as non-convergent       */
#define HELP_FAULT_NET_NONFINITE   4  /* treat the network guess as non-finite    */

typedef struct help_opts {
    int32_t mode;            /* HELP_MODE_*                                     */
    double  beta;            /* diagonal shift of the factor (HELP: 1e-5)       */
    int32_t refine_max;      /* max FP64 refinement steps (FP32 mode)           */
    double  refine_tol;      /* stop when ||r||/||rhs|| <= refine_tol           */
    int32_t pcg_k;           /* partial-Cholesky columns (HELP PCG.h:20 uses 10)*/
    int32_t pcg_max_iter;    /* PCG iteration cap                               */
    double  pcg_tol;         /* PCG stop: ||r|| <= pcg_tol * ||r0||             */
    int32_t pcg_fp32_precond;/* 1: apply the partial-Cholesky factor in FP32,
                                0: in FP64 (vectors FP64 either way)          */
    int32_t fault_flags;     /* HELP_FAULT_*                                    */
    int32_t verbose;
    int32_t symbolic_only;   /* 1: analysis only, no device state (host tests)  */
    int32_t front_path;      /* 0 auto; 1 every front through the tiled path;
                                2 (testing) every small front that fits one
                                wave takes the 64-lane one-wave kernel        */
    int32_t profile;         /* 1: kernel-class profiling from creation on      */
    int32_t refine_fp64;     /* FP64 mode: refinement steps on the FP64 factor
                                (0: one factor solve, as UseCholmod.cpp does)  */
    int32_t inverse_solve;   /* supernodes wider than 64 columns solve their
                                dense triangle with an explicit inverse of L11
                                (formed during the factorization on a second
                                stream): 0 or 1 on, -1 off (block substitution
                                with the 64x64 diagonal-block inverses)        */
    int32_t ordering;        /* fill-reducing ordering of A D^-1 A^T:
                                HELP_ORDER_AMD (CHOLMOD's setting in HELP,
                                Setting_CPU.cpp:21-26) or HELP_ORDER_ND (nested
                                dissection, the family of the METIS ordering
                                HELP's PARDISO backend asks for,
                                UsePardiso.cpp:101)                            */
    int32_t factor_graph;    /* 1: capture the factorization's launches once per
                                precision as a HIP graph and replay it on every
                                renew (0: launch them one by one)             */
    double  refine_switch;   /* FP32 mode: when the solves since the last renew
                                took more than this many refinement steps on
                                average, the next renews factor in FP64 (d's
                                spread only grows along an IPM run, and each
                                step costs a triangular-solve pair); 0: only a
                                refinement failure switches                    */
    int32_t ea_depth;        /* (testing) extend-add child rows per lane in
                                flight: 0 auto (8 in the launches that start a
                                front of at least 2048 rows, else 4); 4 or 8
                                for every launch                               */
} help_opts;
#define HELP_ORDER_AMD 0
#define HELP_ORDER_ND  1

/* Defaults: FP64 direct, beta=1e-5, refine 20 / 1e-12 (FP32 mode; FP64 mode
 * refine_fp64 0), refine_switch 2.5, pcg_k 500, pcg_max_iter 5000, pcg_tol
 * 1e-10, FP32 preconditioner, AMD ordering. */
void help_opts_default(help_opts* o);

typedef struct help_solve_stats {
    int32_t refine_steps;    /* FP64 refinement steps taken                     */
    int32_t pcg_iters;       /* PCG iterations taken                            */
    int32_t fallback;        /* 0 none, 1 FP32->FP64 factor, 2 refine->PCG,
                                3 network x0 -> zero x0                         */
    double  rel_residual;    /* FP64 ||rhs - M x2|| / ||rhs|| at exit           */
    double  ms_rhs, ms_solve, ms_back; /* phase times (HIP events)              */
} help_solve_stats;

typedef struct help_renew_stats {
    double  ms_assemble;     /* ADA^T numeric assembly                          */
    double  ms_factor;       /* numeric factorization                           */
    double  ms_precond;      /* partial-Cholesky build (PCG modes)              */
    int32_t n_tiny_pivots;   /* pivots replaced in the FP32 factor              */
    int32_t factor_bits;     /* 32 or 64: precision of this renew's factor      */
} help_renew_stats;

typedef struct help_gpu help_gpu;   /* opaque handle */

/* A in CSC: colptr[n+1], rowind[nnz] (ascending within each column, as
 * HELP's presolve leaves them: Presolve.cpp:60-75,111-112), val[nnz].
 * Replaces LinearEquation_Construct (UseCholmod.cpp:29-58). Runs the
 * symbolic analysis (ADA^T pattern, ordering, etree, supernodes) once. */
int help_gpu_create(int32_t m, int32_t n, int64_t nnz,
                    const int64_t* colptr, const int32_t* rowind,
                    const double* val, const help_opts* opts,
                    help_gpu** out);

/* Replaces LinearEquation_Destruct (UseCholmod.cpp:60-81). Safe on NULL. */
int help_gpu_destroy(help_gpu* h);

/* Replaces RenewLinearEquation (UseCholmod.cpp:84-149): assemble
 * A D^-1 A^T + beta I from d[n] (> 0) and factor it. */
int help_gpu_renew(help_gpu* h, const double* d, help_renew_stats* st);

/* Replaces SolveLinearEquation (UseCholmod.cpp:157-238): d[n] (same values
 * as the last renew), b1[n], b2[m] in; x1[n], x2[m] out.  Host or device
 * pointers (unified addressing), here and in help_gpu_renew.  In the PCG
 * modes a solve that reaches pcg_max_iter short of pcg_tol writes its last
 * iterate and returns HELP_ERR_NOCONV (the reference's PCG has no cap,
 * PCG.h:19). */
int help_gpu_solve(help_gpu* h, const double* d, const double* b1,
                   const double* b2, double* x1, double* x2,
                   help_solve_stats* st);

/* Two SolveLinearEquation calls (UseCholmod.cpp:157-238) with the same d
 * in one pass: (b1a, b2a) -> (x1a, x2a) and (b1b, b2b) -> (x1b, x2b).  In
 * the plain FP64 direct mode (refine_fp64 = 0) both right-hand sides go
 * through each triangular-solve launch together; with the FP32 factor the
 * two refinements run in lockstep over shared triangular solves (when either
 * misses refine_tol the two single solves run instead, so a refinement
 * failure switches to the FP64 factor as in help_gpu_solve).  The results
 * are identical to two help_gpu_solve calls; the PCG modes make two
 * help_gpu_solve calls.
 * HELP's driver calls the backend once per solve, so the reference has no
 * counterpart: this serves drivers that know two right-hand sides at once
 * (help_hsd_solve_gpu's bottom and predictor solves). */
int help_gpu_solve2(help_gpu* h, const double* d,
                    const double* b1a, const double* b2a, double* x1a, double* x2a,
                    const double* b1b, const double* b2b, double* x1b, double* x2b,
                    help_solve_stats* sta, help_solve_stats* stb);

/* Text of the last error on this handle ("" when none). */
const char* help_gpu_last_error(const help_gpu* h);
/* Text of the last error of help_gpu_create when it returned no handle. */
const char* help_last_create_error(void);

/* Number of visible HIP devices (0 on a host without a GPU). */
int help_device_count(void);

/* ------------------------------------------------------------------ */
/* Introspection used by the parity tests (host arrays in and out).    */
/* ------------------------------------------------------------------ */

/* Sizes of the symbolic analysis. */
typedef struct help_symbolic_info {
    int64_t nnz_adat_upper;  /* upper-triangle nnz of A A^T incl. diagonal     */
    int64_t nnz_l;           /* nnz of L (supernodal, incl. relaxed zeros)     */
    int64_t nnz_l_exact;     /* nnz of L from the exact column counts          */
    int64_t n_supernodes;
    int64_t n_levels;        /* etree levels of the supernodal schedule        */
    int64_t max_front;       /* largest supernode row count                     */
    double  factor_flops;    /* sum over columns of colcount^2                  */
    int64_t n_pairs;         /* ADA^T contribution pairs sum_j c_j(c_j+1)/2     */
} help_symbolic_info;
int help_gpu_symbolic_info(const help_gpu* h, help_symbolic_info* info);

/* perm[m] (new -> old), parent[m] (etree of the permuted matrix, -1 root),
 * colcount[m] (exact column counts of L incl. diagonal). Any may be NULL. */
int help_gpu_symbolic_arrays(const help_gpu* h, int32_t* perm,
                             int32_t* parent, int32_t* colcount);

/* ADA^T upper pattern in HELP's row order (diagonal first, then discovery
 * order; Helper.cpp:273-329): rowptr[m+1], colind[nnz_adat_upper]. */
int help_gpu_adat_pattern(const help_gpu* h, int64_t* rowptr, int32_t* colind);

/* Run only the numeric assembly on the GPU and return the values of the
 * upper pattern above (no beta shift), in that slot order. */
int help_gpu_adat_values(help_gpu* h, const double* d, double* vals);

/* y = A x (trans=0; x[n], y[m]) or y = A^T x (trans=1; x[m], y[n]) on the
 * GPU, summing each output in ascending index order from zero. */
int help_gpu_spmv(help_gpu* h, int32_t trans, const double* x, double* y);

/* Dense dump of column j (permuted numbering) of L after renew:
 * rows[nr], vals[nr]; returns nr via *nr (<= colcount bound). */
int help_gpu_factor_column(help_gpu* h, int32_t j, int32_t* rows,
                           double* vals, int32_t* nr);

/* Kernel-class profiling: with profiling on, every launch of the library is
 * bracketed by HIP events on the handle's stream and its time and algorithmic
 * work (flops for the factor classes, HBM bytes for the others) accumulate
 * per class.  help_gpu_profile resets the counters. */
#define HELP_K_SMALL_FRONTS  0   /* one workgroup per small front (flops)   */
#define HELP_K_EXTEND_ADD    1   /* large-front extend-add (bytes)          */
#define HELP_K_POTRF         2   /* 64x64 diagonal-block Cholesky (flops)   */
#define HELP_K_TRSM          3   /* panel row-block solve (flops)           */
#define HELP_K_GEMM          4   /* MFMA SYRK/GEMM update (flops)           */
#define HELP_K_FWD           5   /* forward solve per supernode (bytes)     */
#define HELP_K_FWD_GEMV      6   /* forward L21 product (bytes)             */
#define HELP_K_BWD_GEMV      7   /* backward L21^T product (bytes)          */
#define HELP_K_BWD           8   /* backward solve per supernode (bytes)    */
#define HELP_K_ADAT          9   /* ADA^T numeric assembly (bytes)          */
#define HELP_K_RHS          10   /* r = A D^-1 b1 - b2 (bytes)              */
#define HELP_K_BACKSUB      11   /* x1 = D^-1 (b1 - A^T x2) (bytes)         */
#define HELP_K_PCG_BUILD    12   /* PCG mode: partial-Cholesky build per renew */
#define HELP_K_PCG_SOLVE    13   /* PCG mode: conjugate gradients per solve  */
#define HELP_K_XINV         14   /* inverse blocks of wide supernodes (flops) */
#define HELP_K_FWD_XINV     15   /* forward solve with the inverse (bytes)   */
#define HELP_K_BWD_XINV     16   /* backward solve with the inverse (bytes)  */
#define HELP_K_COUNT        17
/* ------------------------------------------------------------------ */
/* learned warm start for the PCG mode (no reference counterpart: the  */
/* initial guess of PCG.cpp:309-310 is zero)                           */
/* ------------------------------------------------------------------ */
/* A bipartite message-passing network on A's row/column graph (4 layers,
 * width 64) predicts x0 ~ (A D^-1 A^T)^-1 r from d and r.  bf16 = 1 keeps
 * activations and the weights the kernels use in BF16 (FP32 accumulation);
 * bf16 = 0 keeps them in FP32.  Training: loss_grad (forward + backward on
 * the handle's matrix, target dy) then adam_step. */
typedef struct help_mpnn help_mpnn;
int  help_mpnn_num_params(void);
int  help_mpnn_new(int32_t bf16, uint64_t seed, help_mpnn** out);
void help_mpnn_free(help_mpnn* net);
int  help_mpnn_get_params(const help_mpnn* net, float* out);   /* FP32 master weights */
int  help_mpnn_set_params(help_mpnn* net, const float* in);
/* x0[m] = net(A, d, r) on the handle's matrix; host or device pointers */
int  help_mpnn_predict(help_mpnn* net, help_gpu* h, const double* d, const double* r, double* x0);
/* loss sum_i (y_i - t_i)^2 / sum_i t_i^2 (t_i = dy_i g_i / rms(r)) and its
 * gradient (num_params floats, may be NULL) */
int  help_mpnn_loss_grad(help_mpnn* net, help_gpu* h, const double* d, const double* r,
                         const double* dy, double* loss, float* grad);
int  help_mpnn_adam_step(help_mpnn* net, double lr);
/* HELP_MODE_PCG_LEARNED: the network whose guess starts PCG (the handle
 * does not own it; NULL or a non-finite guess starts from zero and the
 * solve reports fallback = 1) */
int  help_gpu_set_warmstart(help_gpu* h, help_mpnn* net);

/* PCG mode: the partial-Cholesky preconditioner of the last renew.  K is
 * the number of pivots; perm[m] the pivot order (pivots first, PCG_PERM),
 * D[K] the pivots, S[m] the Schur diagonal by row (DIAG_S), L[m*K] the
 * factor columns, column-major by original row (L[c*m + row]).  Any output
 * pointer may be NULL. */
int help_gpu_pcg_factor(help_gpu* h, int32_t* K, int32_t* perm, double* D, double* S, double* L);
/* PCG modes: z[m] = M^-1 r[m] with the partial-Cholesky preconditioner of
 * the last renew, applied as every CG iteration applies it (PCG.cpp:183-196:
 * L^-T D_S^-1 L^-1 with L = [L11 0; L21 I]; FP32 factor, FP64 vectors). */
int help_gpu_pcg_apply(help_gpu* h, const double* r, double* z);
/* help_gpu_pcg_apply without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  r and z must be
 * device memory (HELP_ERR_ARG otherwise); the apply's scratch lives in a
 * static 8 MiB device array per GPU, which limits it to
 * 2 * 1024 + K * ceil(m / 2048) doubles (HELP_ERR_ARG beyond that: K = 500
 * reaches m = 4.2 million).  Calls that overlap on one GPU share that array
 * and race.  Blocking, on the handle's stream, like every entry point. */
int help_gpu_pcg_apply_noalloc(help_gpu* h, const double* r, double* z);

int help_gpu_profile(help_gpu* h, int32_t enable);
int help_gpu_kernel_stats(const help_gpu* h, int32_t kind, double* ms, int64_t* launches,
                          double* work, const char** name);
/* Synchronize the current HIP device (all streams). */
int help_device_synchronize(void);

/* Supernode partition: sfirst[nsuper+1], snrow[nsuper], level[nsuper],
 * tiled[nsuper] (1 when the front takes the tiled many-workgroup path). */
int help_gpu_supernodes(const help_gpu* h, int32_t* sfirst, int32_t* snrow,
                        int32_t* level, int32_t* tiled);

/* Lane-layout probe of the 16x16x4 MFMA used by the factor (fp64 != 0:
 * V_MFMA_F64_16X16X4_F64, else V_MFMA_F32_16X16X4_F32): lane l feeds
 * a[l] and b[l]; out[4*l + r] receives element r of lane l's result. */
int help_selftest_mfma(int32_t fp64, const double* a, const double* b, double* out);

/* Lane-layout probe of V_MFMA_F32_16X16X32_BF16 (the warm-start network's
 * dense layers): lane l feeds a[8l..8l+7] and b[8l..8l+7] (values exact in
 * BF16); out[4*l + r] receives element r of lane l's result. */
int help_selftest_mfma_bf16(const double* a, const double* b, double* out);

/* ------------------------------------------------------------------ */
/* Instance generators (host; deterministic splitmix64/xoshiro256**).  */
/* ------------------------------------------------------------------ */
typedef struct help_lp {
    int32_t m, n;            /* rows, columns                                   */
    int64_t nnz;
    int32_t n_ub, n_lb, n_fr;/* HELP column layout: [UB | LB-only | free]        */
    int64_t* colptr;         /* n+1                                             */
    int32_t* rowind;         /* nnz, ascending per column                       */
    double*  val;            /* nnz                                             */
    double*  b;              /* m                                               */
    double*  c;              /* n                                               */
    double*  ub;             /* n_ub (may be NULL when n_ub = 0)                */
    double   c_intercept;
} help_lp;

#define HELP_GEN_RANDOM        0   /* A=[B|I], B m x nb with k nnz per column   */
#define HELP_GEN_MCF           1   /* multicommodity flow on a grid              */
#define HELP_GEN_STAIRCASE     2   /* staircase, T periods                        */
#define HELP_GEN_ZIPF          3   /* A=[B|I], Zipf(s) row popularity            */

typedef struct help_gen_params {
    int32_t kind;            /* HELP_GEN_*                                      */
    uint64_t seed;
    int32_t m, nb, k;        /* RANDOM / ZIPF: rows, B columns, nnz per column  */
    double  zipf_s;          /* ZIPF exponent                                   */
    int32_t side, commodities, pairs; /* MCF                                    */
    int32_t periods, block_rows, block_cols, block_k; /* STAIRCASE              */
    int32_t ub_cols;         /* RANDOM: the first ub_cols columns of B get an
                                upper bound (HELP's n_UB block)                 */
    int32_t free_cols;       /* RANDOM: free_cols extra free columns after the
                                identity (HELP's n_FR block)                    */
} help_gen_params;

/* Allocates the arrays of *lp (free with help_lp_free). */
int  help_generate(const help_gen_params* p, help_lp* lp);
void help_lp_free(help_lp* lp);
/* Write *lp in fixed-format MPS readable by HELP's MPSRead.cpp, with a
 * BOUNDS section (UP for the n_ub bounded columns, FR for the n_fr free
 * ones) that Crushing.cpp and Presolve_FinalizeModel turn back into the
 * same column layout. */
int  help_write_mps(const help_lp* lp, const char* path, const char* name);
/* Binary CSC instance file for LPs beyond the fixed-MPS limits (HELP's
 * MAX_ROWS / MAX_COLS / MAX_ELEMENTS, LP.h:61-63): magic "HELPCSC1", the
 * sizes and layout counts, colptr, rowind, val, b, c, ub, and an FNV-1a
 * checksum.  help_read_csc_bin allocates *lp (free with help_lp_free) and
 * returns HELP_ERR_ARG on a missing, truncated, corrupted or inconsistent
 * file. */
int  help_write_csc_bin(const help_lp* lp, const char* path);
int  help_read_csc_bin(const char* path, help_lp* lp);

/* ------------------------------------------------------------------ */
/* HSD driver (restatement of HELP HSDSolver.cpp) over any backend.    */
/* ------------------------------------------------------------------ */
typedef struct help_backend {
    void* ctx;
    int (*construct)(void* ctx, const help_lp* lp);
    int (*destruct)(void* ctx);
    int (*renew)(void* ctx, const double* d);
    int (*solve)(void* ctx, const double* d, const double* b1,
                 const double* b2, double* x1, double* x2);
} help_backend;

#define HELP_HSD_OK            0
#define HELP_HSD_STALLED       1
#define HELP_HSD_ITER_LIMIT    2
#define HELP_HSD_UNKNOWN_ERROR 3

typedef struct help_hsd_result {
    int32_t status;          /* HELP_HSD_* (LP.h:163-167)                        */
    int32_t iterations;
    double  primal_obj, dual_obj, rel_gap, primal_infeas, dual_infeas, mu;
    double  seconds;         /* wall time of the HSD_Main span                    */
    double  seconds_renew, seconds_solve;
    double  seconds_construct;   /* backend construct (symbolic analysis, setup)  */
    double  seconds_newton;      /* driver: Newton parameters and residuals       */
    double  seconds_direction;   /* driver: search directions (minus the solves)  */
    double  seconds_step;        /* driver: ratio tests and step update           */
    int32_t n_renew, n_solve;
    int32_t backend_error;   /* first nonzero backend status, 0 if none          */
    int32_t stall_reason;    /* HELP_STALL_*: why status is STALLED               */
    double  newton_ratio;    /* at an inaccurate-direction stall: (r_p + r_d + r_g)
                                of the affine direction / the residual sum it had
                                to fall below 0.1 of (HSDSolver.cpp:469-471)      */
} help_hsd_result;
#define HELP_STALL_NONE        0
#define HELP_STALL_NEGATIVE    1   /* x_j or z_j not positive (HSDSolver.cpp:63-66,77-80) */
#define HELP_STALL_DIRECTION   2   /* inaccurate Newton direction (HSDSolver.cpp:814-818) */
#define HELP_STALL_BACKEND     3   /* a backend call returned an error (backend_error),
                                      e.g. HELP_ERR_NOCONV from PCG at pcg_max_iter */

/* Optional per-iteration log: called after each Newton-parameter update
 * with (iter, pobj, dobj, gap, pinf, dinf). */
typedef void (*help_hsd_log_fn)(void* user, int32_t iter, double pobj,
                                double dobj, double gap, double pinf,
                                double dinf);

/* flags: HELP_HSD_EXACT sums every reduction sequentially in the reference's
 * order (iterates bit-identical to HELP's driver with the same backend);
 * without it long sums use 64 fixed chunks in parallel (deterministic,
 * thread-count independent).  Calls are not reentrant. */
#define HELP_HSD_EXACT 1
int help_hsd_solve(const help_lp* lp, const help_backend* be,
                   int32_t max_iter, int32_t flags, help_hsd_log_fn log,
                   void* log_user, help_hsd_result* res);

/* The same HSD algorithm with every vector resident on the GPU (D update,
 * residuals, directions, ratio test and step update in FP64 kernels; the
 * host keeps the scalars).  be must come from help_gpu_backend_new; its
 * renew/solve are called with device pointers.  Reductions are
 * deterministic tree sums, so iterates agree with help_hsd_solve to
 * rounding.  No counterpart in the reference, whose driver is host code. */
int help_hsd_solve_gpu(const help_lp* lp, const help_backend* be, int32_t max_iter,
                       help_hsd_log_fn log, void* log_user, help_hsd_result* res);

/* The MI355X backend as a help_backend (ctx owned by the returned table;
 * release with help_gpu_backend_free). */
int  help_gpu_backend_new(const help_opts* opts, help_backend* out);
void help_gpu_backend_free(help_backend* be);
/* Last help_gpu handle created by a help_gpu_backend (NULL before construct). */
help_gpu* help_gpu_backend_handle(const help_backend* be);
/* Kernel-class totals over every handle this backend has destructed (with
 * opts.profile = 1), plus its live handle. */
int help_gpu_backend_kernel_stats(const help_backend* be, int32_t kind, double* ms,
                                  int64_t* launches, double* work);
/* The same with fp32_only != 0 restricted to the launches that ran on the
 * FP32 factor (factorization and triangular solves of the FP32 mode). */
int help_gpu_backend_kernel_stats_prec(const help_backend* be, int32_t kind, int32_t fp32_only,
                                       double* ms, int64_t* launches, double* work);
/* Solve accounting over the backend's lifetime: solves, FP64 refinement
 * (factor-preconditioned CG) iterations, and the largest final relative
 * residual ||r - (A D^-1 A^T + beta I) x2|| / ||r||. */
int help_gpu_backend_solve_stats(const help_backend* be, int64_t* n_solve, int64_t* refine_steps,
                                 double* max_rel_residual);
/* help_gpu_solve2 on the backend's live handle, counted as two solves in
 * help_gpu_backend_solve_stats. */
int help_gpu_backend_solve2(help_backend* be, const double* d,
                            const double* b1a, const double* b2a, double* x1a, double* x2a,
                            const double* b1b, const double* b2b, double* x1b, double* x2b);
/* PCG iterations summed over the backend's solves (PCG modes) */
int64_t help_gpu_backend_pcg_iterations(const help_backend* be);
/* HELP_MODE_PCG_LEARNED: the network every handle of this backend starts from */
int help_gpu_backend_set_warmstart(help_backend* be, help_mpnn* net);

#ifdef __cplusplus
}
#endif
#endif /* HELP_MI355X_H */
