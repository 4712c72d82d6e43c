/*
 * CPU baseline backend (TEST INFRASTRUCTURE ONLY): HELP's PARDISO backend
 * restated over Intel MKL PARDISO, loaded at run time from libmkl_rt.
 *
 * Follows UsePardiso.cpp (HELP), which needs mkl.h and so cannot be compiled
 * in this image; the PARDISO entry point is declared here from MKL's
 * documented prototype and resolved with dlopen/dlsym:
 *   construct  UsePardiso.cpp:40-150  CSR of A (= CSC of A^T), the upper
 *              pattern of A A^T (Helper.cpp:273-329), phase 11 with the
 *              iparm table of :99-130 (mtype -2, METIS nested dissection
 *              iparm[1]=3, perturbation 1e-13, scaling, zero-based)
 *   renew      UsePardiso.cpp:167-215  dinv = 1/d, A D^-1 A^T values
 *              (Helper.cpp:350-377: per row i, ascending j, the products
 *              A_ij dinv_j A_kj accumulated densely, OpenMP over rows),
 *              phase 22; no beta shift (the PARDISO backend has none)
 *   solve      UsePardiso.cpp:223-283  r = A (b1/d) - b2, phase 33,
 *              x1 = (b1 - A^T x2)/d
 *   destruct   UsePardiso.cpp:152-165  phase -1
 * One deliberate difference: each row of the upper pattern is sorted by
 * column, as PARDISO's CSR input requires; HELP passes its discovery order
 * with the matrix checker off (UsePardiso.cpp:115).
 *
 * Used only as bench.py's CPU baseline (all host cores).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "help_mi355x.h"

typedef void (*pardiso_fn)(void** pt, const int* maxfct, const int* mnum, const int* mtype,
                           const int* phase, const int* n, const void* a, const int* ia, const int* ja,
                           int* perm, const int* nrhs, int* iparm, const int* msglvl, void* b, void* x,
                           int* error);
typedef void (*set_threads_fn)(int);

static pardiso_fn g_pardiso = NULL;
static char g_err[256] = "";

static double wall(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

typedef struct pb_backend {
    int threads;
    int m, n;
    /* A as CSR (rows of A = the CSC of A^T that UsePardiso.cpp builds) */
    int64_t* rp; int32_t* ci; double* rv;
    /* A as CSC (copy of the model) */
    int64_t* cp; int32_t* ri; double* cv;
    /* upper A D^-1 A^T: PARDISO CSR, rows sorted */
    int* ia; int* ja; double* a;
    int64_t nup;
    double *dinv, *tcol, *trow, *x2;
    void* pt[64];
    int iparm[64];
    int* perm;
    int analyzed;
    double sec_construct, sec_renew, sec_solve;
    int64_t n_renew, n_solve;
} pb_backend;

/* Loads libmkl_rt once; 0 on success. */
static int pb_load(void) {
    if (g_pardiso) return 0;
    const char* names[] = {"libmkl_rt.so", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so",
                           "/opt/conda/lib/libmkl_rt.so.1"};
    void* h = NULL;
    for (int i = 0; i < 4 && !h; ++i) h = dlopen(names[i], RTLD_NOW | RTLD_GLOBAL);
    if (!h) {
        snprintf(g_err, sizeof g_err, "libmkl_rt not found: %s", dlerror());
        return 1;
    }
    g_pardiso = (pardiso_fn)dlsym(h, "pardiso");
    if (!g_pardiso) {
        snprintf(g_err, sizeof g_err, "pardiso symbol not found in libmkl_rt");
        return 1;
    }
    return 0;
}

const char* pb_error(void) { return g_err; }

void* pb_new(int threads) {
    if (pb_load()) return NULL;
    pb_backend* b = (pb_backend*)calloc(1, sizeof(pb_backend));
    b->threads = threads > 0 ? threads : omp_get_max_threads();
    set_threads_fn st = (set_threads_fn)dlsym(RTLD_DEFAULT, "MKL_Set_Num_Threads");
    if (st) st(b->threads);
    return b;
}

static void pb_call(pb_backend* b, int phase, double* rhs, double* x, int* error) {
    int maxfct = 1, mnum = 1, mtype = -2, nrhs = 1, msglvl = 0, n = b->m;
    double dummy = 0;
    g_pardiso(b->pt, &maxfct, &mnum, &mtype, &phase, &n, b->a, b->ia, b->ja, b->perm, &nrhs, b->iparm,
              &msglvl, rhs ? (void*)rhs : (void*)&dummy, x ? (void*)x : (void*)&dummy, error);
}

int pb_destruct(void* ctx) {
    pb_backend* b = (pb_backend*)ctx;
    if (!b) return 0;
    if (b->analyzed) {
        int err = 0;
        pb_call(b, -1, NULL, NULL, &err);
        b->analyzed = 0;
    }
    free(b->rp); free(b->ci); free(b->rv); free(b->cp); free(b->ri); free(b->cv);
    free(b->ia); free(b->ja); free(b->a); free(b->dinv); free(b->tcol); free(b->trow); free(b->x2);
    free(b->perm);
    b->rp = NULL; b->ci = NULL; b->rv = NULL; b->cp = NULL; b->ri = NULL; b->cv = NULL;
    b->ia = NULL; b->ja = NULL; b->a = NULL; b->dinv = NULL; b->tcol = NULL; b->trow = NULL;
    b->x2 = NULL; b->perm = NULL;
    return 0;
}

void pb_free(void* ctx) {
    pb_destruct(ctx);
    free(ctx);
}

static int cmp_i32(const void* x, const void* y) {
    int32_t a = *(const int32_t*)x, c = *(const int32_t*)y;
    return (a > c) - (a < c);
}

int pb_construct(void* ctx, const void* lpv) {
    pb_backend* b = (pb_backend*)ctx;
    const help_lp* lp = (const help_lp*)lpv;
    double t0 = wall();
    pb_destruct(b);
    const int m = lp->m, n = lp->n;
    const int64_t nnz = lp->nnz;
    b->m = m; b->n = n;
    b->cp = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
    b->ri = (int32_t*)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
    b->cv = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
    memcpy(b->cp, lp->colptr, sizeof(int64_t) * (n + 1));
    memcpy(b->ri, lp->rowind, sizeof(int32_t) * nnz);
    memcpy(b->cv, lp->val, sizeof(double) * nnz);
    /* CSR of A: rows ascending by column (the transpose of the sorted CSC) */
    b->rp = (int64_t*)calloc(m + 1, sizeof(int64_t));
    b->ci = (int32_t*)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
    b->rv = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
    for (int64_t p = 0; p < nnz; ++p) b->rp[lp->rowind[p] + 1]++;
    for (int i = 0; i < m; ++i) b->rp[i + 1] += b->rp[i];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (m + 1));
    memcpy(fill, b->rp, sizeof(int64_t) * (m + 1));
    for (int j = 0; j < n; ++j)
        for (int64_t p = lp->colptr[j]; p < lp->colptr[j + 1]; ++p) {
            int64_t q = fill[lp->rowind[p]]++;
            b->ci[q] = j;
            b->rv[q] = lp->val[p];
        }
    free(fill);
    /* upper pattern of A A^T, two passes (count, fill) with per-thread marks */
    int64_t* cnt = (int64_t*)calloc(m + 1, sizeof(int64_t));
#pragma omp parallel num_threads(b->threads)
    {
        int32_t* mark = (int32_t*)malloc(sizeof(int32_t) * m);
        for (int k = 0; k < m; ++k) mark[k] = -1;
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < m; ++i) {
            int64_t c = 0;
            for (int64_t p = b->rp[i]; p < b->rp[i + 1]; ++p) {
                int j = b->ci[p];
                for (int64_t q = b->cp[j]; q < b->cp[j + 1]; ++q) {
                    int k = b->ri[q];
                    if (k >= i && mark[k] != i) { mark[k] = i; ++c; }
                }
            }
            if (b->rp[i] == b->rp[i + 1]) c = 1;   /* empty row: keep a diagonal slot */
            cnt[i + 1] = c;
        }
        free(mark);
    }
    for (int i = 0; i < m; ++i) cnt[i + 1] += cnt[i];
    b->nup = cnt[m];
    if (b->nup > 2147483647LL) {
        snprintf(g_err, sizeof g_err, "ADA^T too large for 32-bit PARDISO indices");
        free(cnt);
        return HELP_ERR_ARG;
    }
    b->ia = (int*)malloc(sizeof(int) * (m + 1));
    b->ja = (int*)malloc(sizeof(int) * b->nup);
    b->a = (double*)malloc(sizeof(double) * b->nup);
    for (int i = 0; i <= m; ++i) b->ia[i] = (int)cnt[i];
#pragma omp parallel num_threads(b->threads)
    {
        int32_t* mark = (int32_t*)malloc(sizeof(int32_t) * m);
        for (int k = 0; k < m; ++k) mark[k] = -1;
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < m; ++i) {
            int64_t c = cnt[i];
            if (b->rp[i] == b->rp[i + 1]) { b->ja[c] = i; continue; }
            for (int64_t p = b->rp[i]; p < b->rp[i + 1]; ++p) {
                int j = b->ci[p];
                for (int64_t q = b->cp[j]; q < b->cp[j + 1]; ++q) {
                    int k = b->ri[q];
                    if (k >= i && mark[k] != i) { mark[k] = i; b->ja[c++] = k; }
                }
            }
            qsort(b->ja + cnt[i], (size_t)(c - cnt[i]), sizeof(int32_t), cmp_i32);
        }
        free(mark);
    }
    free(cnt);
    b->dinv = (double*)malloc(sizeof(double) * n);
    b->tcol = (double*)malloc(sizeof(double) * n);
    b->trow = (double*)malloc(sizeof(double) * m);
    b->x2 = (double*)malloc(sizeof(double) * m);
    b->perm = (int*)malloc(sizeof(int) * m);
    /* iparm as UsePardiso.cpp:99-130 */
    memset(b->pt, 0, sizeof b->pt);
    memset(b->iparm, 0, sizeof b->iparm);
    b->iparm[0] = 1; b->iparm[1] = 3; b->iparm[4] = 2; b->iparm[9] = 13; b->iparm[10] = 1;
    b->iparm[23] = 1; b->iparm[34] = 1;
    for (int64_t p = 0; p < b->nup; ++p) b->a[p] = 1.0;   /* analysis reads the pattern only */
    int err = 0;
    pb_call(b, 11, NULL, NULL, &err);
    if (err) {
        snprintf(g_err, sizeof g_err, "PARDISO phase 11 error %d", err);
        return HELP_ERR_ARG;
    }
    b->analyzed = 1;
    b->iparm[4] = 1;   /* UsePardiso.cpp:148: reuse the permutation */
    b->sec_construct = wall() - t0;
    return HELP_OK;
}

int pb_renew(void* ctx, const double* d) {
    pb_backend* b = (pb_backend*)ctx;
    double t0 = wall();
    const int m = b->m, n = b->n;
    for (int j = 0; j < n; ++j) b->dinv[j] = 1.0 / d[j];
#pragma omp parallel num_threads(b->threads)
    {
        double* acc = (double*)calloc(m, sizeof(double));
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < m; ++i) {
            for (int64_t p = b->rp[i]; p < b->rp[i + 1]; ++p) {
                const int j = b->ci[p];
                const double v = b->rv[p] * b->dinv[j];
                for (int64_t q = b->cp[j]; q < b->cp[j + 1]; ++q) {
                    const int k = b->ri[q];
                    if (k >= i) acc[k] += v * b->cv[q];
                }
            }
            for (int s = b->ia[i]; s < b->ia[i + 1]; ++s) {
                b->a[s] = acc[b->ja[s]];
                acc[b->ja[s]] = 0.0;
            }
        }
        free(acc);
    }
    int err = 0;
    pb_call(b, 22, NULL, NULL, &err);
    b->sec_renew += wall() - t0;
    b->n_renew++;
    if (err) {
        snprintf(g_err, sizeof g_err, "PARDISO phase 22 error %d", err);
        return HELP_ERR_NOT_PD;
    }
    return HELP_OK;
}

int pb_solve(void* ctx, const double* d, const double* b1, const double* b2, double* x1, double* x2) {
    pb_backend* b = (pb_backend*)ctx;
    double t0 = wall();
    const int m = b->m, n = b->n;
#pragma omp parallel for num_threads(b->threads) schedule(static)
    for (int j = 0; j < n; ++j) b->tcol[j] = b1[j] / d[j];
#pragma omp parallel for num_threads(b->threads) schedule(dynamic, 1024)
    for (int i = 0; i < m; ++i) {
        double s = -b2[i];
        for (int64_t p = b->rp[i]; p < b->rp[i + 1]; ++p) s += b->rv[p] * b->tcol[b->ci[p]];
        b->trow[i] = s;
    }
    int err = 0;
    pb_call(b, 33, b->trow, b->x2, &err);
    if (err) {
        snprintf(g_err, sizeof g_err, "PARDISO phase 33 error %d", err);
        return HELP_ERR_NOCONV;
    }
    memcpy(x2, b->x2, sizeof(double) * m);
#pragma omp parallel for num_threads(b->threads) schedule(dynamic, 1024)
    for (int j = 0; j < n; ++j) {
        double s = b1[j];
        for (int64_t p = b->cp[j]; p < b->cp[j + 1]; ++p) s -= b->cv[p] * x2[b->ri[p]];
        x1[j] = s / d[j];
    }
    b->sec_solve += wall() - t0;
    b->n_solve++;
    return HELP_OK;
}

/* seconds spent: 0 construct, 1 renews, 2 solves; 3/4 the renew/solve counts */
double pb_stat(const void* ctx, int which) {
    const pb_backend* b = (const pb_backend*)ctx;
    switch (which) {
        case 0: return b->sec_construct;
        case 1: return b->sec_renew;
        case 2: return b->sec_solve;
        case 3: return (double)b->n_renew;
        case 4: return (double)b->n_solve;
        default: return 0;
    }
}
