/*
 * oracle.c — CPU FP64 restatement of HELP's normal-equation path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the MI355X backend; see
 * oracle.h).  Compiled with gcc -O2 -ffp-contract=off so that every
 * multiply and add rounds separately, as in HELP's own build.
 */
#include "oracle.h"
#include "../include/help_mi355x.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

void or_csc_to_csr(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                   int64_t* rowptr, int32_t* colind, int64_t* cscpos) {
    int64_t nnz = colptr[n];
    memset(rowptr, 0, sizeof(int64_t) * (m + 1));
    for (int64_t p = 0; p < nnz; ++p) rowptr[rowind[p] + 1]++;
    for (int32_t i = 0; i < m; ++i) rowptr[i + 1] += rowptr[i];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (m > 0 ? m : 1));
    memcpy(fill, rowptr, sizeof(int64_t) * m);
    for (int32_t j = 0; j < n; ++j)
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            int64_t t = fill[rowind[p]]++;
            colind[t] = j;
            cscpos[t] = p;
        }
    free(fill);
}

/* Helper.cpp:273-298: first pass of ADAt_Allocate (linker form), counts. */
int64_t or_adat_count(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                      const int64_t* rowptr, const int32_t* colind, const int64_t* cscpos,
                      int64_t* adat_rowptr) {
    (void)n;
    int32_t* flag = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    for (int32_t i = 0; i < m; ++i) flag[i] = -1;
    adat_rowptr[0] = 0;
    for (int32_t i = 0; i < m; ++i) {
        int64_t cnt = 0;
        for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            int32_t j = colind[p];
            for (int64_t q = cscpos[p]; q < colptr[j + 1]; ++q) {
                int32_t k = rowind[q];
                if (flag[k] != i) { flag[k] = i; cnt++; }
            }
        }
        adat_rowptr[i + 1] = adat_rowptr[i] + cnt;
    }
    free(flag);
    return adat_rowptr[m];
}

/* Helper.cpp:306-328: second pass, fills the columns in discovery order. */
void or_adat_fill(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                  const int64_t* rowptr, const int32_t* colind, const int64_t* cscpos,
                  const int64_t* adat_rowptr, int32_t* adat_colind) {
    (void)n;
    int32_t* flag = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    for (int32_t i = 0; i < m; ++i) flag[i] = -1;
    for (int32_t i = 0; i < m; ++i) {
        int64_t s = adat_rowptr[i];
        for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            int32_t j = colind[p];
            for (int64_t q = cscpos[p]; q < colptr[j + 1]; ++q) {
                int32_t k = rowind[q];
                if (flag[k] != i) { flag[k] = i; adat_colind[s++] = k; }
            }
        }
    }
    free(flag);
}

/* Helper.cpp:350-377 (ADAt_Calc, linker form) with dinv_j = 1/d_j computed as
 * in UsePardiso.cpp:172-173.  p_Value[k] += (A_ij * dinv_j) * A_kj, over the
 * row list of i (ascending j) and column j from row i downwards.  The 4-way
 * unrolling of the reference does not change the per-k accumulation order. */
void or_adat_values(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                    const double* val, const int64_t* rowptr, const int32_t* colind,
                    const int64_t* cscpos, const double* d, const int64_t* adat_rowptr,
                    const int32_t* adat_colind, double* adat_val) {
    double* dinv = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    double* acc = (double*)malloc(sizeof(double) * (m > 0 ? m : 1));
    for (int32_t j = 0; j < n; ++j) dinv[j] = 1.0 / d[j];
    for (int32_t i = 0; i < m; ++i) {
        for (int64_t s = adat_rowptr[i]; s < adat_rowptr[i + 1]; ++s) acc[adat_colind[s]] = 0;
        for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            int32_t j = colind[p];
            double vp_dj = val[cscpos[p]] * dinv[j];
            for (int64_t q = cscpos[p]; q < colptr[j + 1]; ++q) acc[rowind[q]] += vp_dj * val[q];
        }
        for (int64_t s = adat_rowptr[i]; s < adat_rowptr[i + 1]; ++s) adat_val[s] = acc[adat_colind[s]];
    }
    free(dinv);
    free(acc);
}

/* Helper.cpp:57-68 */
void or_csrmv_n(int32_t nrow, double alpha, const double* val, const int32_t* colind,
                const int64_t* rowptr, const double* x, double beta, double* y) {
    for (int32_t i = 0; i < nrow; ++i) y[i] *= beta;
    for (int32_t i = 0; i < nrow; ++i) {
        double tmp = 0;
        for (int64_t j = rowptr[i]; j < rowptr[i + 1]; ++j) tmp += val[j] * x[colind[j]];
        y[i] += tmp * alpha;
    }
}

/* Helper.cpp:71-81 */
void or_csrmv_t(int32_t nrow, int32_t ncol, double alpha, const double* val,
                const int32_t* colind, const int64_t* rowptr, const double* x, double beta,
                double* y) {
    for (int32_t i = 0; i < ncol; ++i) y[i] *= beta;
    for (int32_t i = 0; i < nrow; ++i) {
        double tmp = alpha * x[i];
        for (int64_t j = rowptr[i]; j < rowptr[i + 1]; ++j) y[colind[j]] += val[j] * tmp;
    }
}

/* Helper.cpp:110-132 */
void or_a_times(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                const double* val, int32_t sign, const double* x, double* y) {
    (void)m;
    for (int32_t j = 0; j < n; ++j) {
        double xj = x[j];
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            if (sign > 0) y[rowind[p]] += val[p] * xj;
            else y[rowind[p]] -= val[p] * xj;
        }
    }
}

/* Helper.cpp:133-155 */
void or_at_times(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                 const double* val, int32_t sign, const double* x, double* y) {
    (void)m;
    for (int32_t j = 0; j < n; ++j) {
        double yj = y[j];
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            if (sign > 0) yj += val[p] * x[rowind[p]];
            else yj -= val[p] * x[rowind[p]];
        }
        y[j] = yj;
    }
}

/* Lower pattern of P M P^T by columns: for new column j, rows i > j. */
static void permuted_lower(int32_t m, const int64_t* arp, const int32_t* aci, const int32_t* perm,
                           int32_t* iperm, int64_t** cp_out, int32_t** ri_out, int64_t** src_out) {
    for (int32_t k = 0; k < m; ++k) iperm[perm[k]] = k;
    int64_t* cp = (int64_t*)calloc(m + 1, sizeof(int64_t));
    for (int32_t i = 0; i < m; ++i)
        for (int64_t s = arp[i]; s < arp[i + 1]; ++s) {
            int32_t a = iperm[i], b = iperm[aci[s]];
            int32_t c = a < b ? a : b;
            cp[c + 1]++;
        }
    for (int32_t j = 0; j < m; ++j) cp[j + 1] += cp[j];
    int32_t* ri = (int32_t*)malloc(sizeof(int32_t) * (cp[m] > 0 ? cp[m] : 1));
    int64_t* src = (int64_t*)malloc(sizeof(int64_t) * (cp[m] > 0 ? cp[m] : 1));
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (m > 0 ? m : 1));
    memcpy(fill, cp, sizeof(int64_t) * m);
    for (int32_t i = 0; i < m; ++i)
        for (int64_t s = arp[i]; s < arp[i + 1]; ++s) {
            int32_t a = iperm[i], b = iperm[aci[s]];
            int32_t c = a < b ? a : b, r = a < b ? b : a;
            int64_t t = fill[c]++;
            ri[t] = r;
            src[t] = s;
        }
    free(fill);
    *cp_out = cp;
    *ri_out = ri;
    *src_out = src;
}

/* Symbolic factorization by structure merging: struct(L_j) = rows of
 * P M P^T below j in column j, united with struct(L_c) \ {c} of every
 * etree child c; parent(j) = min struct(L_j) \ {j}. */
int64_t or_symbolic(int32_t m, const int64_t* adat_rowptr, const int32_t* adat_colind,
                    const int32_t* perm, int32_t* parent, int32_t* colcount, int64_t* Lp,
                    int32_t* Li) {
    int32_t* iperm = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    int64_t *cp, *src;
    int32_t* ri;
    permuted_lower(m, adat_rowptr, adat_colind, perm, iperm, &cp, &ri, &src);
    int32_t** cols = (int32_t**)calloc(m, sizeof(int32_t*));
    int32_t* len = (int32_t*)calloc(m, sizeof(int32_t));
    int32_t* head = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    int32_t* next = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    int32_t* mark = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    for (int32_t j = 0; j < m; ++j) { head[j] = -1; mark[j] = -1; }
    int64_t total = 0;
    for (int32_t j = 0; j < m; ++j) {
        int32_t cnt = 0;
        buf[cnt++] = j;
        mark[j] = j;
        for (int64_t t = cp[j]; t < cp[j + 1]; ++t) {
            int32_t r = ri[t];
            if (r != j && mark[r] != j) { mark[r] = j; buf[cnt++] = r; }
        }
        for (int32_t c = head[j]; c != -1; c = next[c]) {
            for (int32_t u = 0; u < len[c]; ++u) {
                int32_t r = cols[c][u];
                if (r > j && mark[r] != j) { mark[r] = j; buf[cnt++] = r; }
            }
            free(cols[c]);
            cols[c] = NULL;
        }
        /* insertion-free sort (counting via qsort) */
        int cmp_int(const void* a, const void* b);
        qsort(buf, cnt, sizeof(int32_t), cmp_int);
        cols[j] = (int32_t*)malloc(sizeof(int32_t) * cnt);
        memcpy(cols[j], buf, sizeof(int32_t) * cnt);
        len[j] = cnt;
        colcount[j] = cnt;
        parent[j] = cnt > 1 ? buf[1] : -1;
        if (parent[j] >= 0) { next[j] = head[parent[j]]; head[parent[j]] = j; }
        if (Lp) {
            Lp[j] = total;
            if (Li) memcpy(Li + total, buf, sizeof(int32_t) * cnt);
        }
        total += cnt;
    }
    if (Lp) Lp[m] = total;
    for (int32_t j = 0; j < m; ++j) free(cols[j]);
    free(cols); free(len); free(head); free(next); free(mark); free(buf);
    free(iperm); free(cp); free(ri); free(src);
    return total;
}

int cmp_int(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* Left-looking column Cholesky (George & Liu) of P (M + beta I) P^T. */
int or_cholesky(int32_t m, const int64_t* adat_rowptr, const int32_t* adat_colind,
                const double* adat_val, double beta, const int32_t* perm, const int64_t* Lp,
                const int32_t* Li, double* Lx) {
    int32_t* iperm = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    int64_t *cp, *src;
    int32_t* ri;
    permuted_lower(m, adat_rowptr, adat_colind, perm, iperm, &cp, &ri, &src);
    double* x = (double*)calloc(m > 0 ? m : 1, sizeof(double));
    int32_t* link = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));  /* row lists */
    int64_t* first = (int64_t*)malloc(sizeof(int64_t) * (m > 0 ? m : 1)); /* next row position */
    for (int32_t j = 0; j < m; ++j) link[j] = -1;
    int status = 0;
    for (int32_t j = 0; j < m; ++j) {
        for (int64_t t = cp[j]; t < cp[j + 1]; ++t) x[ri[t]] += adat_val[src[t]];
        x[j] += beta;
        int32_t k = link[j];
        while (k != -1) {
            int32_t knext = link[k];
            int64_t pk = first[k];            /* position of L(j,k) in column k */
            double ljk = Lx[pk];
            for (int64_t q = pk; q < Lp[k + 1]; ++q) x[Li[q]] -= Lx[q] * ljk;
            first[k] = pk + 1;
            if (pk + 1 < Lp[k + 1]) {
                int32_t r = Li[pk + 1];
                link[k] = link[r];
                link[r] = k;
            }
            k = knext;
        }
        double dj = x[j];
        if (!(dj > 0.0)) { status = j + 1; dj = 1e300; }
        double ljj = sqrt(dj);
        Lx[Lp[j]] = ljj;
        x[j] = 0.0;
        for (int64_t q = Lp[j] + 1; q < Lp[j + 1]; ++q) {
            Lx[q] = x[Li[q]] / ljj;
            x[Li[q]] = 0.0;
        }
        first[j] = Lp[j] + 1;
        if (Lp[j] + 1 < Lp[j + 1]) {
            int32_t r = Li[Lp[j] + 1];
            link[j] = link[r];
            link[r] = j;
        }
    }
    free(x); free(link); free(first); free(iperm); free(cp); free(ri); free(src);
    return status;
}

void or_chol_solve(int32_t m, const int32_t* perm, const int64_t* Lp, const int32_t* Li,
                   const double* Lx, const double* b, double* x, double* w) {
    for (int32_t k = 0; k < m; ++k) w[k] = b[perm[k]];          /* UseCholmod.cpp:180-182 */
    for (int32_t j = 0; j < m; ++j) {                            /* L y = P b */
        double yj = w[j] / Lx[Lp[j]];
        w[j] = yj;
        for (int64_t q = Lp[j] + 1; q < Lp[j + 1]; ++q) w[Li[q]] -= Lx[q] * yj;
    }
    for (int32_t j = m - 1; j >= 0; --j) {                       /* L^T z = y */
        double s = w[j];
        for (int64_t q = Lp[j] + 1; q < Lp[j + 1]; ++q) s -= Lx[q] * w[Li[q]];
        w[j] = s / Lx[Lp[j]];
    }
    for (int32_t k = 0; k < m; ++k) x[perm[k]] = w[k];          /* UseCholmod.cpp:216-217 */
}

/* UseCholmod.cpp:157-238:  r = A (b1 ./ d) - b2;  x2 = (ADA^T + beta I)^-1 r;
 * x1 = (b1 - A^T x2) ./ d. */
void or_solve_kkt(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind,
                  const double* val, const int32_t* perm, const int64_t* Lp, const int32_t* Li,
                  const double* Lx, const double* d, const double* b1, const double* b2,
                  double* x1, double* x2, double* work) {
    double* r = work;           /* m */
    double* t = work + m;       /* n */
    for (int32_t i = 0; i < m; ++i) r[i] = -b2[i];               /* :173-174 */
    for (int32_t j = 0; j < n; ++j) t[j] = b1[j] / d[j];         /* :176-177 */
    or_a_times(m, n, colptr, rowind, val, 1, t, r);              /* :179 */
    double* w = (double*)malloc(sizeof(double) * (m > 0 ? m : 1));
    or_chol_solve(m, perm, Lp, Li, Lx, r, x2, w);                /* :180-217 */
    free(w);
    for (int32_t j = 0; j < n; ++j) t[j] = b1[j];                /* :226-227 */
    or_at_times(m, n, colptr, rowind, val, -1, x2, t);           /* :228 */
    for (int32_t j = 0; j < n; ++j) x1[j] = t[j] / d[j];         /* :230-231 */
}

/* ---------------- partial Cholesky preconditioner, PCG.cpp:38-177 -------- */
int or_pchol_build(int32_t m, int32_t n, const int64_t* rowptr, const int32_t* colind,
                   const double* csrval, const double* d, double gamma, double delta, int32_t K,
                   or_pchol* pc) {
    if (K > m) K = m;
    pc->m = m;
    pc->k = K;
    pc->perm = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    pc->L = (double*)calloc((size_t)m * (K > 0 ? K : 1), sizeof(double));
    pc->D = (double*)calloc(K > 0 ? K : 1, sizeof(double));
    pc->diag_s = (double*)malloc(sizeof(double) * (m > 0 ? m : 1));
    pc->l11inv = (double*)calloc((size_t)(K > 0 ? K : 1) * (K > 0 ? K : 1), sizeof(double));
    double* U = (double*)calloc((size_t)m * (K > 0 ? K : 1), sizeof(double));
    double* gr = (double*)calloc(m > 0 ? m : 1, sizeof(double));
    double* sc = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    double* S = pc->diag_s;
    int32_t* P = pc->perm;
    /* Step 1: diagonal (PCG.cpp:40-46) */
    for (int32_t i = 0; i < m; ++i) {
        S[i] = delta;
        for (int64_t j = rowptr[i]; j < rowptr[i + 1]; ++j)
            S[i] += csrval[j] * csrval[j] / (d[colind[j]] + gamma);
    }
    for (int32_t i = 0; i < m; ++i) P[i] = i;
    /* Step 2: greedy partial LDL^T (PCG.cpp:51-103) */
    for (int32_t k = 0; k < K; ++k) {
        int32_t cm = k;
        for (int32_t i = k + 1; i < m; ++i)
            if (S[P[i]] > S[P[cm]]) cm = i;
        int32_t tmp = P[k]; P[k] = P[cm]; P[cm] = tmp;
        int32_t col = P[k];
        for (int64_t j = rowptr[col]; j < rowptr[col + 1]; ++j)
            sc[colind[j]] = csrval[j] / (d[colind[j]] + gamma);
        for (int32_t i = k; i < m; ++i) {
            int32_t ii = P[i];
            gr[ii] = 0;
            for (int64_t j = rowptr[ii]; j < rowptr[ii + 1]; ++j) gr[ii] += csrval[j] * sc[colind[j]];
        }
        gr[col] += delta;
        for (int64_t j = rowptr[col]; j < rowptr[col + 1]; ++j) sc[colind[j]] = 0;
        pc->D[k] = gr[col];
        for (int32_t j = k + 1; j < m; ++j) U[(size_t)P[j] * K + k] = gr[P[j]];
        for (int32_t r = 0; r < k; ++r) {
            double t = pc->L[(size_t)col * K + r];
            pc->D[k] -= U[(size_t)col * K + r] * t;
            for (int32_t j = k + 1; j < m; ++j) U[(size_t)P[j] * K + k] -= U[(size_t)P[j] * K + r] * t;
        }
        pc->L[(size_t)col * K + k] = 1;
        for (int32_t j = k + 1; j < m; ++j) pc->L[(size_t)P[j] * K + k] = U[(size_t)P[j] * K + k] / pc->D[k];
        double td = pc->D[k];
        for (int32_t j = k + 1; j < m; ++j) {
            double l = pc->L[(size_t)P[j] * K + k];
            S[P[j]] -= td * l * l;
        }
    }
    /* Step 4.1: L11^-1 by Gaussian elimination (PCG.cpp:119-133) */
    for (int32_t i = 0; i < K; ++i)
        for (int32_t j = 0; j < K; ++j) pc->l11inv[(size_t)i * K + j] = (double)(i == j);
    for (int32_t j = 0; j < K; ++j)
        for (int32_t i = j + 1; i < K; ++i) {
            double f = pc->L[(size_t)P[i] * K + j];
            for (int32_t k = 0; k <= j; ++k) pc->l11inv[(size_t)i * K + k] -= pc->l11inv[(size_t)j * K + k] * f;
        }
    free(U); free(gr); free(sc);
    return 0;
}

void or_pchol_free(or_pchol* pc) {
    free(pc->perm); free(pc->L); free(pc->D); free(pc->diag_s); free(pc->l11inv);
    memset(pc, 0, sizeof(*pc));
}

/* PCG.cpp:183-196: ret = Inv2^T Inv1^T D^-1 Inv1 Inv2 r, with
 * Inv2 = diag(L11^-1, I) and Inv1 = [I 0; -L21 I] (PCG.cpp:135-176), every
 * CSR product summed along its row in column order. */
void or_pchol_apply(const or_pchol* pc, const double* r, double* ret, double* work) {
    int32_t m = pc->m, K = pc->k;
    const int32_t* P = pc->perm;
    double* t1 = work;
    double* t2 = work + m;
    for (int32_t i = 0; i < K; ++i) {                 /* t1 = Inv2 r */
        double s = 0;
        for (int32_t j = 0; j <= i; ++j) s += pc->l11inv[(size_t)i * K + j] * r[j];
        t1[i] = s;
    }
    for (int32_t i = K; i < m; ++i) t1[i] = r[i];
    for (int32_t i = 0; i < K; ++i) t2[i] = t1[i];   /* t2 = Inv1 t1 */
    for (int32_t i = K; i < m; ++i) {
        double s = 0;
        for (int32_t j = 0; j < K; ++j) s += -pc->L[(size_t)P[i] * K + j] * t1[j];
        s += t1[i];
        t2[i] = s;
    }
    for (int32_t i = 0; i < m; ++i) t2[i] /= pc->diag_s[P[i]];   /* PCG.cpp:192-193 */
    /* t1 = Inv1^T t2 (transposed CSR product: scatter rows in order) */
    for (int32_t i = 0; i < m; ++i) t1[i] = 0;
    for (int32_t i = 0; i < K; ++i) t1[i] += t2[i];
    for (int32_t i = K; i < m; ++i) {
        double ti = t2[i];
        for (int32_t j = 0; j < K; ++j) t1[j] += -pc->L[(size_t)P[i] * K + j] * ti;
        t1[i] += ti;
    }
    for (int32_t i = 0; i < m; ++i) ret[i] = 0;     /* ret = Inv2^T t1 */
    for (int32_t i = 0; i < K; ++i) {
        double ti = t1[i];
        for (int32_t j = 0; j <= i; ++j) ret[j] += pc->l11inv[(size_t)i * K + j] * ti;
    }
    for (int32_t i = K; i < m; ++i) ret[i] += t1[i];
}

static double nrm2(int32_t n, const double* x) {
    double s = 0;
    for (int32_t i = 0; i < n; ++i) s += x[i] * x[i];
    return sqrt(s);
}
static double dot(int32_t n, const double* x, const double* y) {
    double s = 0;
    for (int32_t i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}

/* PCG.cpp:296-395 on G = A (D + gamma I)^-1 A^T + delta I, with the rows of A
 * visited in PCG_PERM order (PCG.cpp:105-110, 323-324, 391-394). */
int or_pcg(int32_t m, int32_t n, const int64_t* rowptr, const int32_t* colind,
           const double* csrval, const double* d, double gamma, double delta, const or_pchol* pc,
           const double* b, const double* x0, double* xout, double tol, int32_t max_iter,
           double* final_rel) {
    const int32_t* P = pc->perm;
    double* x = (double*)malloc(sizeof(double) * m);
    double* r = (double*)malloc(sizeof(double) * m);
    double* z = (double*)malloc(sizeof(double) * m);
    double* p = (double*)malloc(sizeof(double) * m);
    double* q = (double*)malloc(sizeof(double) * m);
    double* tc = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    double* w = (double*)malloc(sizeof(double) * 2 * m);
    for (int32_t i = 0; i < m; ++i) x[i] = x0 ? x0[P[i]] : 0.0;   /* permuted x0 */
    /* r0 = (b - delta x) - A (D+gamma)^-1 A^T x  (PCG.cpp:312-327) */
    for (int32_t j = 0; j < n; ++j) tc[j] = 0;
    for (int32_t i = 0; i < m; ++i) {
        int32_t ii = P[i];
        double t = x[i];
        for (int64_t k = rowptr[ii]; k < rowptr[ii + 1]; ++k) tc[colind[k]] += csrval[k] * t;
    }
    for (int32_t j = 0; j < n; ++j) tc[j] /= d[j] + gamma;
    for (int32_t i = 0; i < m; ++i) r[i] = b[P[i]] - delta * x[i];
    for (int32_t i = 0; i < m; ++i) {
        int32_t ii = P[i];
        double t = 0;
        for (int64_t k = rowptr[ii]; k < rowptr[ii + 1]; ++k) t += csrval[k] * tc[colind[k]];
        r[i] += t * -1.0;
    }
    or_pchol_apply(pc, r, z, w);
    memcpy(p, z, sizeof(double) * m);
    double bn = 0;
    for (int32_t i = 0; i < m; ++i) bn += b[i] * b[i];
    bn = sqrt(bn);
    /* stop on ||r|| < tol ||b|| (with x0 = 0 this is the reference's
     * ||r|| < PREC_THRESHOLD ||r0||, PCG.cpp:335-344) */
    double terminate = bn * tol;
    int it = 0;
    double rn = nrm2(m, r);
    for (;;) {
        rn = nrm2(m, r);
        if (rn < terminate || it >= max_iter) break;
        it++;
        for (int32_t j = 0; j < n; ++j) tc[j] = 0;
        for (int32_t i = 0; i < m; ++i) {
            int32_t ii = P[i];
            double t = p[i];
            for (int64_t k = rowptr[ii]; k < rowptr[ii + 1]; ++k) tc[colind[k]] += csrval[k] * t;
        }
        for (int32_t j = 0; j < n; ++j) tc[j] /= d[j] + gamma;
        for (int32_t i = 0; i < m; ++i) q[i] = delta * p[i];
        for (int32_t i = 0; i < m; ++i) {
            int32_t ii = P[i];
            double t = 0;
            for (int64_t k = rowptr[ii]; k < rowptr[ii + 1]; ++k) t += csrval[k] * tc[colind[k]];
            q[i] += t;
        }
        double au = dot(m, z, r), ad = dot(m, p, q);
        double alpha = au / ad;
        for (int32_t i = 0; i < m; ++i) x[i] += alpha * p[i];
        for (int32_t i = 0; i < m; ++i) r[i] += -alpha * q[i];
        or_pchol_apply(pc, r, z, w);
        double beta = dot(m, z, r) / au;
        for (int32_t i = 0; i < m; ++i) p[i] = beta * p[i] + z[i];
    }
    for (int32_t i = 0; i < m; ++i) xout[P[i]] = x[i];
    if (final_rel) *final_rel = bn > 0 ? rn / bn : 0.0;
    free(x); free(r); free(z); free(p); free(q); free(tc); free(w);
    return it;
}

/* ---------------- CPU FP64 backend with CHOLMOD semantics --------------- */
struct or_backend {
    double beta;
    int32_t m, n;
    int64_t *colptr, *rowptr, *cscpos, *arp, *Lp;
    int32_t *rowind, *colind, *aci, *Li, *perm;
    double *val, *aval, *Lx, *work;
    int32_t have_perm;
    double sec_renew, sec_solve;
    int32_t factor_status;
};

or_backend* or_backend_new(double beta, const int32_t* perm, int32_t perm_len) {
    or_backend* b = (or_backend*)calloc(1, sizeof(or_backend));
    b->beta = beta;
    if (perm && perm_len > 0) {
        b->perm = (int32_t*)malloc(sizeof(int32_t) * perm_len);
        memcpy(b->perm, perm, sizeof(int32_t) * perm_len);
        b->have_perm = perm_len;
    }
    return b;
}

int or_be_destruct(void* ctx) {
    or_backend* b = (or_backend*)ctx;
    if (!b) return 0;
    free(b->colptr); free(b->rowptr); free(b->cscpos); free(b->arp); free(b->Lp);
    free(b->rowind); free(b->colind); free(b->aci); free(b->Li);
    free(b->val); free(b->aval); free(b->Lx); free(b->work);
    b->colptr = b->rowptr = b->cscpos = b->arp = b->Lp = NULL;
    b->rowind = b->colind = b->aci = b->Li = NULL;
    b->val = b->aval = b->Lx = b->work = NULL;
    return 0;
}

void or_backend_free(or_backend* b) {
    if (!b) return;
    or_be_destruct(b);
    free(b->perm);
    free(b);
}

int or_be_construct(void* ctx, const void* lpv) {
    or_backend* b = (or_backend*)ctx;
    const help_lp* lp = (const help_lp*)lpv;
    or_be_destruct(b);
    int32_t m = lp->m, n = lp->n;
    int64_t nnz = lp->colptr[n];
    b->m = m; b->n = n;
    b->colptr = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
    b->rowind = (int32_t*)malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
    b->val = (double*)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    memcpy(b->colptr, lp->colptr, sizeof(int64_t) * (n + 1));
    memcpy(b->rowind, lp->rowind, sizeof(int32_t) * nnz);
    memcpy(b->val, lp->val, sizeof(double) * nnz);
    b->rowptr = (int64_t*)malloc(sizeof(int64_t) * (m + 1));
    b->colind = (int32_t*)malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
    b->cscpos = (int64_t*)malloc(sizeof(int64_t) * (nnz > 0 ? nnz : 1));
    or_csc_to_csr(m, n, b->colptr, b->rowind, b->rowptr, b->colind, b->cscpos);
    b->arp = (int64_t*)malloc(sizeof(int64_t) * (m + 1));
    int64_t nup = or_adat_count(m, n, b->colptr, b->rowind, b->rowptr, b->colind, b->cscpos, b->arp);
    b->aci = (int32_t*)malloc(sizeof(int32_t) * (nup > 0 ? nup : 1));
    b->aval = (double*)malloc(sizeof(double) * (nup > 0 ? nup : 1));
    or_adat_fill(m, n, b->colptr, b->rowind, b->rowptr, b->colind, b->cscpos, b->arp, b->aci);
    if (b->have_perm != m) {
        free(b->perm);
        b->perm = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
        for (int32_t i = 0; i < m; ++i) b->perm[i] = i;
        b->have_perm = m;
    }
    int32_t* par = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    int32_t* cc = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    b->Lp = (int64_t*)malloc(sizeof(int64_t) * (m + 1));
    int64_t nl = or_symbolic(m, b->arp, b->aci, b->perm, par, cc, b->Lp, NULL);
    b->Li = (int32_t*)malloc(sizeof(int32_t) * (nl > 0 ? nl : 1));
    or_symbolic(m, b->arp, b->aci, b->perm, par, cc, b->Lp, b->Li);
    b->Lx = (double*)malloc(sizeof(double) * (nl > 0 ? nl : 1));
    b->work = (double*)malloc(sizeof(double) * (m + n + 1));
    free(par); free(cc);
    b->sec_renew = b->sec_solve = 0;
    return 0;
}

int or_be_renew(void* ctx, const double* d) {
    or_backend* b = (or_backend*)ctx;
    double t0 = now_s();
    or_adat_values(b->m, b->n, b->colptr, b->rowind, b->val, b->rowptr, b->colind, b->cscpos, d,
                   b->arp, b->aci, b->aval);
    b->factor_status = or_cholesky(b->m, b->arp, b->aci, b->aval, b->beta, b->perm, b->Lp, b->Li, b->Lx);
    b->sec_renew += now_s() - t0;
    return b->factor_status ? HELP_ERR_NOT_PD : 0;
}

int or_be_solve(void* ctx, const double* d, const double* b1, const double* b2, double* x1,
                double* x2) {
    or_backend* b = (or_backend*)ctx;
    double t0 = now_s();
    or_solve_kkt(b->m, b->n, b->colptr, b->rowind, b->val, b->perm, b->Lp, b->Li, b->Lx, d, b1, b2,
                 x1, x2, b->work);
    b->sec_solve += now_s() - t0;
    return 0;
}

double or_be_seconds_renew(const or_backend* b) { return b->sec_renew; }
double or_be_seconds_solve(const or_backend* b) { return b->sec_solve; }
