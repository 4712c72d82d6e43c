// HELP backend (LP.h:157-160) over the CPU FP64 oracle (TEST INFRASTRUCTURE
// ONLY): CHOLMOD semantics (UseCholmod.cpp) with the oracle's left-looking
// factor.  The fill-reducing permutation is read from the binary int32 file
// given as argv[2] of the harness (natural order when absent).
#include <cstdio>
#include <vector>

#include "LP.h"
#include "help_mi355x.h"
#include "oracle.h"

const char* g_ref_perm_path = nullptr;
static or_backend* g_be = nullptr;
static help_lp g_lp;
static std::vector<int64_t> g_colptr;
static std::vector<int32_t> g_rowind;
static std::vector<double> g_val;

int LinearEquation_Construct() {
    g_colptr.assign(n_Col + 1, 0);
    g_rowind.clear();
    g_val.clear();
    for (int j = 0; j < n_Col; ++j) {
        for (int p = V_Matrix_Col_Head[j]; p != -1; p = V_Matrix_Col_Next[p]) {
            g_rowind.push_back(V_Matrix_Row[p]);
            g_val.push_back(V_Matrix_Value[p]);
        }
        g_colptr[j + 1] = (int64_t)g_rowind.size();
    }
    std::vector<int32_t> perm;
    if (g_ref_perm_path) {
        FILE* f = std::fopen(g_ref_perm_path, "rb");
        if (f) {
            perm.resize(n_Row);
            if (std::fread(perm.data(), sizeof(int32_t), n_Row, f) != (size_t)n_Row) perm.clear();
            std::fclose(f);
        }
    }
    g_be = or_backend_new(Cholesky_Diagonal_Add, perm.empty() ? nullptr : perm.data(), (int32_t)perm.size());
    g_lp.m = n_Row;
    g_lp.n = n_Col;
    g_lp.nnz = (int64_t)g_rowind.size();
    g_lp.colptr = g_colptr.data();
    g_lp.rowind = g_rowind.data();
    g_lp.val = g_val.data();
    return or_be_construct(g_be, &g_lp);
}

int LinearEquation_Destruct() {
    or_backend_free(g_be);
    g_be = nullptr;
    return 0;
}

void RenewLinearEquation(double* d) { or_be_renew(g_be, d); }

int SolveLinearEquation(double* d, double* b_1, double* b_2, double* x_1, double* x_2) {
    return or_be_solve(g_be, d, b_1, b_2, x_1, x_2);
}
