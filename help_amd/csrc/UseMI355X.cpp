// UseMI355X.cpp — HELP linear-solver backend over the MI355X C ABI.
//
// Drop-in beside UseCholmod.cpp / UsePardiso.cpp / UsePCG_CPU.cpp: it defines
// the four functions HELP's driver calls (LP.h:157-160) and is compiled inside
// a HELP source tree (it includes HELP's LP.h for the global model arrays,
// LP.h:99-109) and linked against libhelp_mi355x.so.
//
// Mode selection: environment variable HELP_MI355X_MODE = fp64 (default),
// fp32 (FP32 factor + FP64 refinement).
#include <cstdlib>
#include <cstring>
#include <vector>

#include "LP.h"
#include "help_mi355x.h"

static help_gpu* g_help_handle = nullptr;
static char g_help_msg[512];

static void help_check(int status, const char* what) {
    if (!status) return;
    const char* detail = g_help_handle ? help_gpu_last_error(g_help_handle) : help_last_create_error();
    std::snprintf(g_help_msg, sizeof(g_help_msg), "MI355X backend %s failed: %s", what, detail);
    CheckError(status, g_help_msg);   // Helper.cpp:24-35: destructs and exits
}

// UseCholmod.cpp:29-58: copy the column lists (sorted by row after
// Presolve_Init) into CSC once, and run the symbolic analysis.
int LinearEquation_Construct() {
    std::vector<int64_t> colptr(n_Col + 1, 0);
    std::vector<int32_t> rowind;
    std::vector<double> val;
    rowind.reserve(n_Element);
    val.reserve(n_Element);
    for (int j = 0; j < n_Col; ++j) {
        for (int p = V_Matrix_Col_Head[j]; p != -1; p = V_Matrix_Col_Next[p]) {
            rowind.push_back(V_Matrix_Row[p]);
            val.push_back(V_Matrix_Value[p]);
        }
        colptr[j + 1] = (int64_t)rowind.size();
    }
    help_opts o;
    help_opts_default(&o);
    o.beta = Cholesky_Diagonal_Add;   // LP.h:139
    const char* mode = std::getenv("HELP_MI355X_MODE");
    if (mode && !std::strcmp(mode, "fp32")) o.mode = HELP_MODE_DIRECT_FP32;
    help_check(help_gpu_create(n_Row, n_Col, (int64_t)rowind.size(), colptr.data(), rowind.data(),
                               val.data(), &o, &g_help_handle),
               "construct");
    return 0;
}

// Safe at any time and idempotent: CheckError may call it before Construct.
int LinearEquation_Destruct() {
    help_gpu_destroy(g_help_handle);
    g_help_handle = nullptr;
    return 0;
}

void RenewLinearEquation(double* d) {
    help_check(help_gpu_renew(g_help_handle, d, nullptr), "renew");
}

int SolveLinearEquation(double* d, double* b_1, double* b_2, double* x_1, double* x_2) {
    help_check(help_gpu_solve(g_help_handle, d, b_1, b_2, x_1, x_2, nullptr), "solve");
    return 0;
}
