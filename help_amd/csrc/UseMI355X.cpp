// UseMI355X.cpp — HELP linear-solver backend over the MI355X C ABI.
//
// Drop-in beside UseCholmod.cpp / UsePardiso.cpp / UsePCG_CPU.cpp: it defines
// the four functions HELP's driver calls (LP.h:157-160) and is compiled inside
// a HELP source tree (it includes HELP's LP.h for the global model arrays,
// LP.h:99-109) and linked against libhelp_mi355x.so.
//
// Mode selection: environment variable HELP_MI355X_MODE = fp64 (default),
// fp32 (FP32 factor + FP64 refinement), pcg (matrix-free PCG with the
// partial-Cholesky preconditioner) or pcg_learned (PCG started from the
// warm-start network whose FP32 weights, help_mpnn_num_params() floats in
// the library's parameter order, are read from the raw file named by
// HELP_MI355X_NET; HELP_MI355X_NET_BF16=0 runs the network in FP32).
// HELP_MI355X_ORDER = amd (default, CHOLMOD's setting) or nd (nested
// dissection) picks the fill-reducing ordering of the direct modes.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "LP.h"
#include "help_mi355x.h"

static help_gpu* g_help_handle = nullptr;
static help_mpnn* g_help_net = nullptr;
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
    else if (mode && !std::strcmp(mode, "pcg")) o.mode = HELP_MODE_PCG;
    else if (mode && !std::strcmp(mode, "pcg_learned")) o.mode = HELP_MODE_PCG_LEARNED;
    else if (mode && std::strcmp(mode, "fp64")) CheckError(1, (char*)"HELP_MI355X_MODE: fp64, fp32, pcg or pcg_learned");
    const char* order = std::getenv("HELP_MI355X_ORDER");
    if (order && !std::strcmp(order, "nd")) o.ordering = HELP_ORDER_ND;
    else if (order && std::strcmp(order, "amd")) CheckError(1, (char*)"HELP_MI355X_ORDER: amd or nd");
    help_check(help_gpu_create(n_Row, n_Col, (int64_t)rowind.size(), colptr.data(), rowind.data(),
                               val.data(), &o, &g_help_handle),
               "construct");
    if (o.mode == HELP_MODE_PCG_LEARNED) {
        const char* path = std::getenv("HELP_MI355X_NET");
        std::vector<float> w(help_mpnn_num_params());
        FILE* f = path ? std::fopen(path, "rb") : nullptr;
        const bool ok = f && std::fread(w.data(), sizeof(float), w.size(), f) == w.size();
        if (f) std::fclose(f);
        if (!ok) CheckError(1, (char*)"HELP_MI355X_NET: cannot read the network weights");
        const char* bf = std::getenv("HELP_MI355X_NET_BF16");
        help_check(help_mpnn_new(!(bf && !std::strcmp(bf, "0")), 0, &g_help_net), "network");
        help_check(help_mpnn_set_params(g_help_net, w.data()), "network weights");
        help_check(help_gpu_set_warmstart(g_help_handle, g_help_net), "warm start");
    }
    return 0;
}

// Safe at any time and idempotent: CheckError may call it before Construct.
int LinearEquation_Destruct() {
    help_gpu_destroy(g_help_handle);
    g_help_handle = nullptr;
    help_mpnn_free(g_help_net);
    g_help_net = nullptr;
    return 0;
}

void RenewLinearEquation(double* d) {
    help_check(help_gpu_renew(g_help_handle, d, nullptr), "renew");
}

int SolveLinearEquation(double* d, double* b_1, double* b_2, double* x_1, double* x_2) {
    help_check(help_gpu_solve(g_help_handle, d, b_1, b_2, x_1, x_2, nullptr), "solve");
    return 0;
}
