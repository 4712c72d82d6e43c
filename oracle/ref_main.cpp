// Harness main for HELP's own driver (TEST INFRASTRUCTURE ONLY).
//
// Runs HELP's pipeline as Main.cpp:37-109 does (Prog_Init, MPS_ReadFile,
// CRUSH_Main, Presolve_Main, HSD_Main, Report_Solution) with the input file
// from argv[1] and no stdin token, without the Matrix.txt debug dump
// (Main.cpp:93), and prints one machine-readable RESULT line.  The backend
// (the four LP.h:157-160 functions) is whichever object it is linked with.
#include <cstdio>
#include <cstring>

#include "LP.h"

extern double HSD_primal_obj, HSD_dual_obj, HSD_relative_gap, HSD_primal_infeas, HSD_dual_infeas;
extern const char* g_ref_perm_path;   // defined by the CPU-oracle adapter (or null)

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s file.mps [perm.bin]\n", argv[0]);
        return 2;
    }
    if (argc > 2) g_ref_perm_path = argv[2];
    CheckError(Prog_Init(), (char*)"Initialization Failed!");
    std::strncpy(Filename, argv[1], MAX_FILENAME - 1);
    CheckError(MPS_ReadFile(), (char*)"MPS ReadFile Failed!");
    CheckError(CRUSH_Main(), (char*)"Crushing Failed!");
    if (LP_Status != LP_STATUS_OK) { Report_Solution(); return 0; }
    CheckError(Presolve_Main(), (char*)"Presolving Failed!");
    if (LP_Status != LP_STATUS_OK) { Report_Solution(); return 0; }
    double t0 = GetTime();
    CheckError(HSD_Main(), (char*)"Homogeneous and Self-Dual Numerical Solving Failed!");
    double sec = GetTime() - t0;
    Report_Solution();
    std::printf("RESULT status=%d pobj=%.17g dobj=%.17g gap=%.6e pinf=%.6e dinf=%.6e seconds=%.6f m=%d n=%d\n",
                HSD_Status, HSD_primal_obj, HSD_dual_obj, HSD_relative_gap, HSD_primal_infeas,
                HSD_dual_infeas, sec, n_Row, n_Col);
    return 0;
}
