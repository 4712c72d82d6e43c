// Harness around HELP's own Helper.cpp (TEST INFRASTRUCTURE ONLY).
//
// Linked with GlobalVar.o and Helper.o compiled from the reference sources
// (oracle/Makefile) into oracle/_ref/libref_helper.so, so the tests can run
// the reference's ADA^T pattern and values (Helper.cpp:273-377, linker form,
// as UsePardiso.cpp calls them) and its products with A and A^T
// (SetATimesVector, Helper.cpp:110-156) on the same instances as the oracle
// and compare bit for bit.
//
// The harness loads a CSC matrix into HELP's global model the way presolve
// leaves it (Presolve.cpp:60-75,109-112): element p is CSC position p, each
// column list runs by ascending row, each row list by ascending column.  The
// transposed CSR copy and the linker of UsePardiso.cpp:57-80 then coincide
// with the CSC arrays (LinkerTocsrAt[p] = p).  HELP's backend interface
// (LP.h:157-160) is given empty bodies: Helper.cpp names
// LinearEquation_Destruct only in CheckError, which these calls never reach.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "LP.h"

int LinearEquation_Construct() { return 0; }
int LinearEquation_Destruct() { return 0; }
void RenewLinearEquation(double*) {}
int SolveLinearEquation(double*, double*, double*, double*, double*) { return 0; }

namespace {
std::vector<int> g_rowptr_at, g_colind_at, g_linker;
std::vector<double> g_val_at;
}

extern "C" {

// Loads A (m x n, CSC with ascending rows) into HELP's globals; 0 on success.
int ref_load(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowind, const double* val) {
    const int64_t nnz = colptr[n];
    if (m > MAX_ROWS || n > MAX_COLS || nnz > MAX_ELEMENTS) return 1;
    n_Row = m;
    n_Col = n;
    n_Element = (int)nnz;
    for (int32_t j = 0; j < n; ++j) {
        V_Matrix_Col_Head[j] = colptr[j] < colptr[j + 1] ? (int)colptr[j] : -1;
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            V_Matrix_Row[p] = rowind[p];
            V_Matrix_Col[p] = j;
            V_Matrix_Value[p] = val[p];
            V_Matrix_Col_Next[p] = p + 1 < colptr[j + 1] ? (int)(p + 1) : -1;
        }
    }
    // row lists by ascending column: walk the columns backwards, prepending
    for (int32_t i = 0; i < m; ++i) V_Matrix_Row_Head[i] = -1;
    for (int32_t j = n - 1; j >= 0; --j)
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            const int i = rowind[p];
            V_Matrix_Row_Next[p] = V_Matrix_Row_Head[i];
            V_Matrix_Row_Head[i] = (int)p;
        }
    g_rowptr_at.assign(colptr, colptr + n + 1);
    g_colind_at.assign(rowind, rowind + nnz);
    g_val_at.assign(val, val + nnz);
    g_linker.resize(nnz);
    for (int64_t p = 0; p < nnz; ++p) g_linker[p] = (int)p;
    return 0;
}

// HELP's ADA^T upper pattern (ADAt_Allocate, linker form): rowptr[m+1] and,
// when colind is non-null, the column indices in HELP's slot order.
// Returns the number of slots.
int64_t ref_adat_pattern(int64_t* rowptr, int32_t* colind) {
    int nnz = 0;
    double* val = nullptr;
    int* col = nullptr;
    std::vector<int> row(n_Row + 1);
    ADAt_Allocate(&nnz, &val, row.data(), &col, g_linker.data(), g_rowptr_at.data(), g_colind_at.data());
    for (int i = 0; i <= n_Row; ++i) rowptr[i] = row[i];
    if (colind)
        for (int s = 0; s < nnz; ++s) colind[s] = col[s];
    std::free(val);
    std::free(col);
    return nnz;
}

// HELP's ADA^T values (ADAt_Calc, linker form) with dinv = 1/d (as
// UsePardiso.cpp:172-174), in the slot order of ref_adat_pattern.
void ref_adat_values(const double* d, double* vals) {
    int nnz = 0;
    double* val = nullptr;
    int* col = nullptr;
    std::vector<int> row(n_Row + 1);
    ADAt_Allocate(&nnz, &val, row.data(), &col, g_linker.data(), g_rowptr_at.data(), g_colind_at.data());
    std::vector<double> dinv(n_Col);
    for (int j = 0; j < n_Col; ++j) dinv[j] = 1.0 / d[j];
    ADAt_Calc(dinv.data(), val, row.data(), col, g_linker.data(), g_val_at.data(), g_rowptr_at.data(),
              g_colind_at.data());
    std::memcpy(vals, val, sizeof(double) * nnz);
    std::free(val);
    std::free(col);
}

// dest += sign * A v (trans = 0) or dest += sign * A^T v (trans = 1).
void ref_a_times(int32_t trans, int32_t sign, double* v, double* dest) { SetATimesVector(trans, sign, v, dest); }

}  // extern "C"
