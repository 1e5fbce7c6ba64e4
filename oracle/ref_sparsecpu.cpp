// extern "C" driver around the reference's own CPU::SparseCSR (header-only,
// gis/SparseSystem/SparseCPU.h), compiled from the reference tree where it lies.
// Test infrastructure: exposes SparseCSR::transpose (SparseCPU.h:369-399) and
// SparseCSR::JacobiPrecondition (SparseCPU.h:428-446) so the oracle's
// restatements can be checked against the reference code itself.
#include "SparseCPU.h"

#include <algorithm>
#include <cstring>

extern "C" {

__attribute__((visibility("default"))) void ref_transpose(int n_rows, int n_cols,
                                                          const int *rowptr, const int *col,
                                                          const double *val, int *t_rowptr,
                                                          int *t_col, double *t_val) {
    CPU::SparseCSR<double, int> A(n_rows, n_cols, const_cast<double *>(val),
                                  const_cast<int *>(rowptr), const_cast<int *>(col));
    CPU::SparseCSR<double, int> At = A.transpose();
    std::copy(At.row_offsets.begin(), At.row_offsets.end(), t_rowptr);
    std::copy(At.column_indices.begin(), At.column_indices.end(), t_col);
    std::copy(At.values.begin(), At.values.end(), t_val);
}

__attribute__((visibility("default"))) int ref_jacobi_precondition(int n, const int *rowptr,
                                                                   const int *col,
                                                                   const double *val,
                                                                   double *inv_diag) {
    CPU::SparseCSR<double, int> A(n, n, const_cast<double *>(val), const_cast<int *>(rowptr),
                                  const_cast<int *>(col));
    try {
        CPU::SparseCSR<double, int> D = A.JacobiPrecondition();
        std::copy(D.values.begin(), D.values.end(), inv_diag);
    } catch (...) {
        return 1;
    }
    return 0;
}
}
