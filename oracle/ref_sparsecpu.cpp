// extern "C" driver around the reference's own CPU::SparseCSR (header-only,
// gis/SparseSystem/SparseCPU.h), compiled from the reference tree where it lies.
// Test infrastructure: exposes SparseCSR::transpose (SparseCPU.h:369-399),
// SparseCSR::JacobiPrecondition (SparseCPU.h:428-446) and the file formats
// (SparseCSR::save/load/readRB/read, SparseCPU.h:105-311; Vector::save/load/
// readMTX, SparseCPU.h:653-689) so the oracle's restatements and the
// package's readers can be checked against the reference code itself.
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

// ---- file formats ----------------------------------------------------------
__attribute__((visibility("default"))) int ref_csr_save(const char *path, int n_rows, int n_cols,
                                                        const int *rowptr, const int *col,
                                                        const double *val) {
    CPU::SparseCSR<double, int> A(n_rows, n_cols, const_cast<double *>(val),
                                  const_cast<int *>(rowptr), const_cast<int *>(col));
    A.save(path);
    return 0;
}

// read (format chosen by suffix, SparseCPU.h:305-311): sizes first, then the
// arrays into caller buffers of those sizes (two calls read the file twice)
__attribute__((visibility("default"))) int ref_csr_read_sizes(const char *path, int *n_rows,
                                                              int *n_cols, int *nnz) {
    try {
        CPU::SparseCSR<double, int> A;
        A.read(path);
        *n_rows = A.num_rows;
        *n_cols = A.num_cols;
        *nnz = A.num_entries;
    } catch (...) {
        return 1;
    }
    return 0;
}

__attribute__((visibility("default"))) int ref_csr_read(const char *path, int *rowptr, int *col,
                                                        double *val) {
    try {
        CPU::SparseCSR<double, int> A;
        A.read(path);
        std::copy(A.row_offsets.begin(), A.row_offsets.end(), rowptr);
        std::copy(A.column_indices.begin(), A.column_indices.end(), col);
        std::copy(A.values.begin(), A.values.end(), val);
    } catch (...) {
        return 1;
    }
    return 0;
}

__attribute__((visibility("default"))) int ref_vec_save(const char *path, int n,
                                                        const double *v) {
    CPU::Vector<double, int> x(n);
    std::copy(v, v + n, x.begin());
    x.save(path);
    return 0;
}

// mtx = 0: binary Vector::load; 1: Vector::readMTX.  out == nullptr: size only.
__attribute__((visibility("default"))) int ref_vec_read(const char *path, int mtx, int *n,
                                                        double *out) {
    try {
        CPU::Vector<double, int> x;
        if (mtx) x.readMTX(path);
        else x.load(path);
        *n = x.size;
        if (out) std::copy(x.begin(), x.end(), out);
    } catch (...) {
        return 1;
    }
    return 0;
}
}
