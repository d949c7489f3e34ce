// Thin C wrapper around the reference's own header-only SparseMatrix<double>
// and Vector<double> (lib_algebra/cpu_algebra/sparsematrix_impl.h,
// vector_impl.h), compiled from the reference checkout into oracle/_ref/.
// Test infrastructure only: it lets tests compare the oracle's CSR semantics
// (insertion, explicit zeros, transposes, SpMV variants) with the real code.
#include <cstdint>
#include <cstring>

#include "lib_algebra/cpu_algebra/sparsematrix_impl.h"
#include "lib_algebra/cpu_algebra/vector_impl.h"
#include "lib_algebra/algebra_common/sparsematrix_util.h"
#include "common/serialization.h"
#include "common/util/binary_buffer.h"

using Mat = ug::SparseMatrix<double>;
using Vec = ug::Vector<double>;

extern "C" {

void* ref_sm_create(int64_t rows, int64_t cols) {
    Mat* m = new Mat();
    m->resize_and_clear(rows, cols);
    return m;
}
void ref_sm_destroy(void* h) { delete (Mat*)h; }
void ref_sm_set(void* h, int64_t r, int64_t c, double v) { (*(Mat*)h)(r, c) = v; }
void ref_sm_add(void* h, int64_t r, int64_t c, double v) { (*(Mat*)h)(r, c) += v; }

// mode 0: as built; 1: set_as_transpose_of; 2: set_as_transpose_of2
void* ref_sm_derive(void* h, int mode) {
    Mat* a = (Mat*)h;
    Mat* b = new Mat();
    if (mode == 0) b->set_as_copy_of(*a);
    else if (mode == 1) b->set_as_transpose_of(*a);
    else b->set_as_transpose_of2(*a);
    return b;
}

// M = A*B*C through AddMultiplyOf into a cleared matrix (the Galerkin
// product of AssembledMultiGridCycle::init_rap_operator)
void* ref_sm_multiply3(void* a, void* b, void* c) {
    const Mat& A = *(const Mat*)a;
    const Mat& C = *(const Mat*)c;
    Mat* m = new Mat();
    m->resize_and_clear(A.num_rows(), C.num_cols());
    ug::AddMultiplyOf(*m, A, *(const Mat*)b, C);
    return m;
}

int64_t ref_sm_rows(void* h) { return (int64_t)((Mat*)h)->num_rows(); }
int64_t ref_sm_nnz(void* h) { return (int64_t)((Mat*)h)->total_num_connections(); }

// CSR export through the row iterators (defragmented order: row by row, sorted columns)
void ref_sm_csr(void* h, int64_t* rowptr, int32_t* col, double* val) {
    const Mat& a = *(const Mat*)h;
    int64_t k = 0;
    rowptr[0] = 0;
    for (size_t r = 0; r < a.num_rows(); ++r) {
        Mat::const_row_iterator itEnd = a.end_row(r);
        for (Mat::const_row_iterator it = a.begin_row(r); it != itEnd; ++it) {
            col[k] = (int32_t)it.index();
            val[k] = it.value();
            ++k;
        }
        rowptr[r + 1] = k;
    }
}

// op 0: axpy dest = alpha*v1 + beta*A*w; 1: apply_ignore_zero_rows; 2: axpy_transposed
void ref_sm_op(void* h, int op, double alpha, const double* v1, double beta, const double* w, double* dest) {
    Mat* a = (Mat*)h;
    const size_t nd = op == 2 ? a->num_cols() : a->num_rows();
    const size_t nw = op == 2 ? a->num_rows() : a->num_cols();
    Vec d(nd), vv(nd), ww(nw);
    for (size_t i = 0; i < nd; ++i) { d[i] = dest[i]; vv[i] = v1 ? v1[i] : dest[i]; }
    for (size_t i = 0; i < nw; ++i) ww[i] = w[i];
    if (op == 0) a->axpy(d, alpha, vv, beta, ww);
    else if (op == 1) a->apply_ignore_zero_rows(d, beta, ww);
    else a->axpy_transposed(d, alpha, vv, beta, ww);
    for (size_t i = 0; i < nd; ++i) dest[i] = d[i];
}

// the bytes Serialize(BinaryBuffer, Vector<double>) produces (the payload of
// SaveToFile, restart_bridge.cpp:146-151); returns their count, copies at most
// cap of them to out
int64_t ref_vec_serialize(int64_t n, const double* a, char* out, int64_t cap) {
    Vec x(n);
    for (int64_t i = 0; i < n; ++i) x[i] = a[i];
    ug::BinaryBuffer b;
    ug::Serialize(b, x);
    const int64_t len = (int64_t)b.write_pos();
    if (out) std::memcpy(out, b.buffer(), (size_t)(len < cap ? len : cap));
    return len;
}

double ref_vec_dot(int64_t n, const double* a, const double* b) {
    Vec x(n), y(n);
    for (int64_t i = 0; i < n; ++i) { x[i] = a[i]; y[i] = b[i]; }
    return x.dotprod(y);
}

double ref_vec_norm(int64_t n, const double* a) {
    Vec x(n);
    for (int64_t i = 0; i < n; ++i) x[i] = a[i];
    return x.norm();
}

}  // extern "C"
