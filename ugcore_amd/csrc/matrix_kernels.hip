// Device side of the SparseMatrix surface beyond the products:
//   * transposes: SparseMatrix::set_as_transpose_of (every stored entry,
//     explicit zeros included) and set_as_transpose_of2 (zeros dropped)
//     (lib_algebra/cpu_algebra/sparsematrix_impl.h:148-240). The transposed
//     rows list their entries by ascending source row, which is what the
//     reference's row-by-row insertion into sorted rows produces;
//   * the transposed product dest = a1*v1 + b1*A^T*w1 (axpy_transposed,
//     sparsematrix_impl.h:341-370) over a cached transpose, one thread per
//     transposed row accumulating in ascending source row as the reference's
//     row loop does, skipping zero entries as the reference does;
//   * Dirichlet rows (SetDirichletRow, algebra_common/sparsematrix_util.h:850-861);
//   * value scaling.
//
// The transpose is a stable radix sort of (column, entry) pairs: entries are
// stored row by row, so a stable sort by column keeps the source rows ascending
// within each column. The transposed row pointers are read off the sorted keys
// (the first position of every key), so no atomics and no scan are involved and
// the result is deterministic.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <cstdint>

#include "ugx_internal.hpp"

namespace ugamd {

namespace {

inline int64_t blocks_for(int64_t n, int threads) {
    int64_t g = (n + threads - 1) / threads;
    return g < 1 ? 1 : g;
}

// the row of every entry (one thread per row)
__global__ __launch_bounds__(256) void k_entry_rows(int64_t nrows, const int32_t* __restrict__ rowptr,
                                                    int32_t* __restrict__ row_of) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    for (int32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) row_of[k] = (int32_t)r;
}

// sort keys: the column, or ncols for a dropped (zero) entry, which sorts
// after every column; values: the entry index
__global__ __launch_bounds__(256) void k_transpose_keys(int64_t nnz, const int32_t* __restrict__ col,
                                                        const double* __restrict__ val, int drop_zeros, int32_t ncols,
                                                        uint32_t* __restrict__ key, int32_t* __restrict__ idx) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    key[k] = (uint32_t)((drop_zeros && val[k] == 0.0) ? ncols : col[k]);
    idx[k] = (int32_t)k;
}

// rowptr_t[c] = first sorted position whose key is >= c, for c = 0..ncols
// (thread k fills the columns between the previous key and its own)
__global__ __launch_bounds__(256) void k_rowptr_from_keys(int64_t nnz, const uint32_t* __restrict__ key,
                                                          int32_t ncols, int32_t* __restrict__ rowptr_t) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k > nnz) return;
    const int64_t lo = k == 0 ? -1 : (int64_t)key[k - 1];
    const int64_t hi = k == nnz ? (int64_t)ncols : (int64_t)key[k];
    for (int64_t c = lo + 1; c <= hi && c <= ncols; ++c) rowptr_t[c] = (int32_t)k;
}

__global__ __launch_bounds__(256) void k_transpose_gather(int64_t n, const int32_t* __restrict__ idx,
                                                          const int32_t* __restrict__ row_of,
                                                          const double* __restrict__ val, double scale,
                                                          int32_t* __restrict__ col_t, double* __restrict__ val_t) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t e = idx[k];
    col_t[k] = row_of[e];
    val_t[k] = scale * val[e];
}

// dest[c] = a1*v1[c] (v1_is_dest: dest[c]*a1; a1 == 0: 0), then for the
// entries of transposed row c in ascending source row r: dest[c] += (b1*a)*w[r]
// for a != 0
__global__ __launch_bounds__(256) void k_axpy_transposed(int64_t n, const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ cl, const double* __restrict__ vl,
                                                         double a1, const double* v1, int v1_is_dest, double b1,
                                                         const double* __restrict__ w, double* dest) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= n) return;
    double d;
    if (a1 == 0.0) d = 0.0;
    else if (v1_is_dest) d = a1 == 1.0 ? dest[c] : dest[c] * a1;
    else d = a1 * v1[c];
    for (int32_t k = rp[c]; k < rp[c + 1]; ++k) {
        const double a = vl[k];
        if (a != 0.0) d = d + (b1 * a) * w[cl[k]];
    }
    dest[c] = d;
}

// SetDirichletRow for every listed row: all entries 0, the diagonal 1 (the
// host checked that every listed row stores its diagonal)
__global__ __launch_bounds__(256) void k_dirichlet_rows(int64_t nlist, const int32_t* __restrict__ rows,
                                                        const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ col, double* __restrict__ val) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nlist) return;
    const int32_t r = rows[t];
    for (int32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) val[k] = col[k] == r ? 1.0 : 0.0;
}

// one flag per listed row: 1 if the row stores no diagonal entry
__global__ __launch_bounds__(256) void k_missing_diagonal(int64_t nlist, const int32_t* __restrict__ rows,
                                                          const int32_t* __restrict__ rowptr,
                                                          const int32_t* __restrict__ col, int* __restrict__ missing) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nlist) return;
    const int32_t r = rows[t];
    bool found = false;
    for (int32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) found = found || col[k] == r;
    if (!found) *missing = 1;
}

__global__ __launch_bounds__(256) void k_scale_values(int64_t n, const double* __restrict__ src, double a,
                                                      double* __restrict__ dst) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) dst[k] = a * src[k];
}

}  // namespace

int transpose_sort_bytes(int64_t nnz, int32_t ncols, size_t* bytes) {
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) <= (int64_t)ncols) ++bits;
    *bytes = 0;
    if (nnz == 0) return 0;
    return rocprim::radix_sort_pairs(nullptr, *bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)nnz, 0, bits) == hipSuccess
               ? 0
               : -1;
}

int transpose_csr(int64_t nrows, int32_t ncols, int64_t nnz, const int32_t* rowptr, const int32_t* col,
                  const double* val, int drop_zeros, uint32_t* key, uint32_t* key2, int32_t* idx, int32_t* idx2,
                  int32_t* row_of, void* temp, size_t temp_bytes, int32_t* rowptr_t, hipStream_t s) {
    if (nnz > 0) {
        hipLaunchKernelGGL(k_entry_rows, dim3(blocks_for(nrows, 256)), dim3(256), 0, s, nrows, rowptr, row_of);
        hipLaunchKernelGGL(k_transpose_keys, dim3(blocks_for(nnz, 256)), dim3(256), 0, s, nnz, col, val, drop_zeros,
                           ncols, key, idx);
        int bits = 1;
        while (bits < 31 && ((int64_t)1 << bits) <= (int64_t)ncols) ++bits;
        size_t tb = temp_bytes;
        if (rocprim::radix_sort_pairs(temp, tb, key, key2, idx, idx2, (size_t)nnz, 0, bits, s) != hipSuccess)
            return -1;
    }
    hipLaunchKernelGGL(k_rowptr_from_keys, dim3(blocks_for(nnz + 1, 256)), dim3(256), 0, s, nnz, key2, ncols,
                       rowptr_t);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int transpose_gather(int64_t n, const int32_t* idx_sorted, const int32_t* row_of, const double* val, double scale,
                     int32_t* col_t, double* val_t, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_transpose_gather, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, idx_sorted, row_of, val,
                       scale, col_t, val_t);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int axpy_transposed(int64_t n, const int32_t* rp, const int32_t* cl, const double* vl, double a1, const double* v1,
                    bool v1_is_dest, double b1, const double* w, double* dest, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_axpy_transposed, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, rp, cl, vl, a1, v1,
                       v1_is_dest ? 1 : 0, b1, w, dest);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int missing_diagonals(int64_t nlist, const int32_t* rows, const int32_t* rowptr, const int32_t* col,
                      int* missing_flag, hipStream_t s) {
    if (nlist <= 0) return 0;
    hipLaunchKernelGGL(k_missing_diagonal, dim3(blocks_for(nlist, 256)), dim3(256), 0, s, nlist, rows, rowptr, col,
                       missing_flag);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dirichlet_rows(int64_t nlist, const int32_t* rows, const int32_t* rowptr, const int32_t* col, double* val,
                   hipStream_t s) {
    if (nlist <= 0) return 0;
    hipLaunchKernelGGL(k_dirichlet_rows, dim3(blocks_for(nlist, 256)), dim3(256), 0, s, nlist, rows, rowptr, col,
                       val);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int scale_values(int64_t n, const double* src, double a, double* dst, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_scale_values, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, src, a, dst);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugamd
