// Oracle (test infrastructure only): CPUAlgebra restatement -- SparseMatrix
// insertion/CSR semantics, SpMV variants, vector reductions, synthetic data.
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

// ---- SparseMatrix restatement -------------------------------------------------

void RowBuilder::resize_and_clear(int64_t r, int64_t c) {
    // sparsematrix_impl.h:90-109
    cols.assign(r, {});
    vals.assign(r, {});
    n_cols = c;
}

double& RowBuilder::at(int64_t r, int64_t c) {
    // SparseMatrix::operator()(r,c) -> get_index: binary search in the sorted
    // row; a missing entry is inserted at its sorted position with value 0
    // (sparsematrix_impl.h:611-700).
    auto& cr = cols[r];
    auto& vr = vals[r];
    auto it = std::lower_bound(cr.begin(), cr.end(), (int)c);
    size_t pos = it - cr.begin();
    if (it == cr.end() || *it != (int)c) {
        cr.insert(it, (int)c);
        vr.insert(vr.begin() + pos, 0.0);
    }
    return vr[pos];
}

CSR defragment(const RowBuilder& B) {
    // copyToNewSize (sparsematrix_impl.h:703-750): rows laid out contiguously
    // in row order, columns ascending, explicit zeros kept.
    CSR A;
    A.n_rows = (int64_t)B.cols.size();
    A.n_cols = B.n_cols;
    A.rowptr.assign(A.n_rows + 1, 0);
    for (int64_t r = 0; r < A.n_rows; ++r) A.rowptr[r + 1] = A.rowptr[r] + (int64_t)B.cols[r].size();
    A.col.resize(A.rowptr.back());
    A.val.resize(A.rowptr.back());
    for (int64_t r = 0; r < A.n_rows; ++r) {
        std::copy(B.cols[r].begin(), B.cols[r].end(), A.col.begin() + A.rowptr[r]);
        std::copy(B.vals[r].begin(), B.vals[r].end(), A.val.begin() + A.rowptr[r]);
    }
    return A;
}

CSR transpose_keep_zeros(const CSR& A) {
    // set_as_transpose_of (sparsematrix_impl.h:148-183): operator()(c, r) = v
    // for every stored entry, zeros included. Row r of A is visited in order,
    // so each transposed row receives ascending column indices.
    CSR T;
    T.n_rows = A.n_cols;
    T.n_cols = A.n_rows;
    T.rowptr.assign(T.n_rows + 1, 0);
    for (int c : A.col) T.rowptr[c + 1]++;
    for (int64_t r = 0; r < T.n_rows; ++r) T.rowptr[r + 1] += T.rowptr[r];
    T.col.resize(A.nnz());
    T.val.resize(A.nnz());
    std::vector<int64_t> fill(T.rowptr.begin(), T.rowptr.end() - 1);
    for (int64_t r = 0; r < A.n_rows; ++r)
        for (int64_t k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k) {
            int64_t p = fill[A.col[k]]++;
            T.col[p] = (int)r;
            T.val[p] = A.val[k];
        }
    return T;
}

CSR transpose_drop_zeros(const CSR& A) {
    // set_as_transpose_of2 (sparsematrix_impl.h:185-254): zero entries skipped
    CSR T;
    T.n_rows = A.n_cols;
    T.n_cols = A.n_rows;
    T.rowptr.assign(T.n_rows + 1, 0);
    for (int64_t k = 0; k < A.nnz(); ++k)
        if (A.val[k] != 0.0) T.rowptr[A.col[k] + 1]++;
    for (int64_t r = 0; r < T.n_rows; ++r) T.rowptr[r + 1] += T.rowptr[r];
    T.col.resize(T.rowptr.back());
    T.val.resize(T.rowptr.back());
    std::vector<int64_t> fill(T.rowptr.begin(), T.rowptr.end() - 1);
    for (int64_t r = 0; r < A.n_rows; ++r)
        for (int64_t k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k) {
            if (A.val[k] == 0.0) continue;
            int64_t p = fill[A.col[k]]++;
            T.col[p] = (int)r;
            T.val[p] = A.val[k];
        }
    return T;
}

CSR multiply_of(const CSR& A, const CSR& B, const CSR& C) {
    if (A.n_cols != B.n_rows || B.n_cols != C.n_rows) throw std::runtime_error("multiply_of: size mismatch");
    CSR M;
    M.n_rows = A.n_rows;
    M.n_cols = C.n_cols;
    M.rowptr.assign(M.n_rows + 1, 0);
    std::vector<double> row(C.n_cols, 0.0);
    std::vector<char> used(C.n_cols, 0);
    std::vector<int> touched;
    for (int64_t i = 0; i < A.n_rows; ++i) {
        touched.clear();
        for (int64_t p = A.rowptr[i]; p < A.rowptr[i + 1]; ++p) {
            if (A.val[p] == 0.0) continue;
            const int k = A.col[p];
            for (int64_t q = B.rowptr[k]; q < B.rowptr[k + 1]; ++q) {
                if (B.val[q] == 0.0) continue;
                const double ab = A.val[p] * B.val[q];
                const int l = B.col[q];
                for (int64_t t = C.rowptr[l]; t < C.rowptr[l + 1]; ++t) {
                    if (C.val[t] == 0.0) continue;
                    const int j = C.col[t];
                    if (!used[j]) {
                        used[j] = 1;
                        touched.push_back(j);
                    }
                    row[j] += ab * C.val[t];
                }
            }
        }
        std::sort(touched.begin(), touched.end());
        for (int j : touched) {
            M.col.push_back(j);
            M.val.push_back(row[j]);
            row[j] = 0.0;
            used[j] = 0;
        }
        M.rowptr[i + 1] = (int64_t)M.col.size();
    }
    return M;
}

void set_row(CSR& A, int64_t r, double v) {
    for (int64_t k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k) A.val[k] = v;
}

double* find_entry(CSR& A, int64_t r, int64_t c) {
    auto b = A.col.begin() + A.rowptr[r], e = A.col.begin() + A.rowptr[r + 1];
    auto it = std::lower_bound(b, e, (int)c);
    if (it == e || *it != (int)c) return nullptr;
    return &A.val[it - A.col.begin()];
}

void axpy(const CSR& A, std::vector<double>& dest, double alpha1, const std::vector<double>& v1,
          double beta1, const std::vector<double>& w1) {
    // sparsematrix_impl.h:291-339; the row sum is accumulated sequentially in
    // CSR order: dest = beta*a_0*w + beta*a_1*w + ...
    const bool same = (&dest == &v1);
    for (int64_t i = 0; i < A.n_rows; ++i) {
        int64_t k = A.rowptr[i], ke = A.rowptr[i + 1];
        if (alpha1 == 0.0) {
            if (k == ke) { dest[i] = 0.0; continue; }
            double s = beta1 * A.val[k] * w1[A.col[k]];
            for (++k; k < ke; ++k) s = s + beta1 * A.val[k] * w1[A.col[k]];
            dest[i] = s;
        } else {
            double s = same ? (alpha1 != 1.0 ? dest[i] * alpha1 : dest[i]) : alpha1 * v1[i];
            for (; k < ke; ++k) s = s + beta1 * A.val[k] * w1[A.col[k]];
            dest[i] = s;
        }
    }
}

void apply_ignore_zero_rows(const CSR& A, std::vector<double>& dest, double beta, const std::vector<double>& w) {
    for (int64_t i = 0; i < A.n_rows; ++i) {
        int64_t k = A.rowptr[i], ke = A.rowptr[i + 1];
        if (k == ke) continue;
        double s = beta * A.val[k] * w[A.col[k]];
        for (++k; k < ke; ++k) s = s + beta * A.val[k] * w[A.col[k]];
        dest[i] = s;
    }
}

void axpy_transposed(const CSR& A, std::vector<double>& dest, double alpha1, const std::vector<double>& v1,
                     double beta1, const std::vector<double>& w1) {
    if (&dest == &v1) {
        if (alpha1 == 0.0) std::fill(dest.begin(), dest.end(), 0.0);
        else if (alpha1 != 1.0) for (auto& x : dest) x *= alpha1;
    } else if (alpha1 == 0.0) {
        std::fill(dest.begin(), dest.end(), 0.0);
    } else {
        for (size_t i = 0; i < dest.size(); ++i) dest[i] = alpha1 * v1[i];
    }
    for (int64_t i = 0; i < A.n_rows; ++i)
        for (int64_t k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k)
            if (A.val[k] != 0.0) dest[A.col[k]] = dest[A.col[k]] + beta1 * A.val[k] * w1[i];
}

void apply_sub(const CSR& A, std::vector<double>& d, const std::vector<double>& x) {
    axpy(A, d, 1.0, d, -1.0, x);
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

double norm(const std::vector<double>& a) { return std::sqrt(dot(a, a)); }

// ---- synthetic data -------------------------------------------------------------

uint64_t splitmix64(uint64_t x) {
    // SplitMix64 (Steele, Lea, Flood 2014), a counter-based mixing function
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

static double unit_open(uint64_t h) { return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

double gauss_from_counter(uint64_t seed, uint64_t counter) {
    // Box-Muller on two counter-derived uniforms in (0,1)
    const uint64_t base = splitmix64(seed) ^ (counter * 2);
    const double u1 = unit_open(splitmix64(base));
    const double u2 = unit_open(splitmix64(base ^ 1ull));
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

double uniform_from_counter(uint64_t seed, uint64_t counter) {
    return unit_open(splitmix64(splitmix64(seed) ^ counter));
}

std::vector<double> make_theta(int n, uint64_t seed, int passes) {
    const int64_t N = (int64_t)n * n * n;
    std::vector<double> t(N), s(N);
    for (int64_t e = 0; e < N; ++e) t[e] = gauss_from_counter(seed, (uint64_t)e);
    for (int p = 0; p < passes; ++p) {
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double acc = 0;
                    int cnt = 0;
                    for (int dk = -1; dk <= 1; ++dk) {
                        int kk = k + dk;
                        if (kk < 0 || kk >= n) continue;
                        for (int dj = -1; dj <= 1; ++dj) {
                            int jj = j + dj;
                            if (jj < 0 || jj >= n) continue;
                            for (int di = -1; di <= 1; ++di) {
                                int ii = i + di;
                                if (ii < 0 || ii >= n) continue;
                                acc += t[((int64_t)kk * n + jj) * n + ii];
                                ++cnt;
                            }
                        }
                    }
                    s[((int64_t)k * n + j) * n + i] = acc / cnt;
                }
        t.swap(s);
    }
    return t;
}

double Coefficient::eval(const double* x) const {
    if (!theta) return 1.0;
    int64_t idx = 0, stride = 1;
    for (int d = 0; d < dim; ++d) {
        int64_t c = (int64_t)std::floor(x[d] * n_fine);
        if (c < 0) c = 0;
        if (c >= n_fine) c = n_fine - 1;
        idx += c * stride;
        stride *= n_fine;
    }
    return std::exp(theta[idx]);
}

}  // namespace orc
