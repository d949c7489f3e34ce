// The GPU1 algebra as a ugcore build registers it: vector and matrix types with
// the SparseMatrix/Vector surface the reference's callers use, and operator
// classes deriving from the reference's own interfaces
//   ILinearOperator<X,Y>   lib_algebra/operator/interface/linear_operator.h:78-143
//   ILinearIterator<X,Y>   lib_algebra/operator/interface/linear_iterator.h:79-198
//   IPreconditioner<A>     lib_algebra/operator/interface/preconditioner.h:99-340
//   MatrixOperator<M,X,Y>  lib_algebra/operator/interface/matrix_operator.h:47-90
// all forwarding to the ugx_* C ABI of libugcore_amd.so (include/ugcore_amd.h).
//
// This header is compiled inside a ugcore build, with ugbase/ on the include
// path, and linked against libugcore_amd.so; it is not part of the library.
// Registration in the bridge (bridge/util_algebra_dependent.h:64-67 lists
// GPUAlgebra under UG_GPU) then makes `InitUG(dim, AlgebraType("GPU", 1))` the
// one-line switch of a script.
//
// The reference's linear solvers CG<TVector> and BiCGStab<TVector>
// (operator/linear_solver/cg.h, bicgstab.h) are templates over the vector
// type: instantiated with GPUVector they run their loop on the host with every
// vector operation (clone_without_values, VecProd, VecScaleAdd, norm) and the
// preconditioner on the device. Their headers need Boost through lib_disc, so
// this repository's compile check (tests/adapter/adapter_check.cpp) covers the
// three interface headers above, not them. The device-resident CG and BiCGStab
// of the library (ugx_solver_*, with the lattice-order fast path) are the
// faster alternative behind the same names.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "common/util/smart_pointer.h"
#include "lib_algebra/algebra_type.h"
#include "lib_algebra/operator/interface/linear_iterator.h"
#include "lib_algebra/operator/interface/linear_operator.h"
#include "lib_algebra/operator/interface/matrix_operator.h"
#include "lib_algebra/operator/interface/preconditioner.h"
#include "ugcore_amd.h"

namespace ug {

// a failed ugx_* call becomes a UGError, as UG_THROW elsewhere in ugcore
inline void ugx_check(int rc, const char* what) {
    if (rc < 0) UG_THROW(what << ": " << ugx_last_error());
}

// ---------------------------------------------------------------------------
// vector_type (cpu_algebra/vector.h:54-230): device storage behind a handle;
// element access from the host is a cold path (one transfer per access)
// ---------------------------------------------------------------------------
class GPUVector {
   public:
    typedef number value_type;
    typedef GPUVector vector_type;
    typedef size_t size_type;

    // host-side element reference: reads and writes one entry
    class element {
       public:
        element(const GPUVector& v, size_t i) : v_(v), i_(i) {}
        operator number() const { return v_.get(i_); }
        element& operator=(number x) {
            const_cast<GPUVector&>(v_).set_value(i_, x);
            return *this;
        }
        element& operator+=(number x) { return *this = number(*this) + x; }

       private:
        const GPUVector& v_;
        size_t i_;
    };

    GPUVector() {}
    explicit GPUVector(size_t n) { resize(n); }
    GPUVector(const GPUVector& o) { *this = o; }
    GPUVector& operator=(const GPUVector& o) {
        if (this == &o) return *this;
        resize(o.size(), false);
        if (n_ > 0)
            ugx_check(ugx_memcpy_dtod(ugx_vector_device_ptr(h_), ugx_vector_device_ptr(o.h_),
                                      (int64_t)(n_ * sizeof(double))),
                      "GPUVector copy");
        return *this;
    }
    virtual ~GPUVector() {
        if (h_) ugx_vector_destroy(h_);
    }

    size_t size() const { return n_; }
    void resize(size_t n, bool bCopyValues = true) {
        if (n == n_ && h_) return;
        void* h = ugx_vector_create((int64_t)n);
        if (!h) UG_THROW("GPUVector::resize: " << ugx_last_error());
        ugx_check(ugx_vector_set(h, 0.0), "GPUVector::resize");
        const size_t keep = n < n_ ? n : n_;
        if (bCopyValues && keep > 0)
            ugx_check(ugx_memcpy_dtod(ugx_vector_device_ptr(h), ugx_vector_device_ptr(h_),
                                      (int64_t)(keep * sizeof(double))),
                      "GPUVector::resize");
        if (h_) ugx_vector_destroy(h_);
        h_ = h;
        n_ = n;
    }
    void reserve(size_t) {}

    bool set(number w) {
        ugx_check(ugx_vector_set(h_, w), "GPUVector::set");
        return true;
    }
    GPUVector& operator=(number w) {
        set(w);
        return *this;
    }
    number norm() const {
        double r = 0.0;
        ugx_check(ugx_norm_f64((int64_t)n_, ugx_vector_device_ptr(h_), &r, nullptr), "GPUVector::norm");
        return r;
    }
    number dotprod(const GPUVector& w) const {
        if (w.n_ != n_) UG_THROW("GPUVector::dotprod: sizes differ");
        double r = 0.0;
        ugx_check(ugx_dot_f64((int64_t)n_, ugx_vector_device_ptr(h_), ugx_vector_device_ptr(w.h_), &r, nullptr),
                  "GPUVector::dotprod");
        return r;
    }
    GPUVector& operator+=(const GPUVector& v) { return axpby(1.0, 1.0, v); }
    GPUVector& operator-=(const GPUVector& v) { return axpby(1.0, -1.0, v); }
    GPUVector& operator*=(number a) { return axpby(a, 0.0, *this); }

    SmartPtr<GPUVector> clone() const { return make_sp(new GPUVector(*this)); }
    SmartPtr<GPUVector> clone_without_values() const { return make_sp(new GPUVector(n_)); }

    element operator[](size_t i) { return element(*this, i); }
    number operator[](size_t i) const { return get(i); }
    number get(size_t i) const {
        check_index(i);
        double x = 0.0;
        ugx_check(ugx_memcpy_dtoh(&x, ugx_vector_device_ptr(h_) + i, sizeof(double)), "GPUVector[]");
        return x;
    }
    void set_value(size_t i, number x) {
        check_index(i);
        ugx_check(ugx_memcpy_htod(ugx_vector_device_ptr(h_) + i, &x, sizeof(double)), "GPUVector[]");
    }

    void* handle() const { return h_; }
    double* device_ptr() const { return h_ ? ugx_vector_device_ptr(h_) : nullptr; }

    // parallel storage types (lib_algebra/parallelization): one process owns
    // the whole vector, so every storage type holds
    bool has_storage_type(uint) const { return true; }
    void set_storage_type(uint) {}
    bool change_storage_type(uint) { return true; }

   private:
    GPUVector& axpby(number a, number b, const GPUVector& v) {
        if (v.n_ != n_) UG_THROW("GPUVector: sizes differ");
        double* d = device_ptr();
        ugx_check(ugx_vec_scale_add2_f64((int64_t)n_, d, a, d, b, v.device_ptr(), nullptr), "GPUVector update");
        return *this;
    }
    void check_index(size_t i) const {
        if (i >= n_) UG_THROW("GPUVector: index " << i << " out of range (size " << n_ << ")");
    }
    void* h_ = nullptr;
    size_t n_ = 0;
};

// operations_vec.h:55-63, 162-175 and their relatives
inline void VecScaleAdd(GPUVector& dest, number a1, const GPUVector& v1, number a2, const GPUVector& v2) {
    ugx_check(ugx_vec_scale_add2_f64((int64_t)dest.size(), dest.device_ptr(), a1, v1.device_ptr(), a2,
                                     v2.device_ptr(), nullptr),
              "VecScaleAdd");
}
inline void VecScaleAdd(GPUVector& dest, number a1, const GPUVector& v1, number a2, const GPUVector& v2, number a3,
                        const GPUVector& v3) {
    ugx_check(ugx_vec_scale_add3_f64((int64_t)dest.size(), dest.device_ptr(), a1, v1.device_ptr(), a2,
                                     v2.device_ptr(), a3, v3.device_ptr(), nullptr),
              "VecScaleAdd");
}
inline void VecScaleAssign(GPUVector& dest, number a1, const GPUVector& v1) { VecScaleAdd(dest, a1, v1, 0.0, v1); }
inline void VecAssign(GPUVector& dest, const GPUVector& v1) { dest = v1; }
inline number VecProd(const GPUVector& a, const GPUVector& b) { return a.dotprod(b); }
inline number VecTwoNorm(const GPUVector& a) { return a.norm(); }

// ---------------------------------------------------------------------------
// matrix_type (cpu_algebra/sparsematrix.h:98-752): a device CSR behind a
// handle. Pattern creation is host-side (resize_and_clear, operator()(r, c)
// creating the entry on first access, defragment/finalize moving the rows to
// the device); after that the device arrays are the truth and values are
// written by the device assembly or the device operations.
// ---------------------------------------------------------------------------
class GPUSparseMatrix {
   public:
    typedef number value_type;
    typedef GPUVector vector_type;

    // begin_row/end_row: iterates over a host copy of the row
    class const_row_iterator {
       public:
        const_row_iterator(const std::vector<int32_t>* c, const std::vector<double>* v, size_t k)
            : c_(c), v_(v), k_(k) {}
        size_t index() const { return (size_t)(*c_)[k_]; }
        const number& value() const { return (*v_)[k_]; }
        const_row_iterator& operator++() {
            ++k_;
            return *this;
        }
        bool operator!=(const const_row_iterator& o) const { return k_ != o.k_; }
        bool operator==(const const_row_iterator& o) const { return k_ == o.k_; }

       private:
        const std::vector<int32_t>* c_;
        const std::vector<double>* v_;
        size_t k_;
    };

    GPUSparseMatrix() : h_(ugx_matrix_create()) {
        if (!h_) UG_THROW("GPUSparseMatrix: " << ugx_last_error());
    }
    GPUSparseMatrix(const GPUSparseMatrix&) = delete;
    GPUSparseMatrix& operator=(const GPUSparseMatrix&) = delete;
    virtual ~GPUSparseMatrix() { ugx_matrix_destroy(h_); }

    // ---- pattern and sizes ----
    bool resize_and_clear(size_t rows, size_t cols) {
        ugx_check(ugx_matrix_resize_and_clear(h_, (int64_t)rows, (int64_t)cols), "resize_and_clear");
        return true;
    }
    bool resize_and_keep_values(size_t rows, size_t cols) {
        ugx_check(ugx_matrix_resize_and_keep_values(h_, (int64_t)rows, (int64_t)cols), "resize_and_keep_values");
        return true;
    }
    void clear_retain_structure() { ugx_check(ugx_matrix_clear_retain_structure(h_), "clear_retain_structure"); }
    size_t num_rows() const { return (size_t)ugx_matrix_num_rows(h_); }
    size_t num_cols() const { return (size_t)ugx_matrix_num_cols(h_); }
    size_t total_num_connections() const { return (size_t)ugx_matrix_nnz(h_); }
    size_t num_connections(size_t r) const { return (size_t)ugx_matrix_row_length(h_, (int64_t)r); }
    // operator()(r, c): insert-or-get on the host build
    value_type& operator()(size_t r, size_t c) {
        double* p = ugx_matrix_entry(h_, (int64_t)r, (int64_t)c);
        if (!p) UG_THROW("GPUSparseMatrix(" << r << ", " << c << "): " << ugx_last_error());
        return *p;
    }
    value_type operator()(size_t r, size_t c) const {
        double v = 0.0;
        ugx_check(ugx_matrix_get(h_, (int64_t)r, (int64_t)c, &v), "GPUSparseMatrix(r, c)");
        return v;
    }
    bool has_connection(size_t r, size_t c) const {
        double v = 0.0;
        const int rc = ugx_matrix_get(h_, (int64_t)r, (int64_t)c, &v);
        ugx_check(rc, "has_connection");
        return rc == 1;
    }
    // the host build becomes the device CSR (sparsematrix_impl.h:703-757)
    void defragment() {
        if (ugx_matrix_is_building(h_)) ugx_check(ugx_matrix_finalize(h_), "defragment");
    }
    const_row_iterator begin_row(size_t r) const {
        load_row(r);
        return const_row_iterator(&row_c_, &row_v_, 0);
    }
    const_row_iterator end_row(size_t r) const {
        if (r != row_r_) load_row(r);
        return const_row_iterator(&row_c_, &row_v_, row_c_.size());
    }

    // ---- products (sparsematrix_impl.h:271-370) ----
    bool axpy(GPUVector& dest, const number& alpha1, const GPUVector& v1, const number& beta1,
              const GPUVector& w1) const {
        ugx_check(ugx_matrix_axpy(h_, dest.handle(), alpha1, v1.handle(), beta1, w1.handle()), "axpy");
        return true;
    }
    bool axpy_transposed(GPUVector& dest, const number& alpha1, const GPUVector& v1, const number& beta1,
                         const GPUVector& w1) const {
        ugx_check(ugx_matrix_axpy_transposed(h_, dest.handle(), alpha1, v1.handle(), beta1, w1.handle()),
                  "axpy_transposed");
        return true;
    }
    bool apply(GPUVector& dest, const GPUVector& v) const { return axpy(dest, 0.0, dest, 1.0, v); }
    bool apply_transposed(GPUVector& dest, const GPUVector& v) const {
        return axpy_transposed(dest, 0.0, dest, 1.0, v);
    }
    bool matmul_minus(GPUVector& res, const GPUVector& x) const { return axpy(res, 1.0, res, -1.0, x); }
    bool apply_ignore_zero_rows(GPUVector& dest, const number& beta1, const GPUVector& w1) const {
        ugx_check(ugx_matrix_apply_ignore_zero_rows(h_, dest.handle(), beta1, w1.handle()),
                  "apply_ignore_zero_rows");
        return true;
    }

    // ---- whole-matrix operations (sparsematrix_impl.h:148-258, 471-496) ----
    void set_as_transpose_of(const GPUSparseMatrix& B, double scale = 1.0) {
        ugx_check(ugx_matrix_set_as_transpose_of(h_, B.h_, scale, 1), "set_as_transpose_of");
    }
    void set_as_transpose_of2(const GPUSparseMatrix& B, double scale = 1.0) {
        ugx_check(ugx_matrix_set_as_transpose_of(h_, B.h_, scale, 0), "set_as_transpose_of2");
    }
    void set_as_copy_of(const GPUSparseMatrix& B, double scale = 1.0) {
        ugx_check(ugx_matrix_set_as_copy_of(h_, B.h_, scale), "set_as_copy_of");
    }
    void scale(double d) { ugx_check(ugx_matrix_scale(h_, d), "scale"); }

    void* handle() const { return h_; }

   private:
    void load_row(size_t r) const {
        const int64_t n = ugx_matrix_row_length(h_, (int64_t)r);
        if (n < 0) UG_THROW("GPUSparseMatrix::begin_row: " << ugx_last_error());
        row_c_.resize((size_t)n);
        row_v_.resize((size_t)n);
        ugx_check(ugx_matrix_row(h_, (int64_t)r, row_c_.data(), row_v_.data()), "begin_row");
        row_r_ = r;
    }
    void* h_;
    mutable size_t row_r_ = (size_t)-1;
    mutable std::vector<int32_t> row_c_;
    mutable std::vector<double> row_v_;
};

// SetDirichletRow (algebra_common/sparsematrix_util.h:850-861) for the device
// matrix: chosen over the generic template by overload resolution
inline void SetDirichletRow(GPUSparseMatrix& A, size_t i) {
    const int32_t r = (int32_t)i;
    ugx_check(ugx_matrix_set_dirichlet_rows(A.handle(), &r, 1), "SetDirichletRow");
}

// whole-vector matrix products through free functions
// (common/operations_mat/matrix_use_global_functions.h)
template <>
struct matrix_algebra_type_traits<GPUSparseMatrix> {
    enum { type = MATRIX_USE_GLOBAL_FUNCTIONS };
};
inline bool MatMultDirect(GPUVector& dest, const number& beta1, const GPUSparseMatrix& A1, const GPUVector& w1) {
    return A1.axpy(dest, 0.0, dest, beta1, w1);
}
inline bool MatMultAddDirect(GPUVector& dest, const number& alpha1, const GPUVector& v1, const number& beta1,
                             const GPUSparseMatrix& A1, const GPUVector& w1) {
    return A1.axpy(dest, alpha1, v1, beta1, w1);
}
inline bool MatMultTransposedDirect(GPUVector& dest, const number& beta1, const GPUSparseMatrix& A1,
                                    const GPUVector& w1) {
    return A1.axpy_transposed(dest, 0.0, dest, beta1, w1);
}
inline bool MatMultTransposedAddDirect(GPUVector& dest, const number& alpha1, const GPUVector& v1,
                                       const number& beta1, const GPUSparseMatrix& A1, const GPUVector& w1) {
    return A1.axpy_transposed(dest, alpha1, v1, beta1, w1);
}

// ---------------------------------------------------------------------------
// the algebra (cpu_algebra_types.h:76-117)
// ---------------------------------------------------------------------------
struct GPUAlgebra {
    typedef GPUSparseMatrix matrix_type;
    typedef GPUVector vector_type;
    static const int blockSize = 1;
    static AlgebraType get_type() { return AlgebraType(AlgebraType::GPU, 1); }
};

typedef MatrixOperator<GPUSparseMatrix, GPUVector> GPUMatrixOperator;

// ---------------------------------------------------------------------------
// Jacobi<GPUAlgebra> (operator/preconditioner/jacobi.h:89-311) as the
// reference's IPreconditioner: preprocess forms diagInv_i = 1/(a_ii / damp)
// with a constant damping folded in, step applies c = diagInv .* d, and apply
// adds a non-constant damping afterwards (jacobi.h:255-300)
// ---------------------------------------------------------------------------
class GPUJacobi : public IPreconditioner<GPUAlgebra> {
    typedef IPreconditioner<GPUAlgebra> base_type;

   public:
    typedef GPUAlgebra algebra_type;
    typedef GPUVector vector_type;
    typedef GPUSparseMatrix matrix_type;

    GPUJacobi() {}
    explicit GPUJacobi(number damp) { this->set_damp(damp); }
    // the constant damping factor (the multigrid folds it into its smoother)
    number constant_damping() const {
        SmartPtr<IDamping<vector_type, vector_type> > d = const_cast<GPUJacobi*>(this)->damping();
        if (!d->constant_damping()) UG_THROW("Jacobi: the damping is not constant");
        return d->damping();
    }
    GPUJacobi(const GPUJacobi& parent) : base_type(parent) {}
    virtual ~GPUJacobi() { release(); }

    virtual SmartPtr<ILinearIterator<vector_type> > clone() { return make_sp(new GPUJacobi(*this)); }
    virtual bool supports_parallel() const { return false; }   // one GPU per process, no layouts
    virtual const char* name() const { return "Jacobi"; }

    virtual bool apply(vector_type& c, const vector_type& d) {
        if (!this->m_bInit) {
            UG_LOG("ERROR in '" << name() << "::apply': Iterator not initialized.\n");
            return false;
        }
        if (c.size() != d.size() || c.size() != n_) UG_THROW(name() << "::apply: sizes differ");
        if (!step(this->approx_operator(), c, d)) return false;
        if (!this->damping()->constant_damping()) {
            const number kappa = this->damping()->damping(c, d, this->approx_operator());
            if (kappa != 1.0) c *= kappa;
        }
        return true;
    }

   protected:
    virtual bool preprocess(SmartPtr<MatrixOperator<matrix_type, vector_type> > pOp) {
        matrix_type& mat = *pOp;
        if (mat.num_rows() != mat.num_cols()) {
            UG_LOG("Square Matrix needed for Jacobi Iteration.\n");
            return false;
        }
        const int32_t *rp = nullptr, *col = nullptr;
        const double* val = nullptr;
        ugx_check(ugx_matrix_device_csr_readonly(mat.handle(), &rp, &col, &val), "Jacobi::preprocess");
        release();
        n_ = mat.num_rows();
        dinv_ = (double*)ugx_malloc((int64_t)((n_ > 0 ? n_ : 1) * sizeof(double)));
        if (!dinv_) UG_THROW("Jacobi::preprocess: " << ugx_last_error());
        const number damp = this->damping()->constant_damping() ? this->damping()->damping() : 1.0;
        const int rc = ugx_jacobi_diag_inv_f64((int64_t)n_, rp, col, val, damp, dinv_, nullptr);
        if (rc == -2) {
            UG_LOG("Jacobi::preprocess: zero diagonal entry\n");
            return false;
        }
        ugx_check(rc, "Jacobi::preprocess");
        return true;
    }
    virtual bool step(SmartPtr<MatrixOperator<matrix_type, vector_type> >, vector_type& c, const vector_type& d) {
        ugx_check(ugx_jacobi_apply_f64((int64_t)n_, c.device_ptr(), dinv_, d.device_ptr(), nullptr), "Jacobi::step");
        return true;
    }
    virtual bool postprocess() { return true; }

   private:
    void release() {
        if (dinv_) ugx_free(dinv_);
        dinv_ = nullptr;
    }
    double* dinv_ = nullptr;
    size_t n_ = 0;
};

// ---------------------------------------------------------------------------
// AssembledMultiGridCycle<TDomain, GPUAlgebra>
// (lib_disc/operator/linear_operator/multi_grid_solver/mg_solver.h:82-590) as
// the reference's ILinearIterator: the hierarchy, level operators, damped
// Jacobi smoother, transfer and base LU live on the device; set-up at init,
// c = B d at apply. The discretisation handle comes from ugx_disc_create.
// ---------------------------------------------------------------------------
class GPUGeometricMultiGrid : public ILinearIterator<GPUVector> {
   public:
    typedef GPUVector vector_type;
    // cycle types of mg_solver.h:339-341
    enum { _V_ = 1, _W_ = 2, _F_ = -1 };

    explicit GPUGeometricMultiGrid(void* disc) : disc_(disc) {}
    GPUGeometricMultiGrid(const GPUGeometricMultiGrid& o)
        : ILinearIterator<GPUVector>(o),
          disc_(o.disc_),
          base_(o.base_),
          nu1_(o.nu1_),
          nu2_(o.nu2_),
          cycle_(o.cycle_),
          omega_(o.omega_),
          precision_(o.precision_),
          rap_(o.rap_) {}
    virtual ~GPUGeometricMultiGrid() { release(); }

    // the set-up of mg_solver.h:130-225, as the Lua bridge exposes it
    void set_base_level(int l) { base_ = l; }
    void set_num_presmooth(int n) { nu1_ = n; }
    void set_num_postsmooth(int n) { nu2_ = n; }
    void set_cycle_type(int t) { cycle_ = t; }
    void set_cycle_type(const std::string& t) {
        if (t == "V") cycle_ = _V_;
        else if (t == "W") cycle_ = _W_;
        else if (t == "F") cycle_ = _F_;
        else UG_THROW("GMG::set_cycle_type: 'V', 'W' or 'F'");
    }
    void set_rap(bool on) { rap_ = on; }
    // the smoother: damped Jacobi with a constant damping (its factor is read)
    void set_smoother(SmartPtr<GPUJacobi> s) { omega_ = s->constant_damping(); }
    // level operators and smoothing in FP64 (0) or FP32 (1); the base LU is FP64
    void set_precision(int p) { precision_ = p; }

    virtual const char* name() const { return "GMG"; }
    virtual bool supports_parallel() const { return false; }
    virtual SmartPtr<ILinearIterator<GPUVector> > clone() { return make_sp(new GPUGeometricMultiGrid(*this)); }

    virtual bool init(SmartPtr<ILinearOperator<GPUVector> > J, const GPUVector&) { return init(J); }
    virtual bool init(SmartPtr<ILinearOperator<GPUVector> > L) {
        SmartPtr<GPUMatrixOperator> pOp = L.template cast_dynamic<GPUMatrixOperator>();
        if (pOp.invalid()) UG_THROW("GMG::init: the operator is not a GPU matrix operator");
        release();
        g_ = ugx_gmg_create(disc_, base_, nu1_, nu2_, cycle_, omega_, precision_);
        if (!g_) UG_THROW("GMG::init: " << ugx_last_error());
        ugx_check(ugx_gmg_set_rap(g_, rap_ ? 1 : 0), "GMG::init");
        ugx_check(ugx_gmg_init(g_, pOp->handle()), "GMG::init");
        op_ = pOp;
        return true;
    }
    virtual bool apply(GPUVector& c, const GPUVector& d) {
        if (!g_) {
            UG_LOG("ERROR in 'GMG::apply': Iterator not initialized.\n");
            return false;
        }
        ugx_check(ugx_gmg_apply(g_, c.handle(), d.handle()), "GMG::apply");
        const number kappa = this->damping()->damping(c, d, op_);
        if (kappa != 1.0) c *= kappa;
        return true;
    }
    virtual bool apply_update_defect(GPUVector& c, GPUVector& d) {
        if (!apply(c, d)) return false;
        op_->apply_sub(d, c);
        return true;
    }
    void* handle() const { return g_; }

   private:
    void release() {
        if (g_) ugx_gmg_destroy(g_);
        g_ = nullptr;
    }
    void* disc_;
    void* g_ = nullptr;
    int base_ = 0, nu1_ = 2, nu2_ = 2, cycle_ = _V_;
    number omega_ = 0.66;
    int precision_ = 0;
    bool rap_ = false;
    SmartPtr<GPUMatrixOperator> op_;
};

// ---------------------------------------------------------------------------
// AssembledLinearOperator<GPUAlgebra>
// (lib_disc/operator/linear_operator/assembled_linear_operator_impl.h:52-143):
// the matrix operator whose init assembles the discretisation into its own
// matrix (and right-hand side) on the device
// ---------------------------------------------------------------------------
class GPUAssembledLinearOperator : public GPUMatrixOperator {
   public:
    explicit GPUAssembledLinearOperator(void* disc) : disc_(disc) {}
    virtual void init(const GPUVector&) { init(); }
    virtual void init() {
        const int64_t n = ugx_disc_num_dofs(disc_, ugx_disc_num_levels(disc_) - 1);
        rhs_.resize((size_t)n, false);
        ugx_check(ugx_disc_assemble_linear(disc_, this->handle(), rhs_.handle()), "AssembledLinearOperator::init");
    }
    const GPUVector& rhs() const { return rhs_; }

   private:
    void* disc_;
    GPUVector rhs_;
};

}  // namespace ug
