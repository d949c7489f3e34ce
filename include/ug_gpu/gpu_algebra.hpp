// GPUAlgebra for ugcore on MI355X (gfx950): the C++ host classes that sit
// above the ugx_* C ABI (include/ugcore_amd.h).
//
// They mirror the reference's algebra and operator interfaces so the bridge
// can register them beside CPUAlgebra under the "GPU1" suffix:
//   GPUAlgebra                  lib_algebra/cpu_algebra_types.h:76-121 (CPUAlgebra / commented GPUAlgebra)
//   GPUVector                   lib_algebra/cpu_algebra/vector.h:90-176
//   GPUSparseMatrix             lib_algebra/cpu_algebra/sparsematrix.h:98-752
//   ILinearOperator             lib_algebra/operator/interface/linear_operator.h:78-143
//   ILinearIterator             lib_algebra/operator/interface/linear_iterator.h:79-198
//   MatrixOperator              lib_algebra/operator/interface/matrix_operator.h:47
//   StdConvCheck                lib_algebra/operator/convergence_check.h:178-262
//   Jacobi                      lib_algebra/operator/preconditioner/jacobi.h:89-311
//   CG, BiCGStab                lib_algebra/operator/linear_solver/cg.h:64-304, bicgstab.h:112-385
//   AssembledMultiGridCycle     lib_disc/operator/linear_operator/multi_grid_solver/mg_solver.h:82-590
//   DomainDiscretization        lib_disc/spatial_disc/domain_disc.h:77,649 (structured device assembly)
//
// Device memory is owned by the objects (hipMalloc at set-up, never inside
// an iteration); all work is ordered on one HIP stream per process.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace ug {
namespace gpu {

struct UGError : public std::runtime_error {
    explicit UGError(const std::string& m) : std::runtime_error(m) {}
};

// ---------------------------------------------------------------------------
// device context and buffers
// ---------------------------------------------------------------------------
struct DeviceContext {
    int device = -1;                 // the device the context was created on
    hipStream_t stream = nullptr;
    double* red_partial = nullptr;   // 65536 partial sums (one per block of a reduction)
    double* red_out = nullptr;       // 8 results
    double* red_host = nullptr;      // pinned host mirror of red_out
    int* flag = nullptr;             // device error flag
};
DeviceContext& context();            // initialises the device on first use
DeviceContext* context_if_ready();   // nullptr before the first context() (allocates nothing)
void sync();

template <typename T>
class DeviceBuffer {
   public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(int64_t n) { resize(n); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void resize(int64_t n);
    void release();
    T* get() const { return p_; }
    int64_t size() const { return n_; }
    void upload(const T* h, int64_t n, int64_t offset = 0);
    void download(T* h, int64_t n, int64_t offset = 0) const;
    void zero();

   private:
    T* p_ = nullptr;
    int64_t n_ = 0;
};

// ---------------------------------------------------------------------------
// algebra types
// ---------------------------------------------------------------------------
class GPUVector {
   public:
    typedef double value_type;
    GPUVector() = default;
    explicit GPUVector(int64_t n) { resize(n); }
    GPUVector(const GPUVector& o);
    GPUVector& operator=(const GPUVector& o);
    void resize(int64_t n);
    int64_t size() const { return n_; }
    void set(double v);
    double norm() const;
    double dotprod(const GPUVector& b) const;
    GPUVector& operator+=(const GPUVector& b);
    GPUVector& operator-=(const GPUVector& b);
    GPUVector& operator*=(double a);
    std::unique_ptr<GPUVector> clone() const { return std::unique_ptr<GPUVector>(new GPUVector(*this)); }
    std::unique_ptr<GPUVector> clone_without_values() const { return std::unique_ptr<GPUVector>(new GPUVector(n_)); }
    double* data() { return buf_.get(); }
    const double* data() const { return buf_.get(); }
    void upload(const double* h) { buf_.upload(h, n_); }
    void download(double* h) const { buf_.download(h, n_); }

   private:
    int64_t n_ = 0;
    DeviceBuffer<double> buf_;
};

// VecScaleAdd (operations_vec.h:55-63, 162-175)
void VecScaleAdd(GPUVector& dest, double a1, const GPUVector& v1, double a2, const GPUVector& v2);
void VecScaleAdd(GPUVector& dest, double a1, const GPUVector& v1, double a2, const GPUVector& v2, double a3,
                 const GPUVector& v3);

class GPUSparseMatrix {
   public:
    typedef double value_type;
    GPUSparseMatrix() = default;
    int64_t num_rows() const { return rows_; }
    int64_t num_cols() const { return cols_; }
    int64_t total_num_connections() const { return nnz_; }

    // ---- pattern creation on the host (sparsematrix_impl.h:90-145, 553-700) ----
    // resize_and_clear starts a host-side build: operator()(r, c) inserts the
    // entry (value 0) on first access and keeps every row sorted by column, as
    // SparseMatrix does; finalize() moves the rows into the device CSR, after
    // which the device arrays are the truth and operator() throws (values are
    // then written by the device assembly or the device operations below).
    void resize_and_clear(int64_t rows, int64_t cols);
    // keeps the entries (downloading a device CSR into a host build); entries
    // in columns >= cols are dropped, rows beyond `rows` are removed or added
    void resize_and_keep_values(int64_t rows, int64_t cols);
    // all values 0, pattern kept (host build or device)
    void clear_retain_structure();
    bool building() const { return building_; }
    double& operator()(int64_t r, int64_t c);
    // the stored value, 0 for an entry that is not stored (host build or device)
    double get(int64_t r, int64_t c) const;
    bool has_connection(int64_t r, int64_t c) const;
    // row access (begin_row/end_row): number of stored entries of row r and a
    // host copy of its columns and values (device CSR: downloaded)
    int64_t num_connections(int64_t r) const;
    void row(int64_t r, int32_t* cols, double* vals) const;

    // ---- device operations (sparsematrix_impl.h:148-240, 341-370) ----
    // this = scale * B^T keeping every stored entry of B, explicit zeros
    // included (set_as_transpose_of) or dropping zero entries
    // (set_as_transpose_of2); rows sorted by column (= source row)
    void set_as_transpose_of(const GPUSparseMatrix& B, double scale = 1.0) { transpose_from(B, scale, false); }
    void set_as_transpose_of2(const GPUSparseMatrix& B, double scale = 1.0) { transpose_from(B, scale, true); }
    void set_as_copy_of(const GPUSparseMatrix& B, double scale = 1.0);
    void scale(double a);
    // dest = a1*v1 + b1*A^T*w1 (v1 may alias dest; w1 must not)
    void axpy_transposed(GPUVector& dest, double a1, const GPUVector& v1, double b1, const GPUVector& w1) const;
    void apply_transposed(GPUVector& dest, const GPUVector& w) const { axpy_transposed(dest, 0.0, dest, 1.0, w); }
    // SetDirichletRow (sparsematrix_util.h:850-861) for every listed row: all
    // entries 0, the diagonal 1; every listed row must store its diagonal
    void set_dirichlet_rows(const int32_t* rows_host, int64_t n);
    // create from host CSR (row pointers int32, columns sorted per row)
    void set_from_host_csr(int64_t rows, int64_t cols, const int32_t* rowptr, const int32_t* col, const double* val);
    // allocate the device pattern (values zero)
    void allocate(int64_t rows, int64_t cols, int64_t nnz);
    void download_csr(int32_t* rowptr, int32_t* col, double* val) const;
    // dest = a1*v1 + b1*A*w1 (sparsematrix_impl.h:291-339)
    void axpy(GPUVector& dest, double a1, const GPUVector& v1, double b1, const GPUVector& w1) const;
    void apply(GPUVector& dest, const GPUVector& w) const { axpy(dest, 0.0, dest, 1.0, w); }
    void matmul_minus(GPUVector& dest, const GPUVector& w) const { axpy(dest, 1.0, dest, -1.0, w); }
    void apply_ignore_zero_rows(GPUVector& dest, double beta, const GPUVector& w) const;
    int32_t* rowptr() const { return rowptr_.get(); }
    int32_t* col() const { return col_.get(); }
    double* val() const { return val_.get(); }
    // ends a host build (resize_and_clear + operator()): the sorted rows become
    // the device CSR (SparseMatrix::defragment, sparsematrix_impl.h:703-757).
    // Otherwise: the values were changed through the raw CSR pointers, so the
    // provenance is dropped (operators derived from an assembly, such as the
    // matrix-free and lattice-order paths of the multigrid, no longer apply to
    // this matrix) and the SpMV layout is rebuilt from the CSR
    void finalize();
    // write access to the CSR values is handed out: the provenance is dropped
    // and SpMV reads the CSR itself until finalize() rebuilds the layout
    void values_edited_externally() {
        require_device("device_csr");
        set_origin(nullptr);
        pattern_id_ = fresh_pattern_id();   // the raw pointers also reach the pattern
        sell_valid_ = false;
        dict_valid_ = false;
    }
    // identifies the current pattern: a fresh process-wide id on every
    // allocation, host build and hand-out of the raw CSR pointers
    uint64_t pattern_id() const { return pattern_id_; }
    // the values changed on the device and the pattern did not: the SpMV
    // layout takes the new values (only the values are rewritten when the
    // layout was built for the current pattern, else it is rebuilt)
    void refresh_layout_values();
    // build the sliced-ELL SpMV layout from the current CSR (provenance kept;
    // assembly and uploads call it); until then SpMV runs on the CSR directly
    void build_spmv_layout();
    bool uses_sliced_ell() const { return sell_valid_; }
    // provenance: the discretisation that assembled the values and a counter
    // bumped on every assembly/upload (used to validate derived operators)
    const void* origin() const { return origin_; }
    uint64_t version() const { return version_; }
    void set_origin(const void* o) { origin_ = o; ++version_; }
    void set_use_sliced_ell(bool on) {
        sell_enabled_ = on;
        if (!on) sell_valid_ = false;
        else if (!sell_valid_) build_spmv_layout();
    }
    // the row permutation that fills the sliced-ELL slices (rows order[64t..
    // 64t+63] form slice t); it groups rows whose columns lie close together so
    // that the x gathers of a wavefront stay local. Results do not depend on
    // it. Host array, checked to be a permutation of [0, rows); nullptr clears.
    // Assembly sets the lattice walk of the discretisation.
    void set_row_order(const int32_t* order, int64_t n);
    // the same from device memory that is a permutation by construction
    void set_row_order_device(const int32_t* order, int64_t n);
    // the slices of the dictionary layout (spmv_dict.hip): slot s (a multiple
    // of 64 of them) holds row order[s] or -1, every row exactly once; slices
    // whose entries touch at most 255 distinct columns (64 lattice points of a
    // 4x4x4 block) take the layout, which stages x at those columns once per
    // slice and stores one byte per column index. Device array (checked on
    // the host unless by_construction, for orders the library generates);
    // nullptr clears.
    void set_slice_order_device(const int32_t* order, int64_t nslots, bool by_construction = false);
    bool uses_dictionary() const { return dict_valid_; }
    // 0: CSR, 1: sliced ELL, 2: the dictionary layout where a slice order
    // exists (else sliced ELL)
    void set_spmv_layout(int layout) {
        dict_enabled_ = layout == 2;
        sell_enabled_ = layout >= 1;
        build_spmv_layout();
    }
    bool has_row_order() const { return rows_ > 0 && row_order_.size() == rows_; }
    int lanes() const { return lanes_; }
    void set_lanes(int g) {
        if (g != 4 && g != 8 && g != 16 && g != 32 && g != 64) throw UGError("set_lanes: 4, 8, 16, 32 or 64");
        lanes_ = g;
    }

   private:
    void require_device(const char* what) const;
    void transpose_from(const GPUSparseMatrix& B, double scale, bool drop_zeros);
    // the values changed on the device: derived operators and the SpMV layout
    // no longer apply
    void values_changed_on_device();
    int64_t rows_ = 0, cols_ = 0, nnz_ = 0;
    int lanes_ = 8;
    // host build (pattern creation): one sorted (column, value) list per row
    bool building_ = false;
    std::vector<std::vector<std::pair<int32_t, double>>> hrows_;
    // transpose of the current values without zero entries (axpy_transposed),
    // valid while t_version_ == version_
    mutable uint64_t t_version_ = ~(uint64_t)0;
    mutable int64_t t_nnz_ = 0;
    mutable DeviceBuffer<int32_t> t_rowptr_, t_col_;
    mutable DeviceBuffer<double> t_val_;
    DeviceBuffer<int32_t> rowptr_, col_;
    DeviceBuffer<double> val_;
    bool sell_enabled_ = true, sell_valid_ = false;
    const void* origin_ = nullptr;
    uint64_t version_ = 0;
    static uint64_t fresh_pattern_id();
    uint64_t pattern_id_ = 0, layout_pattern_id_ = ~(uint64_t)0;
    DeviceBuffer<int32_t> sell_width_, sell_col_, row_order_;
    DeviceBuffer<int64_t> sell_ptr_;
    DeviceBuffer<double> sell_val_;
    // dictionary layout
    bool build_dictionary();
    bool dict_enabled_ = true, dict_valid_ = false;
    DeviceBuffer<int32_t> slice_order_, d_width_, d_dptr_, d_dict_;
    DeviceBuffer<int64_t> d_vptr_, d_iptr_;
    DeviceBuffer<double> d_val_;
    DeviceBuffer<uint32_t> d_idx_;
};

struct AlgebraType {
    enum Type { CPU = 0, GPU = 1 };
    Type type;
    int blocksize;
};

struct GPUAlgebra {
    typedef GPUSparseMatrix matrix_type;
    typedef GPUVector vector_type;
    static const int blockSize = 1;
    static AlgebraType get_type() { return AlgebraType{AlgebraType::GPU, 1}; }
};

// ---------------------------------------------------------------------------
// operator interfaces
// ---------------------------------------------------------------------------
class ILinearOperator {
   public:
    virtual ~ILinearOperator() = default;
    virtual void init() {}
    virtual void apply(GPUVector& f, const GPUVector& u) = 0;       // f = L*u
    virtual void apply_sub(GPUVector& f, const GPUVector& u) = 0;   // f -= L*u
};

class MatrixOperator : public ILinearOperator {
   public:
    explicit MatrixOperator(std::shared_ptr<GPUSparseMatrix> A) : A_(std::move(A)) {}
    void apply(GPUVector& f, const GPUVector& u) override { A_->apply(f, u); }
    void apply_sub(GPUVector& f, const GPUVector& u) override { A_->matmul_minus(f, u); }
    GPUSparseMatrix& get_matrix() { return *A_; }
    std::shared_ptr<GPUSparseMatrix> matrix_ptr() { return A_; }

   private:
    std::shared_ptr<GPUSparseMatrix> A_;
};

class ILinearIterator {
   public:
    virtual ~ILinearIterator() = default;
    virtual const char* name() const = 0;
    virtual bool init(std::shared_ptr<MatrixOperator> J) = 0;
    virtual bool apply(GPUVector& c, const GPUVector& d) = 0;    // c = B*d
    virtual bool apply_update_defect(GPUVector& c, GPUVector& d);  // c = B*d; d -= A*c
   protected:
    std::shared_ptr<MatrixOperator> op_;
};

class StdConvCheck {
   public:
    StdConvCheck(int maxSteps = 100, double minDefect = 1e-50, double relReduction = 1e-10, bool verbose = false)
        : max_steps_(maxSteps), min_defect_(minDefect), rel_reduction_(relReduction), verbose_(verbose) {}
    void set_maximum_steps(int s) { max_steps_ = s; }
    void set_minimum_defect(double d) { min_defect_ = d; }
    void set_reduction(double r) { rel_reduction_ = r; }
    void set_verbose(bool v) { verbose_ = v; }
    void start_defect(double d);
    void start(const GPUVector& d) { start_defect(d.norm()); }
    void update_defect(double d);
    void update(const GPUVector& d) { update_defect(d.norm()); }
    bool iteration_ended() const;
    bool post() const;
    int step() const { return step_; }
    double defect() const { return current_; }
    double reduction() const { return current_ / initial_; }
    const std::vector<double>& defects() const { return defects_; }
    void set_name(const std::string& n) { name_ = n; }

   private:
    int max_steps_;
    double min_defect_, rel_reduction_;
    bool verbose_;
    double initial_ = 0, current_ = 0, last_ = 0;
    int step_ = 0;
    std::vector<double> defects_;
    std::string name_ = "Iteration";
};

class Jacobi : public ILinearIterator {
   public:
    explicit Jacobi(double omega = 1.0) : omega_(omega) {}
    const char* name() const override { return "Jacobi"; }
    bool init(std::shared_ptr<MatrixOperator> J) override;
    bool apply(GPUVector& c, const GPUVector& d) override;
    void set_damp(double w) { omega_ = w; }

   private:
    double omega_;
    DeviceBuffer<double> dinv_;
};

class AssembledMultiGridCycle;

class ILinearOperatorInverse {
   public:
    virtual ~ILinearOperatorInverse() = default;
    virtual const char* name() const = 0;
    virtual bool init(std::shared_ptr<MatrixOperator> J);
    virtual bool apply(GPUVector& x, const GPUVector& b);
    virtual bool apply_return_defect(GPUVector& x, GPUVector& b) = 0;
    void set_preconditioner(std::shared_ptr<ILinearIterator> p) { precond_ = std::move(p); }
    void set_convergence_check(std::shared_ptr<StdConvCheck> c) { conv_ = std::move(c); }
    StdConvCheck& convergence_check() { return *conv_; }
    int step() const { return conv_->step(); }

   protected:
    std::shared_ptr<MatrixOperator> op_;
    std::shared_ptr<ILinearIterator> precond_;
    std::shared_ptr<StdConvCheck> conv_ = std::make_shared<StdConvCheck>();
};

class CG : public ILinearOperatorInverse {
   public:
    const char* name() const override { return "CG"; }
    bool apply(GPUVector& x, const GPUVector& b) override;
    bool apply_return_defect(GPUVector& x, GPUVector& b) override;
    // when on (default) and the preconditioner is a GMG initialised with this
    // operator, the iteration runs on lattice-ordered copies of the vectors;
    // with x and b zero on the Dirichlet DoFs the initial defect is formed
    // there too (r = b - A x on the lattice operator)
    void set_lattice_fast_path(bool on) { fast_ = on; }
    bool last_used_lattice_path() const { return used_fast_; }

   private:
    int lattice_start_kind(const GPUVector& x, const GPUVector& b, int64_t ni);
    // kind 0: src is the defect; 1: src = b and x = 0; 2: src = b, r = b - A x;
    // the final defect is written to defect_out when given
    bool apply_lattice(GPUVector& x, const GPUVector& src, AssembledMultiGridCycle& gmg, int kind,
                       GPUVector* defect_out);
    bool fast_ = true, used_fast_ = false;
    // lattice-path work space, kept across solves (halos stay zero)
    int64_t ws_n_ = -1, ws_h_ = -1;
    DeviceBuffer<double> ws_x_, ws_r_, ws_p_, ws_q_, ws_z_, ws_partial_, ws_sc_;
    DeviceBuffer<float> ws_rf_;   // FP32 copy of r for an FP32 cycle (with halos)
    DeviceBuffer<double> ws_p2_;  // second direction buffer (fused direction update)
};

class BiCGStab : public ILinearOperatorInverse {
   public:
    const char* name() const override { return "BiCGStab"; }
    bool apply_return_defect(GPUVector& x, GPUVector& b) override;
    // restart every n steps (0: never) and whenever |(r0, r)| <=
    // minOrtho * ||r|| ||r0|| (bicgstab.h:386-390)
    void set_restart(int n) {
        if (n < 0) throw UGError("BiCGStab::set_restart: n >= 0");
        num_restarts_ = n;
    }
    void set_min_orthogonality(double m) { min_ortho_ = m; }

   private:
    int num_restarts_ = 0;
    double min_ortho_ = 0.0;
};

// Element-loop assembly from caller-supplied connectivity: the element loop of
// DomainDiscretization::assemble_linear / assemble_jacobian / assemble_defect
// (domain_disc_impl.h:469-948; elem_disc_assemble_util.h:326-413, 577-687,
// 902-1003) over one element type, with the element -> DoF table that
// DoFDistribution::indices gives per element (dof_distribution.cpp:1117-1206)
// and the DoF coordinates. Element types: P1 triangles with FV1 (box method),
// P1 tetrahedra with FV1 and optional full-upwind convection, Q1 hexahedra with
// FE (2x2x2 Gauss); -div(k grad u) (+ div(v u)) = f, k per element, f constant
// or per DoF (hexahedra: constant), homogeneous Dirichlet values on the listed
// DoFs. The CSR pattern (every element coupling, rows sorted), the CSR slot of
// every local entry and an element colouring are built once on the device;
// every assembly is one launch per colour without atomics (elem_assembly.hip).
class ElementDiscretization {
   public:
    enum ElemType { TRIANGLE_FV1 = 0, TETRAHEDRON_FV1 = 1, HEXAHEDRON_Q1 = 2, HEXAHEDRON_Q2 = 3 };
    static int vertices_of(int elem_type) {
        return elem_type == 0 ? 3 : (elem_type == 1 ? 4 : (elem_type == 2 ? 8 : 27));
    }
    static int dim_of(int elem_type) { return elem_type == 0 ? 2 : 3; }
    // elem_dofs: num_elems x vertices_of(type) DoFs in the reference element's
    // local vertex order (Q2: the 27 shapes of LagrangeLSFS<Hexahedron, 2> --
    // vertices, edge midpoints, face centres, centre; the first 8 DoFs'
    // coordinates are the element's corners); dof_coords: num_dofs x
    // dim_of(type). on_device: the two arrays are device pointers (copied),
    // otherwise host arrays.
    ElementDiscretization(int elem_type, int64_t num_elems, const int32_t* elem_dofs, int64_t num_dofs,
                          const double* dof_coords, bool on_device = false);
    int elem_type() const { return type_; }
    int64_t num_elems() const { return ne_; }
    int64_t num_dofs() const { return ndofs_; }
    int64_t nnz() const { return nnz_; }
    int num_colours() const { return (int)colour_start_.size() - 1; }
    // diffusion k per element (host or device array of num_elems; nullptr: 1)
    void set_diffusion(const double* k, bool on_device = false);
    // Q2: k at the 3x3x3 Gauss points of every element (num_elems x 27, x
    // fastest), as the reference evaluates a coefficient function there;
    // replaces set_diffusion (nullptr: back to it)
    void set_diffusion_at_points(const double* kq, bool on_device = false);
    void set_source(double f) { f_const_ = f; has_f_dof_ = false; }
    void set_source_dofs(const double* f_host);          // FV1 only
    void set_velocity(double vx, double vy, double vz);  // tetrahedra only
    // homogeneous Dirichlet DoFs (value 0): rows become identity rows
    // (SetDirichletRow), rhs and defect entries 0
    void set_dirichlet(const int32_t* dofs_host, int64_t n);
    void set_dirichlet_device(const int32_t* dofs_dev, int64_t n);
    // A and b (A is given this discretisation's pattern on its first use)
    void assemble_linear(GPUSparseMatrix& A, GPUVector& b);
    // J = dd/du; the element discretisations here are linear in u, so J is the
    // stiffness matrix of assemble_linear
    void assemble_jacobian(GPUSparseMatrix& J, const GPUVector& u);
    // d = sum_e (A_e u_e - f_e), Dirichlet entries 0
    void assemble_defect(GPUVector& d, const GPUVector& u);
    // a copy of the pattern (rowptr ndofs + 1, col nnz), host arrays
    void download_pattern(int32_t* rowptr, int32_t* col) const;

   private:
    void build();
    void run(double* val, double* rhs, const double* u, double* d);
    // hands the pattern to A (true: A still held it from the last assembly)
    bool give_pattern(GPUSparseMatrix& A);
    void finish_matrix(GPUSparseMatrix& A, bool kept);
    uint64_t asm_pattern_id_ = 0;
    int type_, nv_, dim_;
    int64_t ne_, ndofs_, nnz_ = 0;
    DeviceBuffer<int32_t> dofs_, rowptr_, col_, order_, slot_, dir_;
    DeviceBuffer<double> coords_, kel_, kq_, f_dof_;
    std::vector<int64_t> colour_start_;
    double f_const_ = 0.0;
    bool has_k_ = false, has_kq_ = false, has_f_dof_ = false;
    double vel_[3] = {0.0, 0.0, 0.0};
    bool conv_ = false;
    int64_t ndir_ = 0;
};

// ---------------------------------------------------------------------------
// structured discretisation + multigrid
// ---------------------------------------------------------------------------
struct LevelGrid {
    int dim = 3;
    int n = 0;            // cells per side
    int W = 0;            // points per side (n+1)
    int64_t N = 0;        // points (DoFs)
    int64_t n_inner = 0;  // DoFs [0, n_inner) are not Dirichlet
    DeviceBuffer<int32_t> dof_of_lat, lat_of_dof;
    DeviceBuffer<uint8_t> dir;   // Dirichlet flag per lattice point
    // kind 2: 6 packed tetrahedra per cell; parent-edge code per lattice
    // point (levels >= 1)
    DeviceBuffer<uint16_t> cell_tets;
    DeviceBuffer<uint8_t> pcode;
};

// Refines the lattice hierarchy of domain kind `kind` num_refs times on the
// device and fills each level's size, DoF numbering (ordering 0: the
// reference's LEVEL numbering, 1: lexicographic), Inner count and, for the
// tetrahedral domain, the per-cell tetrahedra and parent-edge codes
// (mesh_kernels.hip).
void build_device_levels(int kind, int num_refs, int ordering, std::vector<std::unique_ptr<LevelGrid>>& levels);

// Domain of lattice topology: kind 0 = unit cube from one hexahedron (Q1 FE,
// -div(k grad u) = f), kind 1 = unit square from two triangles (FV1),
// kind 2 = unit cube as 3x3x3 hexahedra x 6 tetrahedra, regularly refined
// (FV1 convection-diffusion -div(k grad u) + div(v u) = f, full upwind).
class DomainDiscretization {
   public:
    // ordering 0: the reference's LEVEL numbering (lib_grid creation order,
    // subset-major); ordering 1: lexicographic lattice order with the
    // Dirichlet DoFs last (an OrderLex-style renumbering,
    // lib_disc/ordering_strategies)
    // setup 0: the hierarchy and numbering are built on the device
    // (mesh_kernels.hip); setup 1: on the host (hierarchy.cpp), then uploaded.
    // Both give the same numbering.
    DomainDiscretization(int kind, int num_refs, int ordering = 0, int setup = 0);
    int kind() const { return kind_; }
    int dim() const { return kind_ == 1 ? 2 : 3; }
    // level operators symmetric on vectors vanishing on Dirichlet DoFs
    bool symmetric() const { return kind_ != 2; }
    // fine elements (kind 0/1) or fine lattice cells (kind 2) per side
    int fine_cells() const { return (kind_ == 2 ? 3 : 1) << refs_; }
    int num_levels() const { return (int)levels_.size(); }
    int top_level() const { return num_levels() - 1; }
    LevelGrid& level(int l) { return *levels_[l]; }
    // coefficient: theta per fine element (lattice order), k = exp(theta); empty = 1
    void set_theta(const double* theta_host);
    const double* theta_device() const { return has_theta_ ? theta_.get() : nullptr; }
    void set_source_constant(double f) { f_const_ = f; }
    void set_source_dofs(const double* f_host);   // FV1: per top-level DoF
    // assemble the top-level (surface) CSR in DoF numbering and the rhs
    void assemble_linear(GPUSparseMatrix& A, GPUVector& b);
    // J = dd/du (domain_disc_impl.h:469-600): the discretisations here are
    // linear in u, so J is the matrix of assemble_linear (Dirichlet rows set)
    void assemble_jacobian(GPUSparseMatrix& J, const GPUVector& u);
    // d = sum_e (A_e u_e - f_e) with the Dirichlet entries 0
    // (domain_disc_impl.h:646-788), by the element loop over the top level's
    // connectivity (ElementDiscretization built on the device)
    void assemble_defect(GPUVector& d, const GPUVector& u);
    // the element-loop discretisation of the top level (built on first use;
    // rebuilt when theta, the source, velocity or jitter change)
    ElementDiscretization& element_discretization();
    // assemble a level operator in lattice numbering as DIA values (27 or 9
    // diagonals of length N, FP64), Dirichlet rows set
    void assemble_level_dia(int lev, DeviceBuffer<double>& dia, int& nd, std::vector<int32_t>& offsets);
    // coordinates of the level's DoFs in DoF order (2 or 3 per DoF), as the
    // grid function's approximation space would give them to
    // SaveMatrixForConnectionViewer (grid_function_util.h)
    void dof_positions(int lev, double* host_out);
    // the diagonal count and lattice offsets of a level's DIA storage
    void level_dia_layout(int lev, int& nd, std::vector<int32_t>& offsets) const;
    double assemble_seconds() const { return t_assemble_; }
    // theta from device memory (fine-element lattice order)
    void set_theta_device(const double* theta_dev);
    // DirichletBoundary::adjust_defect: zero the Dirichlet entries of a
    // top-level vector (lagrange_dirichlet_boundary_impl.h:920-1035)
    void adjust_defect(GPUVector& d);
    // adjoint gradient g_e = -lambda_e^T (dA_e/dtheta_e) u_e per fine element
    // (kind 0) or fine cell (kind 2); grad_dev holds fine_cells()^3 doubles
    void adjoint_gradient(const GPUVector& u, const GPUVector& lambda, double* grad_dev);
    // kind 2: convection velocity; jitter of the top-level interior vertices
    // by amp * h * (2U - 1) per component (U uniform from the seeded counter
    // generator, counter 3 * lattice index + component)
    void set_velocity(double vx, double vy, double vz);
    void set_jitter(uint64_t seed, double amp);
    // operator of the adjoint solve: AT = (A with the Dirichlet columns of its
    // interior rows zeroed)^T, identity Dirichlet rows, on A's pattern (which
    // must be structurally symmetric); A from assemble_linear
    void assemble_adjoint_operator(const GPUSparseMatrix& A, GPUSparseMatrix& AT);
    const double* velocity() const { return vel_; }
    uint64_t jitter_seed() const { return jseed_; }
    double jitter_amp() const { return jamp_; }
    double source_constant() const { return f_const_; }
    int num_refs() const { return refs_; }

   private:
    void build_host_levels(int kind, int num_refs, int ordering);
    int dim_of_kind() const { return kind_ == 1 ? 2 : 3; }
    // the element-loop discretisation of the top level and the jitter its
    // coordinates were made with
    std::unique_ptr<ElementDiscretization> elem_;
    uint64_t elem_jseed_ = 0;
    double elem_jamp_ = -1.0;
    int kind_;
    int refs_;
    std::vector<std::unique_ptr<LevelGrid>> levels_;
    DeviceBuffer<double> theta_, f_dof_;
    bool has_theta_ = false, has_f_ = false;
    double f_const_ = 1.0;
    double vel_[3] = {0.0, 0.0, 0.0};
    uint64_t jseed_ = 0;
    double jamp_ = 0.0;
    double t_assemble_ = 0.0;
    // pattern id of the matrix this discretisation last assembled: a matrix
    // that still holds it keeps its pattern (and SpMV layout) on re-assembly
    uint64_t asm_pattern_id_ = 0;
    // Q1 unit-cube tables of the row-wise lattice assembly and exp(theta) per
    // fine element (recomputed after the next set_theta)
    DeviceBuffer<double> q1_tab_, kel_;
    bool kel_valid_ = false;
    // row start of every top-level lattice point in the lattice pattern's CSR
    // (kinds 0, 1: the pattern depends on the numbering alone)
    DeviceBuffer<int32_t> rp_lat_;
    const double* exp_theta_device();
    const double* q1_tables_device();
};

// Q2 Lagrange on the unit cube split into one hexahedron and refined
// num_refs times: -div(k grad u) = f, u = 0 on the boundary, k = exp(theta)
// per top-level element (evaluated at the Gauss points of every level's
// elements, as the reference evaluates a coefficient function), f constant.
// Every level has the reference's LEVEL DoF order (vertices, edges, faces,
// volumes; subset by subset in creation order), its operator is assembled by
// the element loop (ElementDiscretization, Q2 element matrices on the matrix
// cores) and the transfers are StdTransfer's (q2_hierarchy.hip). With
// MatrixMultiGrid they give the Q2 geometric multigrid.
class Q2Discretization {
   public:
    explicit Q2Discretization(int num_refs);
    int num_levels() const { return (int)L_.size(); }
    int64_t num_dofs(int lev) const { return L_.at(lev).N; }
    void set_theta(const double* theta_host);   // (2^num_refs)^3 values, x fastest; nullptr: k = 1
    void set_source(double f) { f_ = f; }
    // DoF coordinates (N x 3) and Dirichlet flags (N) of a level, host arrays
    void dof_positions(int lev, double* out) const;
    void dirichlet_flags(int lev, uint8_t* out) const;
    void assemble_linear(int lev, GPUSparseMatrix& A, GPUVector& b);
    void prolongation(int lev, GPUSparseMatrix& P) const;                        // lev-1 -> lev
    void restriction(int lev, const GPUSparseMatrix& P, GPUSparseMatrix& R) const;  // P of the same level

   private:
    struct Level {
        int n = 0;
        int64_t N = 0;
        DeviceBuffer<int32_t> dof_of_node, node_of_dof;
        DeviceBuffer<uint8_t> dir;
        std::unique_ptr<ElementDiscretization> elem;
        bool coef_valid = false;
    };
    std::vector<Level> L_;
    int refs_ = 0;
    double f_ = 0.0;
    DeviceBuffer<double> theta_;
    bool has_theta_ = false;
};

// Geometric multigrid over level operators and grid transfers handed in as
// matrices: the cycle of AssembledMultiGridCycle (mg_solver_impl.hpp:175-275
// apply, 1686-1816 presmooth + restriction, 1819-1964 prolongation +
// postsmooth, 1968-2086 base solve, 2090-2136 lmgc) with damped Jacobi
// smoothers and the dense LU on the base level. Level l = 0 .. top: A[l] the
// level operator, P[l] (l >= 1) the prolongation from l-1 to l (rows on level
// l), R[l] the restriction from l to l-1 (StdTransfer's matrices,
// std_transfer_impl.h:602-806, Dirichlet rows already adjusted). The top level
// uses the operator passed to init (the solver's), as the reference copies the
// surface operator to its top level. FP64 throughout; every matrix goes
// through the SpMV layouts of GPUSparseMatrix.
class MatrixMultiGrid : public ILinearIterator {
   public:
    enum CycleType { _F_ = -1, _V_ = 1, _W_ = 2 };
    MatrixMultiGrid() = default;
    const char* name() const override { return "GMG"; }
    void set_levels(std::vector<std::shared_ptr<GPUSparseMatrix>> A, std::vector<std::shared_ptr<GPUSparseMatrix>> P,
                    std::vector<std::shared_ptr<GPUSparseMatrix>> R);
    void set_base_level(int l) { base_ = l; }
    void set_num_presmooth(int n) { nu1_ = n; }
    void set_num_postsmooth(int n) { nu2_ = n; }
    void set_cycle_type(int t) { cycle_ = t; }
    void set_smoother_damp(double w) { omega_ = w; }
    bool init(std::shared_ptr<MatrixOperator> J) override;
    bool apply(GPUVector& c, const GPUVector& d) override;

   private:
    void lmgc(int lev, int type);
    void presmooth_and_restriction(int lev);
    void prolongation_and_postsmooth(int lev);
    void base_solve(int lev);
    GPUSparseMatrix& Aop(int lev) { return lev == top_ ? op_->get_matrix() : *A_[lev]; }
    std::vector<std::shared_ptr<GPUSparseMatrix>> A_, P_, R_;
    int base_ = 0, top_ = -1, nu1_ = 2, nu2_ = 2, cycle_ = 1;
    double omega_ = 0.66;
    std::vector<std::unique_ptr<Jacobi>> smoother_;
    std::vector<GPUVector> sc_, sd_, st_;
    int lu_n_ = 0;
    DeviceBuffer<double> lu_linvT_, lu_uinvT_;
    DeviceBuffer<int32_t> lu_perm_;
};

enum class Precision { F64 = 0, F32 = 1 };

class AssembledMultiGridCycle : public ILinearIterator {
   public:
    explicit AssembledMultiGridCycle(std::shared_ptr<DomainDiscretization> disc);
    ~AssembledMultiGridCycle() override;
    const char* name() const override { return "GMG"; }
    // cycle types of AssembledMultiGridCycle (mg_solver.h:339-341)
    static const int _V_ = 1;
    static const int _W_ = 2;
    static const int _F_ = -1;
    void set_base_level(int l) {
        if (l < 0) throw UGError("GMG::set_base_level: level must be >= 0");
        base_lev_ = l;
    }
    void set_num_presmooth(int n) {
        if (n < 0) throw UGError("GMG::set_num_presmooth: number of steps must be >= 0");
        nu1_ = n;
    }
    void set_num_postsmooth(int n) {
        if (n < 0) throw UGError("GMG::set_num_postsmooth: number of steps must be >= 0");
        nu2_ = n;
    }
    // _V_ (1), _W_ (2), any gamma >= 1 (gamma coarse cycles per level), or
    // _F_ (-1: one F-cycle then one V-cycle on the next coarser level,
    // mg_solver_impl.hpp:2115-2118); 0 and other negative values are refused
    // (they would run no coarse correction at all)
    void set_cycle_type(int c) {
        if (c != _F_ && c < 1) throw UGError("GMG::set_cycle_type: cycle must be 1 (V), 2 (W), gamma > 2 or -1 (F)");
        cycle_ = c;
    }
    int cycle_type() const { return cycle_; }
    void set_smoother_damp(double w) { omega_ = w; }
    void set_precision(Precision p) { prec_ = p; }
    // precondition the adjoint operator: level operators A_l,elim^T (the
    // transposed Dirichlet-eliminated level matrices), transfer unchanged
    void set_transposed(bool t) { transposed_ = t; }
    bool transposed() const { return transposed_; }
    // apply the top-level operator of the hexahedral discretisation from its
    // per-element coefficients (matrix-free) instead of the stored diagonals
    // (default on; the stored level matrices are kept either way)
    void set_matrix_free(bool on) { matrix_free_ = on; }
    // Galerkin coarse operators A_{l-1} = R_l A_l P_l from the assembled top
    // level instead of the assembled level matrices (set_rap, mg_solver.h:149;
    // init_rap_operator, mg_solver_impl.hpp:829-1013)
    void set_rap(bool on) { rap_ = on; }
    bool rap() const { return rap_; }
    // replay the lattice-order cycle (the CG fast path) from captured HIP
    // graphs (default) or launch its kernels one by one; same kernels, same
    // results. Before init.
    void set_graphs(bool on) { graphs_ = on; }
    // matrix-free top level of the lattice entry: run the last full
    // post-smoothing step together with the correction update that ends the
    // cycle and its conversion (default) or as two passes; the same
    // correction, the (z, r) product summed in another order. Before init.
    void set_fused_smoothing(bool on) { fused_smoothing_ = on; }
    bool init(std::shared_ptr<MatrixOperator> J) override;
    bool apply(GPUVector& c, const GPUVector& d) override;
    // testing access: level matrix (FP64 DIA values in lattice numbering)
    void download_level_dia(int lev, double* out) const;
    int level_num_diags(int lev) const;
    // nsteps fused smoothing steps on level `lev` (timing / inspection)
    void smooth_level(int lev, int nsteps);
    // Lattice-order interface used by the CG fast path. Valid while A is the
    // matrix this cycle was initialised with (same assembly). Lattice vectors
    // are FP64 buffers of N entries with `lattice_halo()` zero entries on
    // each side (allocated by the caller).
    bool lattice_ready(const GPUSparseMatrix& A) const;
    int64_t lattice_size() const;
    int64_t lattice_halo() const;
    const int32_t* dof_of_lattice() const;
    const int32_t* lattice_of_dof() const;
    int64_t num_inner_dofs() const;
    // c = B d for a defect d in lattice order that vanishes on Dirichlet DoFs;
    // d carries the halo (the first smoothing step stages its planes from it);
    // with cd_out (device), also (c, d) into cd_out[0] (partial: device
    // scratch of max_partials() doubles). d_lat_f32 (optional, with the halo):
    // (float) d, read instead of d by the first step when
    // lattice_uses_f32_defect(); same results
    void apply_lattice(double* c_lat, const double* d_lat, double* cd_out = nullptr, double* partial = nullptr,
                       const float* d_lat_f32 = nullptr);
    bool lattice_uses_f32_defect() const;
    // y = alpha*v + beta*A*x with the surface operator in lattice order
    // (x must vanish on Dirichlet DoFs and carry the halo); with yx_out
    // (device) also (y, x) into yx_out[0] (partial: device scratch of
    // max_partials() doubles), false if that fusion is not available
    // p_new = z + beta p_old (beta: device scalar, the arithmetic of the CG's
    // direction update) and q = A p_new with (q, p_new) into qp_out[0] in one
    // pass; p_new and p_old are distinct lattice vectors with the halo; false
    // when the top level is not matrix-free
    bool surface_direction_axpy_lattice(double* q, const double* p_old, const double* z, const double* beta,
                                        double* p_new, double* qp_out, double* partial);
    bool surface_axpy_lattice(double* y, double alpha, const double* v, double beta, const double* x,
                              double* yx_out = nullptr, double* partial = nullptr);
    static constexpr int max_partials() { return 65536; }
    struct Impl;

   private:
    std::shared_ptr<DomainDiscretization> disc_;
    int base_lev_ = 0, nu1_ = 2, nu2_ = 2, cycle_ = 1;
    double omega_ = 0.66;
    Precision prec_ = Precision::F64;
    bool transposed_ = false;
    bool matrix_free_ = true;
    bool rap_ = false;
    bool graphs_ = true;
    bool fused_smoothing_ = true;
    Impl* impl_ = nullptr;
    const GPUSparseMatrix* mat_ = nullptr;
    uint64_t mat_version_ = 0;
};

// ---------------------------------------------------------------------------
// file formats (io.cpp): ConnectionViewer text output (connection_viewer_
// output.h:74-141, 402-422) and the serial SaveToFile / ReadFromFile vector
// file (restart_bridge.cpp:49-75, 146-158; vector_impl.h:347-361)
// ---------------------------------------------------------------------------
void cv_write_matrix(const std::string& path, int64_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                     const double* positions, int dim);
void cv_write_vector(const std::string& path, int64_t n, const double* values, const double* positions, int dim);
void vector_file_write(const std::string& path, int64_t n, const double* values);
int64_t vector_file_length(const std::string& path);
void vector_file_read(const std::string& path, int64_t n, double* values);
void SaveToFile(const GPUVector& v, const std::string& filename);
void ReadFromFile(GPUVector& v, const std::string& filename);

}  // namespace gpu
}  // namespace ug
