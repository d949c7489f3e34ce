// ug4py: the pybind11 bindings of the GPU1 algebra (the module the reference's
// USE_PYBIND11 option builds, cmake/ug/pybind11.cmake:33-55, has no sources in
// the reference). Host code over the C ABI of libugcore_amd.so only; names
// follow the Lua registry (DomainDiscretization, GeometricMultiGrid, CG,
// BiCGStab, Jacobi, SparseMatrix, Vector, ElementDiscretization).
//
// Device memory of other frameworks is exchanged by address: Vector.copy_from /
// copy_to take a device pointer (a torch tensor's data_ptr()), and
// DomainDiscretization.set_theta_device / adjoint_gradient read and write
// device pointers. Those addresses are valid here only when this library and
// the framework share one HIP runtime (ugcore_amd.lib() arranges that when
// torch is installed).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>

#include "ugcore_amd.h"

namespace py = pybind11;

namespace {

struct UGError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(int rc, const char* what) {
    if (rc < 0) throw UGError(std::string(what) + ": " + ugx_last_error());
}
void* check_ptr(void* p, const char* what) {
    if (!p) throw UGError(std::string(what) + ": " + ugx_last_error());
    return p;
}

typedef py::array_t<double, py::array::c_style | py::array::forcecast> DArray;
typedef py::array_t<int32_t, py::array::c_style | py::array::forcecast> IArray;

// owning wrappers of the C ABI handles
struct Vector {
    void* h;
    explicit Vector(int64_t n) : h(check_ptr(ugx_vector_create(n), "Vector")) { check(ugx_vector_set(h, 0.0), "Vector"); }
    ~Vector() { ugx_vector_destroy(h); }
    Vector(const Vector&) = delete;
    Vector& operator=(const Vector&) = delete;
    int64_t size() const { return ugx_vector_size(h); }
    uintptr_t device_ptr() const { return (uintptr_t)ugx_vector_device_ptr(h); }
    void set(double v) { check(ugx_vector_set(h, v), "Vector.set"); }
    void upload(DArray a) {
        if (a.size() != size()) throw py::value_error("Vector.upload: size mismatch");
        check(ugx_vector_upload(h, a.data()), "Vector.upload");
    }
    DArray numpy() const {
        DArray a(size());
        check(ugx_vector_download(h, a.mutable_data()), "Vector.numpy");
        return a;
    }
    double norm() const {
        double r = 0.0;
        check(ugx_norm_f64(size(), ugx_vector_device_ptr(h), &r, nullptr), "Vector.norm");
        return r;
    }
    double dotprod(const Vector& o) const {
        if (o.size() != size()) throw py::value_error("Vector.dotprod: size mismatch");
        double r = 0.0;
        check(ugx_dot_f64(size(), ugx_vector_device_ptr(h), ugx_vector_device_ptr(o.h), &r, nullptr),
              "Vector.dotprod");
        return r;
    }
    // device-to-device copies with another framework's memory (synchronous)
    void copy_from(uintptr_t src) {
        check(ugx_memcpy_dtod(ugx_vector_device_ptr(h), (const void*)src, 8 * size()), "Vector.copy_from");
    }
    void copy_to(uintptr_t dst) const {
        check(ugx_memcpy_dtod((void*)dst, ugx_vector_device_ptr(h), 8 * size()), "Vector.copy_to");
    }
};

struct SparseMatrix {
    void* h;
    SparseMatrix() : h(check_ptr(ugx_matrix_create(), "SparseMatrix")) {}
    ~SparseMatrix() { ugx_matrix_destroy(h); }
    SparseMatrix(const SparseMatrix&) = delete;
    SparseMatrix& operator=(const SparseMatrix&) = delete;
    int64_t num_rows() const { return ugx_matrix_num_rows(h); }
    int64_t num_cols() const { return ugx_matrix_num_cols(h); }
    int64_t nnz() const { return ugx_matrix_nnz(h); }
    py::tuple download_csr() const {
        IArray rp(num_rows() + 1), col(nnz());
        DArray val(nnz());
        check(ugx_matrix_download_csr(h, rp.mutable_data(), col.mutable_data(), val.mutable_data()),
              "SparseMatrix.download_csr");
        return py::make_tuple(rp, col, val);
    }
    void upload_csr(IArray rp, IArray col, DArray val, int64_t ncols) {
        const int64_t rows = (int64_t)rp.size() - 1;
        if (rows < 0 || col.size() != val.size() || rp.data()[rows] != (int32_t)col.size())
            throw py::value_error("SparseMatrix.upload_csr: inconsistent CSR");
        check(ugx_matrix_upload_csr(h, rows, ncols < 0 ? rows : ncols, rp.data(), col.data(), val.data()),
              "SparseMatrix.upload_csr");
    }
    void axpy(Vector& dest, double a, Vector& v1, double b, Vector& w) {
        check(ugx_matrix_axpy(h, dest.h, a, v1.h, b, w.h), "SparseMatrix.axpy");
    }
    void apply(Vector& dest, Vector& w) { axpy(dest, 0.0, dest, 1.0, w); }
    void apply_sub(Vector& d, Vector& x) { check(ugx_matrix_apply_sub(h, d.h, x.h), "SparseMatrix.apply_sub"); }
    void axpy_transposed(Vector& dest, double a, Vector& v1, double b, Vector& w) {
        check(ugx_matrix_axpy_transposed(h, dest.h, a, v1.h, b, w.h), "SparseMatrix.axpy_transposed");
    }
    void resize_and_clear(int64_t r, int64_t c) { check(ugx_matrix_resize_and_clear(h, r, c), "resize_and_clear"); }
    void add(int64_t r, int64_t c, double v) {
        double* p = ugx_matrix_entry(h, r, c);
        if (!p) throw UGError(std::string("SparseMatrix.add: ") + ugx_last_error());
        *p += v;
    }
    double get(int64_t r, int64_t c) const {
        double v = 0.0;
        check(ugx_matrix_get(h, r, c, &v), "SparseMatrix.get");
        return v;
    }
    void finalize() { check(ugx_matrix_finalize(h), "SparseMatrix.finalize"); }
    void set_as_transpose_of(const SparseMatrix& B, double scale) {
        check(ugx_matrix_set_as_transpose_of(h, B.h, scale, 1), "set_as_transpose_of");
    }
};

struct DomainDiscretization {
    void* h;
    int kind;
    DomainDiscretization(int kind_, int refs) : h(check_ptr(ugx_disc_create(kind_, refs), "DomainDiscretization")),
                                                kind(kind_) {}
    ~DomainDiscretization() { ugx_disc_destroy(h); }
    DomainDiscretization(const DomainDiscretization&) = delete;
    DomainDiscretization& operator=(const DomainDiscretization&) = delete;
    int num_levels() const { return ugx_disc_num_levels(h); }
    int64_t num_dofs() const { return ugx_disc_num_dofs(h, num_levels() - 1); }
    int64_t num_elements() const {
        int64_t n = int64_t(1) << (num_levels() - 1);
        if (kind == 2) n *= 3;
        return kind == 1 ? n * n : n * n * n;
    }
    void set_theta(DArray t) {
        if (t.size() != num_elements()) throw py::value_error("set_theta: one value per fine element");
        check(ugx_disc_set_theta(h, t.data()), "set_theta");
    }
    void set_theta_device(uintptr_t p) { check(ugx_disc_set_theta_device(h, (const double*)p), "set_theta_device"); }
    void set_source(double f) { check(ugx_disc_set_source_const(h, f), "set_source"); }
    void assemble_linear(SparseMatrix& A, Vector& b) { check(ugx_disc_assemble_linear(h, A.h, b.h), "assemble_linear"); }
    void assemble_jacobian(SparseMatrix& J, Vector& u) {
        check(ugx_disc_assemble_jacobian(h, J.h, u.h), "assemble_jacobian");
    }
    void assemble_defect(Vector& d, Vector& u) { check(ugx_disc_assemble_defect(h, d.h, u.h), "assemble_defect"); }
    void adjust_defect(Vector& d) { check(ugx_disc_adjust_defect(h, d.h), "adjust_defect"); }
    // g_e = -lambda_e^T (dA_e/dtheta_e) u_e into the device array at grad_ptr
    void adjoint_gradient(Vector& u, Vector& lam, uintptr_t grad_ptr) {
        check(ugx_disc_adjoint_gradient(h, u.h, lam.h, (double*)grad_ptr), "adjoint_gradient");
    }
    IArray dof_of_lattice() const {
        int64_t W = (int64_t(1) << (num_levels() - 1)) * (kind == 2 ? 3 : 1) + 1;
        IArray a(kind == 1 ? W * W : W * W * W);
        check(ugx_disc_dof_of_lattice(h, num_levels() - 1, a.mutable_data()), "dof_of_lattice");
        return a;
    }
};

struct GeometricMultiGrid {
    void* h;
    GeometricMultiGrid(DomainDiscretization& d, int base_level, int nu1, int nu2, int cycle, double omega,
                       const std::string& precision)
        : h(check_ptr(ugx_gmg_create(d.h, base_level, nu1, nu2, cycle, omega,
                                     precision == "f32" ? 1 : (precision == "f64" ? 0 : -1)),
                      "GeometricMultiGrid")) {}
    ~GeometricMultiGrid() { ugx_gmg_destroy(h); }
    GeometricMultiGrid(const GeometricMultiGrid&) = delete;
    GeometricMultiGrid& operator=(const GeometricMultiGrid&) = delete;
    void init(SparseMatrix& A) { check(ugx_gmg_init(h, A.h), "GeometricMultiGrid.init"); }
    void apply(Vector& c, Vector& d) { check(ugx_gmg_apply(h, c.h, d.h), "GeometricMultiGrid.apply"); }
};

// Q2 hierarchy of the unit cube (ugx_q2disc_*)
struct Q2Discretization {
    void* h;
    int refs;
    explicit Q2Discretization(int num_refs)
        : h(check_ptr(ugx_q2disc_create(num_refs), "Q2Discretization")), refs(num_refs) {}
    ~Q2Discretization() { ugx_q2disc_destroy(h); }
    Q2Discretization(const Q2Discretization&) = delete;
    Q2Discretization& operator=(const Q2Discretization&) = delete;
    int num_levels() const { return ugx_q2disc_num_levels(h); }
    int64_t num_dofs(int level) const { return ugx_q2disc_num_dofs(h, level); }
    void set_theta(DArray theta) {
        if (theta.size() != (int64_t(1) << (3 * refs))) throw py::value_error("one theta per top-level element");
        check(ugx_q2disc_set_theta(h, theta.data()), "set_theta");
    }
    void set_source(double f) { check(ugx_q2disc_set_source(h, f), "set_source"); }
    void assemble_linear(int level, SparseMatrix& A, Vector& b) {
        check(ugx_q2disc_assemble_linear(h, level, A.h, b.h), "assemble_linear");
    }
    void prolongation(int level, SparseMatrix& P) { check(ugx_q2disc_prolongation(h, level, P.h), "prolongation"); }
    void restriction(int level, SparseMatrix& P, SparseMatrix& R) {
        check(ugx_q2disc_restriction(h, level, P.h, R.h), "restriction");
    }
};

// AssembledMultiGridCycle over level matrices (ugx_mgmg_*); keeps the
// matrices' Python objects alive
struct MatrixMultiGrid {
    void* h;
    py::list keep;
    MatrixMultiGrid(py::list A, py::list P, py::list R, int base_level, int nu1, int nu2, int cycle, double omega) {
        const size_t n = A.size();
        if (P.size() != n || R.size() != n) throw py::value_error("one A, P and R entry per level");
        std::vector<void*> a(n), p(n), r(n);
        auto handle = [](py::handle o) -> void* { return o.is_none() ? nullptr : o.cast<SparseMatrix&>().h; };
        for (size_t l = 0; l < n; ++l) {
            a[l] = handle(A[l]);
            p[l] = handle(P[l]);
            r[l] = handle(R[l]);
        }
        h = check_ptr(ugx_mgmg_create((int)n, a.data(), p.data(), r.data(), base_level, nu1, nu2, cycle, omega),
                      "MatrixMultiGrid");
        keep.append(A);
        keep.append(P);
        keep.append(R);
    }
    ~MatrixMultiGrid() { ugx_mgmg_destroy(h); }
    MatrixMultiGrid(const MatrixMultiGrid&) = delete;
    MatrixMultiGrid& operator=(const MatrixMultiGrid&) = delete;
    void init(SparseMatrix& A) { check(ugx_mgmg_init(h, A.h), "MatrixMultiGrid.init"); }
    void apply(Vector& c, Vector& d) { check(ugx_mgmg_apply(h, c.h, d.h), "MatrixMultiGrid.apply"); }
};

struct Jacobi {
    void* h;
    explicit Jacobi(double omega) : h(check_ptr(ugx_jacobi_create(omega), "Jacobi")) {}
    ~Jacobi() { ugx_jacobi_destroy(h); }
    Jacobi(const Jacobi&) = delete;
    Jacobi& operator=(const Jacobi&) = delete;
};

struct Solver {
    void* h;
    py::object precond;   // keeps the preconditioner alive
    Solver(int type, py::object pre, int max_steps, double reduction, double min_defect)
        : h(check_ptr(ugx_solver_create(type), "Solver")), precond(pre) {
        if (!pre.is_none()) {
            if (py::isinstance<GeometricMultiGrid>(pre))
                check(ugx_solver_set_preconditioner(h, 1, pre.cast<GeometricMultiGrid&>().h), "set_preconditioner");
            else if (py::isinstance<Jacobi>(pre))
                check(ugx_solver_set_preconditioner(h, 2, pre.cast<Jacobi&>().h), "set_preconditioner");
            else if (py::isinstance<MatrixMultiGrid>(pre))
                check(ugx_solver_set_preconditioner(h, 3, pre.cast<MatrixMultiGrid&>().h), "set_preconditioner");
            else
                throw py::type_error("preconditioner: GeometricMultiGrid, MatrixMultiGrid, Jacobi or None");
        }
        check(ugx_solver_set_conv_check(h, max_steps, min_defect, reduction, 0), "set_conv_check");
    }
    ~Solver() { ugx_solver_destroy(h); }
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;
    void init(SparseMatrix& A) { check(ugx_solver_init(h, A.h), "init"); }
    bool apply(Vector& x, Vector& b) {
        const int rc = ugx_solver_apply(h, x.h, b.h);
        check(rc, "apply");
        return rc == 1;
    }
    int steps() const { return ugx_solver_steps(h); }
};

struct CG : Solver {
    CG(py::object pre, int max_steps, double reduction, double min_defect)
        : Solver(0, pre, max_steps, reduction, min_defect) {}
};
struct BiCGStab : Solver {
    BiCGStab(py::object pre, int max_steps, double reduction, double min_defect)
        : Solver(1, pre, max_steps, reduction, min_defect) {}
};

struct ElementDiscretization {
    void* h;
    int64_t ne, nd;
    ElementDiscretization(int type, IArray dofs, DArray coords) {
        const int nv = type == 0 ? 3 : (type == 1 ? 4 : (type == 2 ? 8 : 27)), dim = type == 0 ? 2 : 3;
        if (type < 0 || type > 3) throw py::value_error("element type 0, 1, 2 or 3");
        if (dofs.ndim() != 2 || dofs.shape(1) != nv || coords.ndim() != 2 || coords.shape(1) != dim)
            throw py::value_error("connectivity shapes do not match the element type");
        ne = dofs.shape(0);
        nd = coords.shape(0);
        h = check_ptr(ugx_elemdisc_create(type, ne, dofs.data(), nd, coords.data()), "ElementDiscretization");
    }
    ~ElementDiscretization() { ugx_elemdisc_destroy(h); }
    ElementDiscretization(const ElementDiscretization&) = delete;
    ElementDiscretization& operator=(const ElementDiscretization&) = delete;
    void set_diffusion(DArray k) {
        if (k.size() != ne) throw py::value_error("one diffusion value per element");
        check(ugx_elemdisc_set_diffusion(h, k.data()), "set_diffusion");
    }
    void set_diffusion_at_points(DArray kq) {
        if (kq.size() != 27 * ne) throw py::value_error("27 values per element");
        check(ugx_elemdisc_set_diffusion_at_points(h, kq.data()), "set_diffusion_at_points");
    }
    void set_source(double f) { check(ugx_elemdisc_set_source_const(h, f), "set_source"); }
    void set_dirichlet(IArray d) { check(ugx_elemdisc_set_dirichlet(h, d.data(), d.size()), "set_dirichlet"); }
    void assemble_linear(SparseMatrix& A, Vector& b) {
        check(ugx_elemdisc_assemble_linear(h, A.h, b.h), "assemble_linear");
    }
    void assemble_defect(Vector& d, Vector& u) { check(ugx_elemdisc_assemble_defect(h, d.h, u.h), "assemble_defect"); }
};

}  // namespace

PYBIND11_MODULE(ug4py, m) {
    m.doc() = "pybind11 bindings of the MI355X GPU1 algebra for ugcore (over libugcore_amd.so)";
    static py::exception<UGError> exc(m, "UGError");
    m.def("init", [] { check(ugx_init(), "init"); });
    m.def("synchronize", [] { check(ugx_synchronize(), "synchronize"); });
    m.def("stream", [] { return (uintptr_t)ugx_stream(); }, "the library's hipStream_t");

    py::class_<Vector>(m, "Vector")
        .def(py::init<int64_t>())
        .def("size", &Vector::size)
        .def("set", &Vector::set)
        .def("upload", &Vector::upload)
        .def("numpy", &Vector::numpy)
        .def("norm", &Vector::norm)
        .def("dotprod", &Vector::dotprod)
        .def("device_ptr", &Vector::device_ptr)
        .def("copy_from", &Vector::copy_from, "device-to-device copy from a device address")
        .def("copy_to", &Vector::copy_to, "device-to-device copy to a device address");
    py::class_<SparseMatrix>(m, "SparseMatrix")
        .def(py::init<>())
        .def("num_rows", &SparseMatrix::num_rows)
        .def("num_cols", &SparseMatrix::num_cols)
        .def("nnz", &SparseMatrix::nnz)
        .def("download_csr", &SparseMatrix::download_csr)
        .def("upload_csr", &SparseMatrix::upload_csr, py::arg("rowptr"), py::arg("col"), py::arg("val"),
             py::arg("num_cols") = -1)
        .def("axpy", &SparseMatrix::axpy)
        .def("apply", &SparseMatrix::apply)
        .def("apply_sub", &SparseMatrix::apply_sub)
        .def("axpy_transposed", &SparseMatrix::axpy_transposed)
        .def("resize_and_clear", &SparseMatrix::resize_and_clear)
        .def("add", &SparseMatrix::add)
        .def("get", &SparseMatrix::get)
        .def("finalize", &SparseMatrix::finalize)
        .def("set_as_transpose_of", &SparseMatrix::set_as_transpose_of, py::arg("B"), py::arg("scale") = 1.0);
    py::class_<DomainDiscretization>(m, "DomainDiscretization")
        .def(py::init<int, int>(), py::arg("kind"), py::arg("num_refs"))
        .def("num_levels", &DomainDiscretization::num_levels)
        .def("num_dofs", &DomainDiscretization::num_dofs)
        .def("num_elements", &DomainDiscretization::num_elements)
        .def("set_theta", &DomainDiscretization::set_theta)
        .def("set_theta_device", &DomainDiscretization::set_theta_device)
        .def("set_source", &DomainDiscretization::set_source)
        .def("assemble_linear", &DomainDiscretization::assemble_linear)
        .def("assemble_jacobian", &DomainDiscretization::assemble_jacobian)
        .def("assemble_defect", &DomainDiscretization::assemble_defect)
        .def("adjust_defect", &DomainDiscretization::adjust_defect)
        .def("adjoint_gradient", &DomainDiscretization::adjoint_gradient)
        .def("dof_of_lattice", &DomainDiscretization::dof_of_lattice);
    py::class_<GeometricMultiGrid>(m, "GeometricMultiGrid")
        .def(py::init<DomainDiscretization&, int, int, int, int, double, const std::string&>(), py::arg("disc"),
             py::arg("base_level") = 0, py::arg("nu1") = 2, py::arg("nu2") = 2, py::arg("cycle") = 1,
             py::arg("omega") = 0.66, py::arg("precision") = "f64", py::keep_alive<1, 2>())
        .def("init", &GeometricMultiGrid::init)
        .def("apply", &GeometricMultiGrid::apply);
    py::class_<Jacobi>(m, "Jacobi").def(py::init<double>(), py::arg("omega") = 0.66);
    py::class_<Q2Discretization>(m, "Q2Discretization")
        .def(py::init<int>(), py::arg("num_refs"))
        .def("num_levels", &Q2Discretization::num_levels)
        .def("num_dofs", &Q2Discretization::num_dofs)
        .def("set_theta", &Q2Discretization::set_theta)
        .def("set_source", &Q2Discretization::set_source)
        .def("assemble_linear", &Q2Discretization::assemble_linear)
        .def("prolongation", &Q2Discretization::prolongation)
        .def("restriction", &Q2Discretization::restriction);
    py::class_<MatrixMultiGrid>(m, "MatrixMultiGrid")
        .def(py::init<py::list, py::list, py::list, int, int, int, int, double>(), py::arg("A"), py::arg("P"),
             py::arg("R"), py::arg("base_level") = 0, py::arg("nu1") = 2, py::arg("nu2") = 2, py::arg("cycle") = 1,
             py::arg("omega") = 0.66)
        .def("init", &MatrixMultiGrid::init)
        .def("apply", &MatrixMultiGrid::apply);
    py::class_<Solver>(m, "Solver")
        .def("init", &Solver::init, py::keep_alive<1, 2>())
        .def("apply", &Solver::apply)
        .def("steps", &Solver::steps);
    py::class_<CG, Solver>(m, "CG").def(py::init<py::object, int, double, double>(), py::arg("precond"),
                                        py::arg("max_steps") = 100, py::arg("reduction") = 1e-10,
                                        py::arg("min_defect") = 1e-50);
    py::class_<BiCGStab, Solver>(m, "BiCGStab")
        .def(py::init<py::object, int, double, double>(), py::arg("precond"), py::arg("max_steps") = 100,
             py::arg("reduction") = 1e-10, py::arg("min_defect") = 1e-50);
    py::class_<ElementDiscretization>(m, "ElementDiscretization")
        .def(py::init<int, IArray, DArray>(), py::arg("elem_type"), py::arg("elem_dofs"), py::arg("dof_coords"))
        .def("set_diffusion", &ElementDiscretization::set_diffusion)
        .def("set_diffusion_at_points", &ElementDiscretization::set_diffusion_at_points)
        .def("set_source", &ElementDiscretization::set_source)
        .def("set_dirichlet", &ElementDiscretization::set_dirichlet)
        .def("assemble_linear", &ElementDiscretization::assemble_linear)
        .def("assemble_defect", &ElementDiscretization::assemble_defect);
}
