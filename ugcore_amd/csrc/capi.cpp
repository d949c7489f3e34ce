// C ABI (include/ugcore_amd.h) over the kernels and the GPUAlgebra classes.
// No exception crosses this boundary: failures return negative status codes
// and leave a message for ugx_last_error().
#include "ugcore_amd.h"

#include <memory>
#include <string>

#include "ug_gpu/gpu_algebra.hpp"
#include "hierarchy.hpp"
#include "ugx_internal.hpp"

#include <algorithm>

using namespace ug::gpu;

namespace {

thread_local std::string g_err;

hipStream_t S(void* s) { return s ? (hipStream_t)s : context().stream; }

template <typename F>
int guard(F&& f) {
    try {
        return f();
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

template <typename F>
void* guard_ptr(F&& f) {
    try {
        return f();
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

int launched(int rc) {
    if (rc != 0) {
        g_err = "kernel launch failed";
        return -1;
    }
    return 0;
}

struct MatrixH {
    std::shared_ptr<GPUSparseMatrix> A = std::make_shared<GPUSparseMatrix>();
    std::shared_ptr<MatrixOperator> op = std::make_shared<MatrixOperator>(A);
};
struct VectorH {
    GPUVector v;
};
struct DiscH {
    std::shared_ptr<DomainDiscretization> d;
};
struct ElemH {
    std::unique_ptr<ElementDiscretization> e;
};
struct GmgH {
    std::shared_ptr<AssembledMultiGridCycle> g;
};
struct JacobiH {
    std::shared_ptr<Jacobi> j;
};
struct SolverH {
    std::shared_ptr<ILinearOperatorInverse> s;
};
struct Q2H {
    std::unique_ptr<Q2Discretization> q;
};
struct MgmgH {
    std::shared_ptr<MatrixMultiGrid> g;
};

// ---- the allocation-free dot product (ugx_dot_f64_noalloc) -------------------------
// Its partial sums live in this file-scope device array instead of the context's
// buffer: one double per workgroup of the reduction grid, which is capped at
// 65536 workgroups, so 512 KiB per device, held until the process ends. In host
// code `scratch` names a host-side stand-in, never the device memory: the device
// address of each device's copy is taken with hipGetSymbolAddress (at ugx_init,
// or at the entry point's first call on a device) and only that address is used.
constexpr int kNoAllocPartials = 65536;
static __device__ __align__(256) double scratch[kNoAllocPartials];

constexpr int kMaxDevices = 64;
double* g_scratch_dev[kMaxDevices] = {};   // the address of `scratch` per device

// the device address of `scratch` on device `dev`, which must be current
double* scratch_address(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw UGError("device index out of range for the no-alloc scratch table");
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            throw UGError("hipGetSymbolAddress failed for the no-alloc scratch array");
        g_scratch_dev[dev] = (double*)p;
    }
    return g_scratch_dev[dev];
}

}  // namespace

extern "C" {

int ugx_init(void) {
    return guard([] {
        auto& c = context();
        scratch_address(c.device);
        return 0;
    });
}
const char* ugx_last_error(void) { return g_err.c_str(); }
void* ugx_stream(void) {
    try {
        return (void*)context().stream;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
int ugx_synchronize(void) {
    return guard([] {
        sync();
        return 0;
    });
}

// ---- device buffers / events ---------------------------------------------------

void* ugx_malloc(int64_t bytes) {
    return guard_ptr([&] {
        context();
        void* p = nullptr;
        if (bytes <= 0 || hipMalloc(&p, (size_t)bytes) != hipSuccess) throw UGError("ugx_malloc failed");
        return p;
    });
}
int ugx_free(void* p) {
    return guard([&] {
        if (p && hipFree(p) != hipSuccess) throw UGError("ugx_free failed");
        return 0;
    });
}
int ugx_memcpy_htod(void* dst, const void* src, int64_t bytes) {
    return guard([&] {
        auto& c = context();
        if (hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, c.stream) != hipSuccess ||
            hipStreamSynchronize(c.stream) != hipSuccess)
            throw UGError("ugx_memcpy_htod failed");
        return 0;
    });
}
int ugx_memcpy_dtoh(void* dst, const void* src, int64_t bytes) {
    return guard([&] {
        auto& c = context();
        if (hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
            hipStreamSynchronize(c.stream) != hipSuccess)
            throw UGError("ugx_memcpy_dtoh failed");
        return 0;
    });
}
int ugx_memcpy_dtod(void* dst, const void* src, int64_t bytes) {
    return guard([&] {
        auto& c = context();
        if (hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, c.stream) != hipSuccess ||
            hipStreamSynchronize(c.stream) != hipSuccess)
            throw UGError("ugx_memcpy_dtod failed");
        return 0;
    });
}
void* ugx_event_create(void) {
    return guard_ptr([] {
        context();
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) throw UGError("ugx_event_create failed");
        return (void*)e;
    });
}
int ugx_event_destroy(void* ev) {
    return guard([&] { return hipEventDestroy((hipEvent_t)ev) == hipSuccess ? 0 : -1; });
}
int ugx_event_record(void* ev, void* stream) {
    return guard([&] {
        if (hipEventRecord((hipEvent_t)ev, S(stream)) != hipSuccess) throw UGError("ugx_event_record failed");
        return 0;
    });
}
int ugx_event_elapsed_ms(void* begin, void* end, double* ms) {
    return guard([&] {
        float f = 0.f;
        if (hipEventSynchronize((hipEvent_t)end) != hipSuccess ||
            hipEventElapsedTime(&f, (hipEvent_t)begin, (hipEvent_t)end) != hipSuccess)
            throw UGError("ugx_event_elapsed_ms failed");
        *ms = (double)f;
        return 0;
    });
}

// ---- kernels -----------------------------------------------------------------

int ugx_spmv_csr_f64(int64_t nrows, const int32_t* rowptr, const int32_t* col, const double* val, double alpha,
                     const double* v1, double beta, const double* w, double* dest, void* stream) {
    return guard([&] {
        if (nrows < 0 || !rowptr || !col || !val || !w || !dest) throw UGError("ugx_spmv_csr_f64: bad arguments");
        return launched(ugamd::spmv_csr<double, double>(nrows, rowptr, col, val, alpha, v1 ? v1 : dest, beta, w, dest,
                                                        0, 16, S(stream)));
    });
}
int ugx_spmv_csr_f32(int64_t nrows, const int32_t* rowptr, const int32_t* col, const float* val, double alpha,
                     const float* v1, double beta, const float* w, float* dest, void* stream) {
    return guard([&] {
        if (nrows < 0 || !rowptr || !col || !val || !w || !dest) throw UGError("ugx_spmv_csr_f32: bad arguments");
        return launched(ugamd::spmv_csr<float, float>(nrows, rowptr, col, val, alpha, v1 ? v1 : dest, beta, w, dest,
                                                      0, 16, S(stream)));
    });
}
int ugx_spmv_csr_ignore_zero_rows_f64(int64_t nrows, const int32_t* rowptr, const int32_t* col, const double* val,
                                      double beta, const double* w, double* dest, void* stream) {
    return guard([&] {
        if (nrows < 0 || !rowptr || !col || !val || !w || !dest) throw UGError("ugx_spmv_csr_ignore_zero_rows_f64: bad arguments");
        return launched(ugamd::spmv_csr<double, double>(nrows, rowptr, col, val, 0.0, dest, beta, w, dest, 1, 16,
                                                        S(stream)));
    });
}

int ugx_vec_scale_add2_f64(int64_t n, double* dest, double a1, const double* v1, double a2, const double* v2,
                           void* stream) {
    return guard([&] { return launched(ugamd::vec_scale_add<double>(n, dest, a1, v1, a2, v2, 0.0, nullptr, S(stream))); });
}
int ugx_vec_scale_add3_f64(int64_t n, double* dest, double a1, const double* v1, double a2, const double* v2,
                           double a3, const double* v3, void* stream) {
    return guard([&] { return launched(ugamd::vec_scale_add<double>(n, dest, a1, v1, a2, v2, a3, v3, S(stream))); });
}
int ugx_vec_set_f64(int64_t n, double* x, double value, void* stream) {
    return guard([&] { return launched(ugamd::vec_set<double>(n, x, value, S(stream))); });
}
int ugx_dot_f64(int64_t n, const double* a, const double* b, double* result_host, void* stream) {
    return guard([&] {
        auto& c = context();
        hipStream_t s = S(stream);
        if (ugamd::dot<double>(n, a, b, c.red_partial, c.red_out, false, s) != 0) throw UGError("dot launch failed");
        if (hipMemcpyAsync(c.red_host, c.red_out, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            throw UGError("dot readback failed");
        *result_host = c.red_host[0];
        return 0;
    });
}
int ugx_dot_f64_noalloc(int64_t n, const double* a, const double* b, double* result_dev, void* stream) {
    return guard([&] {
        if (n < 0) throw UGError("ugx_dot_f64_noalloc: negative length");
        if (!result_dev || (n > 0 && (!a || !b))) throw UGError("ugx_dot_f64_noalloc: null pointer");
        const int np = ugamd::dot_partials(n);
        if (np > kNoAllocPartials)
            throw UGError("ugx_dot_f64_noalloc: the reduction needs more partial sums than the 65536 of its scratch");
        hipStream_t s = (hipStream_t)stream;
        if (!s) {
            // the library's stream exists only after the set-up, which allocates
            DeviceContext* c = context_if_ready();
            if (!c) throw UGError("ugx_dot_f64_noalloc: pass a stream or call ugx_init() first");
            s = c->stream;
        }
        int dev = -1;
        if (hipGetDevice(&dev) != hipSuccess) throw UGError("ugx_dot_f64_noalloc: hipGetDevice failed");
        double* part = scratch_address(dev);
        // same kernels and grid as ugx_dot_f64: the same result bit for bit
        if (ugamd::dot<double>(n, a, b, part, result_dev, false, s) != 0)
            throw UGError("ugx_dot_f64_noalloc: launch failed");
        return 0;
    });
}
int ugx_norm_f64(int64_t n, const double* a, double* result_host, void* stream) {
    return guard([&] {
        auto& c = context();
        hipStream_t s = S(stream);
        if (ugamd::dot<double>(n, a, a, c.red_partial, c.red_out, true, s) != 0) throw UGError("norm launch failed");
        if (hipMemcpyAsync(c.red_host, c.red_out, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            throw UGError("norm readback failed");
        *result_host = c.red_host[0];
        return 0;
    });
}
int ugx_jacobi_diag_inv_f64(int64_t nrows, const int32_t* rowptr, const int32_t* col, const double* val, double omega,
                            double* dinv, void* stream) {
    return guard([&] {
        auto& c = context();
        hipStream_t s = S(stream);
        if (hipMemsetAsync(c.flag, 0, sizeof(int), s) != hipSuccess) throw UGError("flag reset failed");
        if (ugamd::jacobi_diag_csr<double, double>(nrows, rowptr, col, val, omega, dinv, c.flag, s) != 0)
            throw UGError("jacobi launch failed");
        int bad = 0;
        if (hipMemcpyAsync(&bad, c.flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            throw UGError("jacobi readback failed");
        if (bad) {
            g_err = "Jacobi: zero diagonal entry";
            return -2;
        }
        return 0;
    });
}
int ugx_jacobi_apply_f64(int64_t n, double* c, const double* dinv, const double* d, void* stream) {
    return guard([&] { return launched(ugamd::diag_mult<double, double>(n, c, dinv, d, S(stream))); });
}

// ---- domain discretisation ---------------------------------------------------------

int ugx_hierarchy_numbering(int kind, int num_refs, int level, int32_t* out, int64_t* n_inner) {
    return guard([&] {
        if (num_refs < 0 || num_refs > 10 || level < 0 || level > num_refs || kind < 0 || kind > 2)
            throw UGError("ugx_hierarchy_numbering: level out of range");
        auto nums = ugamd::build_level_numbering(kind, level);
        const auto& L = nums.back();
        std::copy(L.dof_of_lattice.begin(), L.dof_of_lattice.end(), out);
        if (n_inner) *n_inner = L.n_inner;
        return 0;
    });
}

int ugx_hierarchy_tets(int num_refs, int level, uint16_t* cell_tets, uint8_t* parent_edge) {
    return guard([&] {
        if (num_refs < 0 || num_refs > 7 || level < 0 || level > num_refs)
            throw UGError("ugx_hierarchy_tets: level out of range");
        auto nums = ugamd::build_level_numbering(2, level);
        const auto& L = nums.back();
        if (cell_tets) std::copy(L.cell_tets.begin(), L.cell_tets.end(), cell_tets);
        if (parent_edge && level > 0) std::copy(L.parent_edge.begin(), L.parent_edge.end(), parent_edge);
        return 0;
    });
}

void* ugx_disc_create(int kind, int num_refs) {
    return guard_ptr([&] {
        auto* h = new DiscH();
        h->d = std::make_shared<DomainDiscretization>(kind, num_refs);
        return (void*)h;
    });
}
void* ugx_disc_create_ordered(int kind, int num_refs, int ordering) {
    return guard_ptr([&] {
        auto* h = new DiscH();
        h->d = std::make_shared<DomainDiscretization>(kind, num_refs, ordering);
        return (void*)h;
    });
}
void* ugx_disc_create_setup(int kind, int num_refs, int ordering, int setup) {
    return guard_ptr([&] {
        auto* h = new DiscH();
        h->d = std::make_shared<DomainDiscretization>(kind, num_refs, ordering, setup);
        return (void*)h;
    });
}
int ugx_disc_level_tets(void* disc, int level, uint16_t* cell_tets, uint8_t* parent_codes) {
    return guard([&] {
        auto* h = (DiscH*)disc;
        if (h->d->kind() != 2) throw UGError("level_tets: the domain has no tetrahedra");
        if (level < 0 || level >= h->d->num_levels()) throw UGError("level_tets: level out of range");
        const LevelGrid& L = h->d->level(level);
        if (cell_tets) L.cell_tets.download(cell_tets, L.cell_tets.size());
        if (parent_codes) {
            if (level == 0) throw UGError("level_tets: level 0 has no parent codes");
            L.pcode.download(parent_codes, L.N);
        }
        return 0;
    });
}
void ugx_disc_destroy(void* disc) { delete (DiscH*)disc; }
int ugx_disc_num_levels(void* disc) { return ((DiscH*)disc)->d->num_levels(); }
int64_t ugx_disc_num_dofs(void* disc, int level) {
    auto* h = (DiscH*)disc;
    if (level < 0 || level >= h->d->num_levels()) return -1;
    return h->d->level(level).N;
}
int64_t ugx_disc_num_inner(void* disc, int level) {
    auto* h = (DiscH*)disc;
    if (level < 0 || level >= h->d->num_levels()) return -1;
    return h->d->level(level).n_inner;
}
int ugx_disc_dof_of_lattice(void* disc, int level, int32_t* out) {
    return guard([&] {
        auto& L = ((DiscH*)disc)->d->level(level);
        L.dof_of_lat.download(out, L.N);
        return 0;
    });
}
int ugx_disc_dof_positions(void* disc, int level, double* out) {
    return guard([&] {
        ((DiscH*)disc)->d->dof_positions(level, out);
        return 0;
    });
}
int ugx_disc_set_theta(void* disc, const double* theta) {
    return guard([&] {
        ((DiscH*)disc)->d->set_theta(theta);
        return 0;
    });
}
int ugx_disc_set_theta_device(void* disc, const double* theta) {
    return guard([&] {
        ((DiscH*)disc)->d->set_theta_device(theta);
        return 0;
    });
}
int ugx_disc_adjust_defect(void* disc, void* vec) {
    return guard([&] {
        ((DiscH*)disc)->d->adjust_defect(((VectorH*)vec)->v);
        return 0;
    });
}
int ugx_disc_adjoint_gradient(void* disc, void* u, void* lambda, double* grad) {
    return guard([&] {
        if (!grad) throw UGError("ugx_disc_adjoint_gradient: null output");
        ((DiscH*)disc)->d->adjoint_gradient(((VectorH*)u)->v, ((VectorH*)lambda)->v, grad);
        return 0;
    });
}
int ugx_disc_set_velocity(void* disc, double vx, double vy, double vz) {
    return guard([&] {
        ((DiscH*)disc)->d->set_velocity(vx, vy, vz);
        return 0;
    });
}
int ugx_disc_set_jitter(void* disc, uint64_t seed, double amp) {
    return guard([&] {
        ((DiscH*)disc)->d->set_jitter(seed, amp);
        return 0;
    });
}
int ugx_q2_hex_element_matrices(int64_t n, const double* corners, const double* k, double f_const, double* ke,
                                double* fe) {
    return guard([&] {
        if (n < 0) throw UGError("ugx_q2_hex_element_matrices: negative element count");
        if (n > 0 && (!corners || !ke)) throw UGError("ugx_q2_hex_element_matrices: null corners or output");
        if (ugamd::q2_hex_elements(n, corners, k, f_const, ke, fe, context().stream) != 0)
            throw UGError("ugx_q2_hex_element_matrices: launch failed");
        return 0;
    });
}
int ugx_q2_assemble_lattice(int n, const double* theta, double f_const, void* matrix, void* rhs) {
    return guard([&] {
        // nnz = (8n+1)^3 (an even node couples to 5 nodes per direction, an
        // odd one to 3: 2n+1 + 4n + 2n = 8n+1 per direction) must fit int32:
        // checked before anything is allocated
        if (n < 1 || n > 161) throw UGError("ugx_q2_assemble_lattice: n must be in 1..161 (nnz must fit int32)");
        if (!matrix || !rhs) throw UGError("ugx_q2_assemble_lattice: null matrix or vector");
        auto& c = context();
        const int W = 2 * n + 1;
        const int64_t N = (int64_t)W * W * W;
        DeviceBuffer<int32_t> cnt(N);
        if (ugamd::q2_lattice_count(n, cnt.get(), c.stream) != 0) throw UGError("ugx_q2_assemble_lattice: launch failed");
        std::vector<int32_t> h(N + 1, 0);
        cnt.download(h.data() + 1, N);
        int64_t acc = 0;
        for (int64_t i = 1; i <= N; ++i) {
            acc += h[i];
            if (acc >= ((int64_t)1 << 31)) throw UGError("ugx_q2_assemble_lattice: nnz exceeds int32 range");
            h[i] = (int32_t)acc;
        }
        GPUSparseMatrix& A = *((MatrixH*)matrix)->A;
        GPUVector& b = ((VectorH*)rhs)->v;
        A.allocate(N, N, acc);
        if (hipMemcpyAsync(A.rowptr(), h.data(), sizeof(int32_t) * (N + 1), hipMemcpyHostToDevice, c.stream) !=
                hipSuccess ||
            hipMemsetAsync(A.val(), 0, sizeof(double) * acc, c.stream) != hipSuccess)
            throw UGError("ugx_q2_assemble_lattice: copy failed");
        b.resize(N);
        b.set(0.0);
        if (ugamd::q2_lattice_fill(n, A.rowptr(), A.col(), c.stream) != 0 ||
            ugamd::q2_lattice_assemble(n, theta, f_const, A.rowptr(), A.col(), A.val(), b.data(), c.stream) != 0)
            throw UGError("ugx_q2_assemble_lattice: launch failed");
        if (hipStreamSynchronize(c.stream) != hipSuccess) throw UGError("ugx_q2_assemble_lattice: stream failed");
        A.set_origin(nullptr);
        A.build_spmv_layout();
        return 0;
    });
}
int ugx_disc_adjoint_operator(void* disc, void* a, void* at) {
    return guard([&] {
        ((DiscH*)disc)->d->assemble_adjoint_operator(*((MatrixH*)a)->A, *((MatrixH*)at)->A);
        return 0;
    });
}
int ugx_disc_set_source_const(void* disc, double f) {
    ((DiscH*)disc)->d->set_source_constant(f);
    return 0;
}
int ugx_disc_set_source_dofs(void* disc, const double* f) {
    return guard([&] {
        ((DiscH*)disc)->d->set_source_dofs(f);
        return 0;
    });
}
int ugx_disc_assemble_linear(void* disc, void* matrix, void* rhs) {
    return guard([&] {
        ((DiscH*)disc)->d->assemble_linear(*((MatrixH*)matrix)->A, ((VectorH*)rhs)->v);
        return 0;
    });
}
double ugx_disc_assemble_seconds(void* disc) { return ((DiscH*)disc)->d->assemble_seconds(); }
int ugx_disc_assemble_jacobian(void* disc, void* matrix, void* u) {
    return guard([&] {
        ((DiscH*)disc)->d->assemble_jacobian(*((MatrixH*)matrix)->A, ((VectorH*)u)->v);
        return 0;
    });
}
int ugx_disc_assemble_defect(void* disc, void* d, void* u) {
    return guard([&] {
        ((DiscH*)disc)->d->assemble_defect(((VectorH*)d)->v, ((VectorH*)u)->v);
        return 0;
    });
}

// ---- element-loop assembly from connectivity -------------------------------------------

void* ugx_elemdisc_create(int elem_type, int64_t num_elems, const int32_t* elem_dofs, int64_t num_dofs,
                          const double* dof_coords) {
    return guard_ptr([&] {
        auto* h = new ElemH();
        try {
            h->e.reset(new ElementDiscretization(elem_type, num_elems, elem_dofs, num_dofs, dof_coords, false));
        } catch (...) {
            delete h;
            throw;
        }
        return (void*)h;
    });
}
void ugx_elemdisc_destroy(void* e) { delete (ElemH*)e; }
int64_t ugx_elemdisc_nnz(void* e) { return ((ElemH*)e)->e->nnz(); }
int ugx_elemdisc_num_colours(void* e) { return ((ElemH*)e)->e->num_colours(); }
int ugx_elemdisc_set_diffusion(void* e, const double* k) {
    return guard([&] {
        ((ElemH*)e)->e->set_diffusion(k, false);
        return 0;
    });
}
int ugx_elemdisc_set_diffusion_at_points(void* e, const double* kq) {
    return guard([&] {
        ((ElemH*)e)->e->set_diffusion_at_points(kq, false);
        return 0;
    });
}
int ugx_elemdisc_set_source_const(void* e, double f) {
    ((ElemH*)e)->e->set_source(f);
    return 0;
}
int ugx_elemdisc_set_source_dofs(void* e, const double* f) {
    return guard([&] {
        if (!f) throw UGError("ugx_elemdisc_set_source_dofs: null source");
        ((ElemH*)e)->e->set_source_dofs(f);
        return 0;
    });
}
int ugx_elemdisc_set_velocity(void* e, double vx, double vy, double vz) {
    return guard([&] {
        ((ElemH*)e)->e->set_velocity(vx, vy, vz);
        return 0;
    });
}
int ugx_elemdisc_set_dirichlet(void* e, const int32_t* dofs, int64_t n) {
    return guard([&] {
        if (n < 0 || (n > 0 && !dofs)) throw UGError("ugx_elemdisc_set_dirichlet: bad list");
        ((ElemH*)e)->e->set_dirichlet(dofs, n);
        return 0;
    });
}
int ugx_elemdisc_assemble_linear(void* e, void* matrix, void* rhs) {
    return guard([&] {
        ((ElemH*)e)->e->assemble_linear(*((MatrixH*)matrix)->A, ((VectorH*)rhs)->v);
        return 0;
    });
}
int ugx_elemdisc_assemble_jacobian(void* e, void* matrix, void* u) {
    return guard([&] {
        ((ElemH*)e)->e->assemble_jacobian(*((MatrixH*)matrix)->A, ((VectorH*)u)->v);
        return 0;
    });
}
int ugx_elemdisc_assemble_defect(void* e, void* d, void* u) {
    return guard([&] {
        ((ElemH*)e)->e->assemble_defect(((VectorH*)d)->v, ((VectorH*)u)->v);
        return 0;
    });
}

// ---- matrices / vectors --------------------------------------------------------------

void* ugx_matrix_create(void) {
    return guard_ptr([] { return (void*)new MatrixH(); });
}
void ugx_matrix_destroy(void* m) { delete (MatrixH*)m; }
int64_t ugx_matrix_num_rows(void* m) { return ((MatrixH*)m)->A->num_rows(); }
int64_t ugx_matrix_nnz(void* m) { return ((MatrixH*)m)->A->total_num_connections(); }
int ugx_matrix_upload_csr(void* m, int64_t rows, int64_t cols, const int32_t* rowptr, const int32_t* col,
                          const double* val) {
    return guard([&] {
        ((MatrixH*)m)->A->set_from_host_csr(rows, cols, rowptr, col, val);
        return 0;
    });
}
int ugx_matrix_download_csr(void* m, int32_t* rowptr, int32_t* col, double* val) {
    return guard([&] {
        ((MatrixH*)m)->A->download_csr(rowptr, col, val);
        return 0;
    });
}
int ugx_matrix_device_csr(void* m, int32_t** rowptr, int32_t** col, double** val) {
    return guard([&] {
        auto& A = *((MatrixH*)m)->A;
        // the caller may write through these pointers: from here on the matrix is
        // no longer the assembly's (no matrix-free or lattice-order operator is
        // derived from it) and SpMV reads the CSR until ugx_matrix_finalize
        A.values_edited_externally();
        *rowptr = A.rowptr();
        *col = A.col();
        *val = A.val();
        return 0;
    });
}
int ugx_matrix_axpy(void* m, void* dest, double alpha, void* v1, double beta, void* w) {
    return guard([&] {
        ((MatrixH*)m)->A->axpy(((VectorH*)dest)->v, alpha, ((VectorH*)v1)->v, beta, ((VectorH*)w)->v);
        return 0;
    });
}
int ugx_matrix_set_spmv_lanes(void* m, int lanes) {
    return guard([&] {
        ((MatrixH*)m)->A->set_lanes(lanes);
        return 0;
    });
}
int ugx_matrix_set_spmv_layout(void* m, int layout) {
    return guard([&] {
        if (layout < 0 || layout > 2) throw UGError("ugx_matrix_set_spmv_layout: 0 (CSR), 1 (sliced ELL) or 2 (dictionary)");
        auto& A = *((MatrixH*)m)->A;
        A.set_spmv_layout(layout);
        return 0;
    });
}
int ugx_matrix_spmv_layout(void* m) {
    auto& A = *((MatrixH*)m)->A;
    return A.uses_dictionary() ? 2 : (A.uses_sliced_ell() ? 1 : 0);
}
int ugx_matrix_set_slice_order(void* m, const int32_t* order, int64_t nslots) {
    return guard([&] {
        auto& A = *((MatrixH*)m)->A;
        if (!order) {
            A.set_slice_order_device(nullptr, 0);
            return 0;
        }
        if (nslots <= 0 || nslots % 64 != 0)
            throw UGError("ugx_matrix_set_slice_order: the slots must be a positive multiple of 64");
        DeviceBuffer<int32_t> d(nslots);
        d.upload(order, nslots);
        A.set_slice_order_device(d.get(), nslots);   // checked on the host
        return 0;
    });
}
int ugx_matrix_set_row_order(void* m, const int32_t* order, int64_t n) {
    return guard([&] {
        ((MatrixH*)m)->A->set_row_order(order, n);
        return 0;
    });
}
int ugx_matrix_has_row_order(void* m) { return ((MatrixH*)m)->A->has_row_order() ? 1 : 0; }
int ugx_matrix_finalize(void* m) {
    return guard([&] {
        ((MatrixH*)m)->A->finalize();
        return 0;
    });
}
int ugx_matrix_apply_sub(void* m, void* d, void* x) {
    return guard([&] {
        ((MatrixH*)m)->A->matmul_minus(((VectorH*)d)->v, ((VectorH*)x)->v);
        return 0;
    });
}
int64_t ugx_matrix_num_cols(void* m) { return ((MatrixH*)m)->A->num_cols(); }
int ugx_matrix_device_csr_readonly(void* m, const int32_t** rowptr, const int32_t** col, const double** val) {
    return guard([&] {
        auto& A = *((MatrixH*)m)->A;
        if (A.building()) throw UGError("ugx_matrix_device_csr_readonly: the matrix is being built on the host");
        *rowptr = A.rowptr();
        *col = A.col();
        *val = A.val();
        return 0;
    });
}
int ugx_matrix_apply_ignore_zero_rows(void* m, void* dest, double beta, void* w) {
    return guard([&] {
        ((MatrixH*)m)->A->apply_ignore_zero_rows(((VectorH*)dest)->v, beta, ((VectorH*)w)->v);
        return 0;
    });
}
int ugx_matrix_resize_and_clear(void* m, int64_t rows, int64_t cols) {
    return guard([&] {
        ((MatrixH*)m)->A->resize_and_clear(rows, cols);
        return 0;
    });
}
int ugx_matrix_resize_and_keep_values(void* m, int64_t rows, int64_t cols) {
    return guard([&] {
        ((MatrixH*)m)->A->resize_and_keep_values(rows, cols);
        return 0;
    });
}
int ugx_matrix_clear_retain_structure(void* m) {
    return guard([&] {
        ((MatrixH*)m)->A->clear_retain_structure();
        return 0;
    });
}
int ugx_matrix_is_building(void* m) { return ((MatrixH*)m)->A->building() ? 1 : 0; }
double* ugx_matrix_entry(void* m, int64_t r, int64_t c) {
    try {
        return &(*((MatrixH*)m)->A)(r, c);
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
int ugx_matrix_get(void* m, int64_t r, int64_t c, double* value) {
    return guard([&] {
        auto& A = *((MatrixH*)m)->A;
        const bool stored = A.has_connection(r, c);
        *value = A.get(r, c);
        return stored ? 1 : 0;
    });
}
int64_t ugx_matrix_row_length(void* m, int64_t r) {
    try {
        return ((MatrixH*)m)->A->num_connections(r);
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}
int ugx_matrix_row(void* m, int64_t r, int32_t* cols, double* vals) {
    return guard([&] {
        ((MatrixH*)m)->A->row(r, cols, vals);
        return 0;
    });
}
int ugx_matrix_set_as_transpose_of(void* m, void* src, double scale, int keep_zeros) {
    return guard([&] {
        auto& A = *((MatrixH*)m)->A;
        const auto& B = *((MatrixH*)src)->A;
        if (keep_zeros) A.set_as_transpose_of(B, scale);
        else A.set_as_transpose_of2(B, scale);
        return 0;
    });
}
int ugx_matrix_set_as_copy_of(void* m, void* src, double scale) {
    return guard([&] {
        ((MatrixH*)m)->A->set_as_copy_of(*((MatrixH*)src)->A, scale);
        return 0;
    });
}
int ugx_matrix_scale(void* m, double a) {
    return guard([&] {
        ((MatrixH*)m)->A->scale(a);
        return 0;
    });
}
int ugx_matrix_axpy_transposed(void* m, void* dest, double alpha, void* v1, double beta, void* w) {
    return guard([&] {
        ((MatrixH*)m)->A->axpy_transposed(((VectorH*)dest)->v, alpha, ((VectorH*)v1)->v, beta, ((VectorH*)w)->v);
        return 0;
    });
}
int ugx_matrix_set_dirichlet_rows(void* m, const int32_t* rows_host, int64_t n) {
    return guard([&] {
        ((MatrixH*)m)->A->set_dirichlet_rows(rows_host, n);
        return 0;
    });
}

void* ugx_vector_create(int64_t n) {
    return guard_ptr([&] {
        auto* h = new VectorH();
        h->v.resize(n);
        h->v.set(0.0);
        return (void*)h;
    });
}
void ugx_vector_destroy(void* v) { delete (VectorH*)v; }
int64_t ugx_vector_size(void* v) { return ((VectorH*)v)->v.size(); }
double* ugx_vector_device_ptr(void* v) { return ((VectorH*)v)->v.data(); }
int ugx_vector_upload(void* v, const double* host) {
    return guard([&] {
        ((VectorH*)v)->v.upload(host);
        return 0;
    });
}
int ugx_vector_download(void* v, double* host) {
    return guard([&] {
        ((VectorH*)v)->v.download(host);
        return 0;
    });
}
int ugx_vector_set(void* v, double value) {
    return guard([&] {
        ((VectorH*)v)->v.set(value);
        return 0;
    });
}

// ---- GMG / Jacobi / solvers ---------------------------------------------------------------

int ugx_gmg_set_matrix_free(void* g, int on) {
    return guard([&] {
        ((GmgH*)g)->g->set_matrix_free(on != 0);
        return 0;
    });
}
int ugx_gmg_set_graphs(void* g, int on) {
    return guard([&] {
        ((GmgH*)g)->g->set_graphs(on != 0);
        return 0;
    });
}
int ugx_gmg_set_fused_smoothing(void* g, int on) {
    return guard([&] {
        ((GmgH*)g)->g->set_fused_smoothing(on != 0);
        return 0;
    });
}
int ugx_gmg_set_rap(void* g, int on) {
    return guard([&] {
        ((GmgH*)g)->g->set_rap(on != 0);
        return 0;
    });
}
int ugx_gmg_set_transposed(void* g, int transposed) {
    return guard([&] {
        ((GmgH*)g)->g->set_transposed(transposed != 0);
        return 0;
    });
}
// ---- Q2 hierarchy and the multigrid over matrices ----------------------------------
void* ugx_q2disc_create(int num_refs) {
    return guard_ptr([&] {
        std::unique_ptr<Q2H> h(new Q2H());
        h->q.reset(new Q2Discretization(num_refs));
        return (void*)h.release();
    });
}
void ugx_q2disc_destroy(void* q) { delete (Q2H*)q; }
int ugx_q2disc_num_levels(void* q) { return ((Q2H*)q)->q->num_levels(); }
int64_t ugx_q2disc_num_dofs(void* q, int level) {
    auto& d = *((Q2H*)q)->q;
    return level >= 0 && level < d.num_levels() ? d.num_dofs(level) : -1;
}
static void q2_level_check(Q2Discretization& d, int level) {
    if (level < 0 || level >= d.num_levels()) throw UGError("Q2 discretization: level out of range");
}
int ugx_q2disc_dof_positions(void* q, int level, double* out) {
    return guard([&] {
        auto& d = *((Q2H*)q)->q;
        q2_level_check(d, level);
        if (!out) throw UGError("ugx_q2disc_dof_positions: null output");
        d.dof_positions(level, out);
        return 0;
    });
}
int ugx_q2disc_dirichlet(void* q, int level, uint8_t* out) {
    return guard([&] {
        auto& d = *((Q2H*)q)->q;
        q2_level_check(d, level);
        if (!out) throw UGError("ugx_q2disc_dirichlet: null output");
        d.dirichlet_flags(level, out);
        return 0;
    });
}
int ugx_q2disc_set_theta(void* q, const double* theta) {
    return guard([&] {
        ((Q2H*)q)->q->set_theta(theta);
        return 0;
    });
}
int ugx_q2disc_set_source(void* q, double f) {
    ((Q2H*)q)->q->set_source(f);
    return 0;
}
int ugx_q2disc_assemble_linear(void* q, int level, void* matrix, void* rhs) {
    return guard([&] {
        auto& d = *((Q2H*)q)->q;
        q2_level_check(d, level);
        MatrixH* m = (MatrixH*)matrix;
        d.assemble_linear(level, *m->A, ((VectorH*)rhs)->v);
        return 0;
    });
}
int ugx_q2disc_prolongation(void* q, int level, void* P) {
    return guard([&] {
        ((Q2H*)q)->q->prolongation(level, *((MatrixH*)P)->A);
        return 0;
    });
}
int ugx_q2disc_restriction(void* q, int level, void* P, void* R) {
    return guard([&] {
        ((Q2H*)q)->q->restriction(level, *((MatrixH*)P)->A, *((MatrixH*)R)->A);
        return 0;
    });
}
void* ugx_mgmg_create(int num_levels, void* const* A, void* const* P, void* const* R, int base_level, int nu1,
                      int nu2, int cycle, double omega) {
    return guard_ptr([&] {
        if (num_levels < 1 || !A || !P || !R) throw UGError("ugx_mgmg_create: levels missing");
        std::vector<std::shared_ptr<GPUSparseMatrix>> a(num_levels), p(num_levels), r(num_levels);
        for (int l = 0; l < num_levels; ++l) {
            if (!A[l]) throw UGError("ugx_mgmg_create: null level operator");
            a[l] = ((MatrixH*)A[l])->A;
            if (l > 0) {
                if (!P[l] || !R[l]) throw UGError("ugx_mgmg_create: null transfer matrix");
                p[l] = ((MatrixH*)P[l])->A;
                r[l] = ((MatrixH*)R[l])->A;
            }
        }
        std::unique_ptr<MgmgH> h(new MgmgH());
        h->g = std::make_shared<MatrixMultiGrid>();
        h->g->set_levels(a, p, r);
        h->g->set_base_level(base_level);
        h->g->set_num_presmooth(nu1);
        h->g->set_num_postsmooth(nu2);
        h->g->set_cycle_type(cycle);
        h->g->set_smoother_damp(omega);
        return (void*)h.release();
    });
}
void ugx_mgmg_destroy(void* g) { delete (MgmgH*)g; }
int ugx_mgmg_init(void* g, void* matrix) {
    return guard([&] { return ((MgmgH*)g)->g->init(((MatrixH*)matrix)->op) ? 0 : -1; });
}
int ugx_mgmg_apply(void* g, void* c, void* d) {
    return guard([&] { return ((MgmgH*)g)->g->apply(((VectorH*)c)->v, ((VectorH*)d)->v) ? 0 : -1; });
}

void* ugx_gmg_create(void* disc, int base_level, int nu1, int nu2, int cycle, double omega, int precision) {
    return guard_ptr([&] {
        if (!disc) throw UGError("ugx_gmg_create: null discretization");
        if (precision != 0 && precision != 1) throw UGError("ugx_gmg_create: precision must be 0 (FP64) or 1 (FP32)");
        std::unique_ptr<GmgH> h(new GmgH());
        h->g = std::make_shared<AssembledMultiGridCycle>(((DiscH*)disc)->d);
        h->g->set_base_level(base_level);
        h->g->set_num_presmooth(nu1);
        h->g->set_num_postsmooth(nu2);
        h->g->set_cycle_type(cycle);
        h->g->set_smoother_damp(omega);
        h->g->set_precision(precision == 1 ? Precision::F32 : Precision::F64);
        return (void*)h.release();
    });
}
void ugx_gmg_destroy(void* g) { delete (GmgH*)g; }
int ugx_gmg_init(void* g, void* matrix) {
    return guard([&] { return ((GmgH*)g)->g->init(((MatrixH*)matrix)->op) ? 0 : -1; });
}
int ugx_gmg_apply(void* g, void* c, void* d) {
    return guard([&] { return ((GmgH*)g)->g->apply(((VectorH*)c)->v, ((VectorH*)d)->v) ? 0 : -1; });
}
int ugx_gmg_smooth_level(void* g, int level, int nsteps) {
    return guard([&] {
        ((GmgH*)g)->g->smooth_level(level, nsteps);
        return 0;
    });
}
int ugx_gmg_surface_apply_lattice(void* g, double* y, const double* x) {
    return guard([&] {
        ((GmgH*)g)->g->surface_axpy_lattice(y, 0.0, y, 1.0, x);
        return 0;
    });
}
int ugx_gmg_surface_direction_lattice(void* g, double* q, const double* p_old, const double* z, const double* beta,
                                      double* p_new, double* qp, double* partial) {
    return guard([&] {
        if (!q || !p_old || !z || !beta || !p_new || !qp || !partial)
            throw UGError("gmg_surface_direction_lattice: null pointer");
        if (!((GmgH*)g)->g->surface_direction_axpy_lattice(q, p_old, z, beta, p_new, qp, partial))
            throw UGError("gmg_surface_direction_lattice: needs the matrix-free top level");
        return 0;
    });
}
int ugx_gmg_apply_lattice(void* g, double* c, const double* d, const float* d_f32) {
    return guard([&] {
        auto& gm = *((GmgH*)g)->g;
        if (!c || !d) throw UGError("gmg_apply_lattice: null vector");
        gm.apply_lattice(c, d, nullptr, nullptr, d_f32);
        return 0;
    });
}
int ugx_gmg_lattice_uses_f32_defect(void* g) { return ((GmgH*)g)->g->lattice_uses_f32_defect() ? 1 : 0; }
int64_t ugx_gmg_lattice_halo(void* g) { return ((GmgH*)g)->g->lattice_halo(); }
int ugx_gmg_level_num_diags(void* g, int level) { return ((GmgH*)g)->g->level_num_diags(level); }
int ugx_gmg_level_dia(void* g, int level, double* out) {
    return guard([&] {
        ((GmgH*)g)->g->download_level_dia(level, out);
        return 0;
    });
}

void* ugx_jacobi_create(double omega) {
    return guard_ptr([&] {
        auto* h = new JacobiH();
        h->j = std::make_shared<Jacobi>(omega);
        return (void*)h;
    });
}
void ugx_jacobi_destroy(void* j) { delete (JacobiH*)j; }

void* ugx_solver_create(int type) {
    return guard_ptr([&] {
        auto* h = new SolverH();
        if (type == 0) h->s = std::make_shared<CG>();
        else if (type == 1) h->s = std::make_shared<BiCGStab>();
        else throw UGError("ugx_solver_create: unknown solver type");
        return (void*)h;
    });
}
void ugx_solver_destroy(void* s) { delete (SolverH*)s; }
int ugx_solver_set_preconditioner(void* s, int kind, void* p) {
    return guard([&] {
        auto& solver = *((SolverH*)s)->s;
        if (kind == 0 || !p) solver.set_preconditioner(nullptr);
        else if (kind == 1) solver.set_preconditioner(((GmgH*)p)->g);
        else if (kind == 2) solver.set_preconditioner(((JacobiH*)p)->j);
        else if (kind == 3) solver.set_preconditioner(((MgmgH*)p)->g);
        else throw UGError("ugx_solver_set_preconditioner: unknown kind");
        return 0;
    });
}
int ugx_solver_set_conv_check(void* s, int max_steps, double min_defect, double rel_reduction, int verbose) {
    return guard([&] {
        ((SolverH*)s)->s->set_convergence_check(
            std::make_shared<StdConvCheck>(max_steps, min_defect, rel_reduction, verbose != 0));
        return 0;
    });
}
int ugx_solver_init(void* s, void* matrix) {
    return guard([&] {
        if (!((SolverH*)s)->s->init(((MatrixH*)matrix)->op))
            throw UGError("solver init: the preconditioner's set-up failed (Jacobi: zero diagonal entry)");
        return 0;
    });
}
int ugx_solver_apply(void* s, void* x, void* b) {
    return guard([&] { return ((SolverH*)s)->s->apply(((VectorH*)x)->v, ((VectorH*)b)->v) ? 1 : 0; });
}
int ugx_solver_apply_return_defect(void* s, void* x, void* b) {
    return guard([&] { return ((SolverH*)s)->s->apply_return_defect(((VectorH*)x)->v, ((VectorH*)b)->v) ? 1 : 0; });
}
int ugx_solver_steps(void* s) { return ((SolverH*)s)->s->step(); }
int ugx_solver_set_restart(void* s, int n) {
    return guard([&] {
        auto* b = dynamic_cast<BiCGStab*>(((SolverH*)s)->s.get());
        if (!b) throw UGError("ugx_solver_set_restart: BiCGStab only");
        b->set_restart(n);
        return 0;
    });
}
int ugx_solver_set_min_orthogonality(void* s, double m) {
    return guard([&] {
        auto* b = dynamic_cast<BiCGStab*>(((SolverH*)s)->s.get());
        if (!b) throw UGError("ugx_solver_set_min_orthogonality: BiCGStab only");
        b->set_min_orthogonality(m);
        return 0;
    });
}
int ugx_solver_set_lattice_path(void* s, int on) {
    auto* cg = dynamic_cast<CG*>(((SolverH*)s)->s.get());
    if (!cg) {
        g_err = "ugx_solver_set_lattice_path: CG only";
        return -1;
    }
    cg->set_lattice_fast_path(on != 0);
    return 0;
}
int ugx_solver_used_lattice_path(void* s) {
    auto* cg = dynamic_cast<CG*>(((SolverH*)s)->s.get());
    return cg && cg->last_used_lattice_path() ? 1 : 0;
}
int ugx_solver_defects(void* s, double* out, int max_n) {
    const auto& d = ((SolverH*)s)->s->convergence_check().defects();
    const int n = (int)d.size();
    if (out)
        for (int i = 0; i < n && i < max_n; ++i) out[i] = d[i];
    return n;
}

// ---- file formats -------------------------------------------------------------------------

int ugx_cv_write_matrix(const char* path, int64_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                        const double* positions, int dim) {
    return guard([&] {
        if (!path || n < 0 || (n > 0 && (!rowptr || !positions))) throw UGError("ugx_cv_write_matrix: bad arguments");
        cv_write_matrix(path, n, rowptr, col, val, positions, dim);
        return 0;
    });
}
int ugx_cv_write_vector(const char* path, int64_t n, const double* values, const double* positions, int dim) {
    return guard([&] {
        if (!path || n < 0 || (n > 0 && (!values || !positions))) throw UGError("ugx_cv_write_vector: bad arguments");
        cv_write_vector(path, n, values, positions, dim);
        return 0;
    });
}
int ugx_vector_file_write(const char* path, int64_t n, const double* values) {
    return guard([&] {
        if (!path || n < 0 || (n > 0 && !values)) throw UGError("ugx_vector_file_write: bad arguments");
        vector_file_write(path, n, values);
        return 0;
    });
}
int64_t ugx_vector_file_length(const char* path) {
    try {
        if (!path) throw UGError("ugx_vector_file_length: no path");
        return vector_file_length(path);
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}
int ugx_vector_file_read(const char* path, int64_t n, double* values) {
    return guard([&] {
        if (!path || n < 0 || (n > 0 && !values)) throw UGError("ugx_vector_file_read: bad arguments");
        vector_file_read(path, n, values);
        return 0;
    });
}
int ugx_vector_save(void* v, const char* path) {
    return guard([&] {
        if (!path) throw UGError("ugx_vector_save: no path");
        SaveToFile(((VectorH*)v)->v, path);
        return 0;
    });
}
int ugx_vector_read(void* v, const char* path) {
    return guard([&] {
        if (!path) throw UGError("ugx_vector_read: no path");
        ReadFromFile(((VectorH*)v)->v, path);
        return 0;
    });
}

}  // extern "C"
