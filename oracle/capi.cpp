// Oracle (test infrastructure only): C entry points used by tests/ (ctypes)
// and by the CPU-baseline leg of bench.py.
#include <chrono>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "oracle.hpp"

using namespace orc;

namespace {

struct Problem {
    MultiGrid mg;
    int kind = 0;
    int n_fine = 0;
    std::vector<double> theta;          // per fine element (lattice order)
    std::vector<double> f_top;          // FV1 source per top-level DoF (kind 1)
    std::vector<CSR> A, P, R;
    std::vector<std::vector<double>> b;
    std::vector<double> x;
    TetParams tp;                       // kind 2: velocity, jitter
    BiCGStabOptions bicg;               // BiCGStab restarts
    bool record_x = false;              // keep every iterate of the next solves
    SolveResult last;
    double t_assemble = 0, t_setup = 0, t_solve = 0;
    std::string err;
    // Q2 Lagrange instead of Q1 (kind 0): numbering of every level
    bool q2 = false;
    std::vector<Q2Numbering> q2num;
};

thread_local std::string g_err;

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void copy_csr(const CSR& A, int64_t* rowptr, int32_t* col, double* val) {
    if (rowptr) std::memcpy(rowptr, A.rowptr.data(), sizeof(int64_t) * A.rowptr.size());
    if (col) std::memcpy(col, A.col.data(), sizeof(int32_t) * A.col.size());
    if (val) std::memcpy(val, A.val.data(), sizeof(double) * A.val.size());
}

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// ---- mesh hierarchy -------------------------------------------------------------

void* orc_problem_create(int kind, int num_refs) {
    try {
        auto* p = new Problem();
        p->kind = kind;
        p->mg = build_hierarchy(kind, num_refs);
        p->n_fine = (kind == 2 ? 3 : 1) << num_refs;
        p->A.resize(num_refs + 1);
        p->P.resize(num_refs + 1);
        p->R.resize(num_refs + 1);
        p->b.resize(num_refs + 1);
        return p;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_problem_destroy(void* h) { delete (Problem*)h; }

int orc_num_levels(void* h) { return (int)((Problem*)h)->mg.levels.size(); }

int64_t orc_num_dofs(void* h, int lev) { return ((Problem*)h)->mg.levels[lev].n_dof; }

int64_t orc_num_elements(void* h, int lev) {
    auto* p = (Problem*)h;
    const Level& L = p->mg.levels[lev];
    return p->mg.dim == 3 ? (int64_t)L.vols.size() : (int64_t)L.faces.size();
}

// lattice coordinates (units of the finest lattice) per DoF, 3 int64 each
void orc_dof_lattice(void* h, int lev, int64_t* out) {
    const Level& L = ((Problem*)h)->mg.levels[lev];
    for (int d = 0; d < L.n_dof; ++d) {
        const auto& p = L.vpos[L.vtx_of_dof[d]];
        out[3 * d] = p[0];
        out[3 * d + 1] = p[1];
        out[3 * d + 2] = p[2];
    }
}

void orc_dirichlet_mask(void* h, int lev, uint8_t* out) {
    auto m = dirichlet_mask(((Problem*)h)->mg, lev);
    std::memcpy(out, m.data(), m.size());
}

// element -> DoF table in element creation order, local vertex order
void orc_elements(void* h, int lev, int32_t* out) {
    auto* p = (Problem*)h;
    const Level& L = p->mg.levels[lev];
    if (p->mg.dim == 3) {
        const int nv = p->kind == 2 ? 4 : 8;
        for (size_t e = 0; e < L.vols.size(); ++e)
            for (int i = 0; i < nv; ++i) out[nv * e + i] = L.dof[L.vols[e][i]];
    } else {
        for (size_t e = 0; e < L.faces.size(); ++e)
            for (int i = 0; i < 3; ++i) out[3 * e + i] = L.dof[L.faces[e][i]];
    }
}

// ---- Q2 (kind 0) ---------------------------------------------------------------------
// switches the problem to Q2 Lagrange: later orc_assemble calls form the Q2
// level operators and transfers in the Q2 LEVEL numbering
int orc_q2_enable(void* h) {
    auto* p = (Problem*)h;
    try {
        if (p->kind != 0) throw std::runtime_error("oracle: Q2 needs the hexahedral hierarchy (kind 0)");
        p->q2 = true;
        p->q2num.clear();
        for (const auto& L : p->mg.levels) p->q2num.push_back(q2_numbering(L));
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int64_t orc_q2_num_dofs(void* h, int lev) { return ((Problem*)h)->q2num[lev].n_dof; }

// DoF positions in doubled finest-lattice units, 3 int64 each, and the
// Dirichlet flags
void orc_q2_dof_positions(void* h, int lev, int64_t* pos2, uint8_t* dir) {
    const auto& Q = ((Problem*)h)->q2num[lev];
    for (int d = 0; d < Q.n_dof; ++d) {
        for (int c = 0; c < 3; ++c) pos2[3 * d + c] = Q.pos2[d][c];
        if (dir) dir[d] = Q.dir[d];
    }
}

// ---- synthetic data ---------------------------------------------------------------

void orc_make_theta(int n, uint64_t seed, int passes, double* out) {
    auto t = make_theta(n, seed, passes);
    std::memcpy(out, t.data(), sizeof(double) * t.size());
}

double orc_gauss(uint64_t seed, uint64_t counter) { return gauss_from_counter(seed, counter); }

void orc_set_theta(void* h, const double* theta) {
    auto* p = (Problem*)h;
    const int64_t n = (int64_t)p->n_fine * p->n_fine * (p->mg.dim == 3 ? p->n_fine : 1);
    p->theta.assign(theta, theta + n);
}

// kind 2: convection velocity and top-level jitter (amplitude in units of h)
void orc_set_tet_params(void* h, double vx, double vy, double vz, uint64_t jitter_seed, double jitter_amp) {
    auto* p = (Problem*)h;
    p->tp.vel[0] = vx;
    p->tp.vel[1] = vy;
    p->tp.vel[2] = vz;
    p->tp.jitter_seed = jitter_seed;
    p->tp.jitter_amp = jitter_amp;
}

// positions (jittered on the top level) per DoF of level lev, 3 doubles each
void orc_dof_positions(void* h, int lev, double* out) {
    auto* p = (Problem*)h;
    const Level& L = p->mg.levels[lev];
    for (int d = 0; d < L.n_dof; ++d) tet_vertex_position(p->mg, lev, L.vtx_of_dof[d], p->tp, out + 3 * d);
}

void orc_set_source_top(void* h, const double* f) {
    auto* p = (Problem*)h;
    const int top = (int)p->mg.levels.size() - 1;
    p->f_top.assign(f, f + p->mg.levels[top].n_dof);
}

// ---- assembly -------------------------------------------------------------------------

int orc_assemble(void* h, int lev, double f_const) {
    auto* p = (Problem*)h;
    try {
        Coefficient k;
        k.n_fine = p->n_fine;
        k.dim = p->mg.dim;
        k.theta = p->theta.empty() ? nullptr : p->theta.data();
        double t0 = now();
        if (p->kind == 0 && p->q2) {
            assemble_q2_hex(p->mg, lev, p->q2num[lev], k, f_const, p->A[lev], p->b[lev]);
        } else if (p->kind == 0) {
            assemble_q1_hex(p->mg, lev, k, f_const, p->A[lev], p->b[lev]);
        } else if (p->kind == 2) {
            assemble_fv1_tet(p->mg, lev, p->theta, p->tp, f_const, p->A[lev], p->b[lev]);
        } else {
            const int top = (int)p->mg.levels.size() - 1;
            std::vector<double> f(p->mg.levels[lev].n_dof, 0.0);
            if (lev == top && !p->f_top.empty()) f = p->f_top;
            assemble_fv1_tri(p->mg, lev, k, f, p->A[lev], p->b[lev]);
        }
        p->t_assemble += now() - t0;
        if (lev > 0 && p->q2) {
            p->P[lev] = prolongation_q2(p->mg, lev, p->q2num[lev - 1], p->q2num[lev]);
            p->R[lev] = restriction_q2(p->mg, lev, p->q2num[lev - 1], p->q2num[lev], p->P[lev]);
        } else if (lev > 0) {
            p->P[lev] = prolongation_p1(p->mg, lev);
            p->R[lev] = restriction_from(p->mg, lev, p->P[lev]);
        }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Galerkin coarse operators A_{l-1} = R_l A_l P_l from the top level down
// (AssembledMultiGridCycle::init_rap_operator, mg_solver_impl.hpp:829-1013);
// call after every level is assembled
int orc_set_rap(void* h) {
    auto* p = (Problem*)h;
    try {
        const int top = (int)p->mg.levels.size() - 1;
        for (int lev = top; lev >= 1; --lev) p->A[lev - 1] = multiply_of(p->R[lev], p->A[lev], p->P[lev]);
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void orc_set_bicgstab_options(void* h, int num_restarts, double min_ortho) {
    auto* p = (Problem*)h;
    p->bicg.num_restarts = num_restarts;
    p->bicg.min_ortho = min_ortho;
}

int64_t orc_matrix_nnz(void* h, int which, int lev) {
    auto* p = (Problem*)h;
    const CSR& A = which == 0 ? p->A[lev] : which == 1 ? p->P[lev] : p->R[lev];
    return A.nnz();
}

int64_t orc_matrix_rows(void* h, int which, int lev) {
    auto* p = (Problem*)h;
    const CSR& A = which == 0 ? p->A[lev] : which == 1 ? p->P[lev] : p->R[lev];
    return A.n_rows;
}

// which: 0 = level matrix A, 1 = prolongation P (lev-1 -> lev), 2 = restriction R
void orc_matrix_get(void* h, int which, int lev, int64_t* rowptr, int32_t* col, double* val) {
    auto* p = (Problem*)h;
    const CSR& A = which == 0 ? p->A[lev] : which == 1 ? p->P[lev] : p->R[lev];
    copy_csr(A, rowptr, col, val);
}

void orc_rhs_get(void* h, int lev, double* out) {
    auto* p = (Problem*)h;
    std::memcpy(out, p->b[lev].data(), sizeof(double) * p->b[lev].size());
}

// ---- solve ------------------------------------------------------------------------------
// solver: 0 = CG, 1 = BiCGStab. precond: 0 = none, 1 = GMG.
int orc_solve(void* h, int solver, int precond, int base_lev, int nu1, int nu2, int cycle,
              double omega, double rel_red, double min_def, int max_steps) {
    auto* p = (Problem*)h;
    try {
        const int top = (int)p->mg.levels.size() - 1;
        double t0 = now();
        GMG g;
        if (precond) {
            g.base_lev = base_lev;
            g.top_lev = top;
            g.nu1 = nu1;
            g.nu2 = nu2;
            g.cycle = cycle;
            g.omega = omega;
            g.A = p->A;
            g.P = p->P;
            g.R = p->R;
            g.init();
        }
        double t1 = now();
        p->x.assign(p->A[top].n_rows, 0.0);
        ConvCheck cc;
        cc.max_steps = max_steps;
        cc.min_defect = min_def;
        cc.rel_reduction = rel_red;
        p->last = solver == 0 ? cg_solve(p->A[top], precond ? &g : nullptr, p->x, p->b[top], cc, p->record_x)
                              : bicgstab_solve(p->A[top], precond ? &g : nullptr, p->x, p->b[top], cc, p->bicg,
                                               p->record_x);
        double t2 = now();
        p->t_setup = t1 - t0;
        p->t_solve = t2 - t1;
        return p->last.ok ? p->last.steps : -1 - p->last.steps;
    } catch (std::exception& e) {
        g_err = e.what();
        return -100000;
    }
}

// one application c = GMG(d) on the top level (AssembledMultiGridCycle::apply)
int orc_gmg_apply(void* h, int base_lev, int nu1, int nu2, int cycle, double omega, const double* d, double* c) {
    auto* p = (Problem*)h;
    try {
        const int top = (int)p->mg.levels.size() - 1;
        GMG g;
        g.base_lev = base_lev;
        g.top_lev = top;
        g.nu1 = nu1;
        g.nu2 = nu2;
        g.cycle = cycle;
        g.omega = omega;
        g.A = p->A;
        g.P = p->P;
        g.R = p->R;
        g.init();
        const int64_t n = p->A[top].n_rows;
        std::vector<double> dv(d, d + n), cv(n, 0.0);
        g.apply(cv, dv);
        std::memcpy(c, cv.data(), sizeof(double) * n);
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// solve A x = rhs on the top level (rhs given; Dirichlet entries zeroed as
// DirichletBoundary::adjust_defect does), GMG-CG; returns steps (<0 failure)
int orc_solve_rhs(void* h, const double* rhs, double* x_out, int base_lev, double rel_red, int zero_dirichlet) {
    auto* p = (Problem*)h;
    try {
        const int top = (int)p->mg.levels.size() - 1;
        GMG g;
        g.base_lev = base_lev;
        g.top_lev = top;
        g.A = p->A;
        g.P = p->P;
        g.R = p->R;
        g.init();
        const int64_t n = p->A[top].n_rows;
        auto dm = p->q2 ? std::vector<uint8_t>(p->q2num[top].dir) : dirichlet_mask(p->mg, top);
        std::vector<double> b(rhs, rhs + n), x(n, 0.0);
        if (zero_dirichlet)
            for (int64_t i = 0; i < n; ++i)
                if (dm[i]) b[i] = 0.0;
        ConvCheck cc;
        cc.max_steps = 200;
        cc.rel_reduction = rel_red;
        auto r = cg_solve(p->A[top], &g, x, b, cc);
        std::memcpy(x_out, x.data(), sizeof(double) * n);
        return r.ok ? r.steps : -1 - r.steps;
    } catch (std::exception& e) {
        g_err = e.what();
        return -100000;
    }
}

void orc_adjoint_gradient(void* h, const double* u, const double* lambda, double* g_out) {
    auto* p = (Problem*)h;
    Coefficient k;
    k.n_fine = p->n_fine;
    k.dim = 3;
    k.theta = p->theta.empty() ? nullptr : p->theta.data();
    const int top = (int)p->mg.levels.size() - 1;
    const int64_t n = p->mg.levels[top].n_dof;
    std::vector<double> uu(u, u + n), ll(lambda, lambda + n), g;
    if (p->kind == 2)
        adjoint_gradient_fv1_tet(p->mg, p->theta, p->tp, uu, ll, g);
    else
        adjoint_gradient_q1_hex(p->mg, k, uu, ll, g);
    std::memcpy(g_out, g.data(), sizeof(double) * g.size());
}

// adjoint solve A^T lambda = rhs on the top level for an operator with
// identity Dirichlet rows: lambda on the interior solves A_II^T lambda_I =
// rhs_I, i.e. the transpose of the Dirichlet-eliminated operator (Dirichlet
// columns of interior rows zeroed), with rhs Dirichlet entries zeroed.
// GMG V(nu1,nu2) on the transposed eliminated level operators (P, R
// unchanged), BiCGStab (solver 1) or CG (solver 0). Returns steps (<0 failure).
int orc_solve_adjoint(void* h, int solver, const double* rhs, double* lambda_out, int base_lev, int nu1, int nu2,
                      double omega, double rel_red, int max_steps) {
    auto* p = (Problem*)h;
    try {
        const int top = (int)p->mg.levels.size() - 1;
        GMG g;
        g.base_lev = base_lev;
        g.top_lev = top;
        g.nu1 = nu1;
        g.nu2 = nu2;
        g.omega = omega;
        g.A.resize(p->A.size());
        for (int l = base_lev; l <= top; ++l) {
            CSR E = p->A[l];
            auto dm = dirichlet_mask(p->mg, l);
            for (int64_t r = 0; r < E.n_rows; ++r) {
                if (dm[r]) continue;
                for (int64_t q = E.rowptr[r]; q < E.rowptr[r + 1]; ++q)
                    if (dm[E.col[q]]) E.val[q] = 0.0;
            }
            g.A[l] = transpose_keep_zeros(E);
        }
        g.P = p->P;
        g.R = p->R;
        g.init();
        const int64_t n = p->A[top].n_rows;
        auto dm = p->q2 ? std::vector<uint8_t>(p->q2num[top].dir) : dirichlet_mask(p->mg, top);
        std::vector<double> b(rhs, rhs + n), x(n, 0.0);
        for (int64_t i = 0; i < n; ++i)
            if (dm[i]) b[i] = 0.0;
        ConvCheck cc;
        cc.max_steps = max_steps;
        cc.rel_reduction = rel_red;
        auto r = solver == 0 ? cg_solve(g.A[top], &g, x, b, cc) : bicgstab_solve(g.A[top], &g, x, b, cc, p->bicg);
        std::memcpy(lambda_out, x.data(), sizeof(double) * n);
        p->last = r;
        return r.ok ? r.steps : -1 - r.steps;
    } catch (std::exception& e) {
        g_err = e.what();
        return -100000;
    }
}

// iterate recording for the solves of orc_solve (tests pin the Krylov
// methods by properties of their iterates)
void orc_set_record_x(void* h, int on) { ((Problem*)h)->record_x = on != 0; }
int orc_iterates_count(void* h) { return (int)((Problem*)h)->last.xs.size(); }
int orc_iterate_get(void* h, int k, double* out) {
    auto* p = (Problem*)h;
    if (k < 0 || k >= (int)p->last.xs.size()) return -1;
    std::memcpy(out, p->last.xs[k].data(), sizeof(double) * p->last.xs[k].size());
    return 0;
}

// CG (solver 0) or BiCGStab (1) without a preconditioner on a caller's CSR
// (row pointers int64) from x (in/out); returns steps if converged, -1-steps
// if not (breakdown, NaN defect or step limit), -100000 on an exception.
// Failure-handling tests drive it with singular and NaN inputs.
int orc_csr_solve(int solver, int64_t n, const int64_t* rowptr, const int32_t* col, const double* val,
                  const double* b, double* x, double rel_red, double min_def, int max_steps) {
    try {
        CSR A;
        A.n_rows = n;
        A.n_cols = n;
        A.rowptr.assign(rowptr, rowptr + n + 1);
        A.col.assign(col, col + rowptr[n]);
        A.val.assign(val, val + rowptr[n]);
        std::vector<double> xv(x, x + n), bv(b, b + n);
        ConvCheck cc;
        cc.max_steps = max_steps;
        cc.min_defect = min_def;
        cc.rel_reduction = rel_red;
        SolveResult r = solver == 0 ? cg_solve(A, nullptr, xv, bv, cc) : bicgstab_solve(A, nullptr, xv, bv, cc);
        std::memcpy(x, xv.data(), sizeof(double) * n);
        return r.ok ? r.steps : -1 - r.steps;
    } catch (std::exception& e) {
        g_err = e.what();
        return -100000;
    }
}

void orc_solution_get(void* h, double* out) {
    auto* p = (Problem*)h;
    std::memcpy(out, p->x.data(), sizeof(double) * p->x.size());
}

int orc_defects_get(void* h, double* out, int max_n) {
    auto* p = (Problem*)h;
    int n = (int)p->last.defects.size();
    if (out)
        for (int i = 0; i < n && i < max_n; ++i) out[i] = p->last.defects[i];
    return n;
}

void orc_timings(void* h, double* out3) {
    auto* p = (Problem*)h;
    out3[0] = p->t_assemble;
    out3[1] = p->t_setup;
    out3[2] = p->t_solve;
}

// ---- CSR semantics probes (insertion, transpose, SpMV variants) -------------------------

void* orc_builder_create(int64_t rows, int64_t cols) {
    auto* b = new RowBuilder();
    b->resize_and_clear(rows, cols);
    return b;
}
void orc_builder_set(void* h, int64_t r, int64_t c, double v) { ((RowBuilder*)h)->at(r, c) = v; }
void orc_builder_add(void* h, int64_t r, int64_t c, double v) { ((RowBuilder*)h)->at(r, c) += v; }
void orc_builder_destroy(void* h) { delete (RowBuilder*)h; }

// mode: 0 = as built, 1 = transpose keeping zeros, 2 = transpose dropping zeros
int64_t orc_builder_csr(void* h, int mode, int64_t* rowptr, int32_t* col, double* val) {
    CSR A = defragment(*(RowBuilder*)h);
    if (mode == 1) A = transpose_keep_zeros(A);
    else if (mode == 2) A = transpose_drop_zeros(A);
    copy_csr(A, rowptr, col, val);
    return A.nnz();
}

// op: 0 = axpy (dest = a*v1 + b*A*w), 1 = apply_ignore_zero_rows (dest = b*A*w, empty rows untouched),
//     2 = axpy_transposed (dest = a*v1 + b*A^T*w)
void orc_csr_op(int op, int64_t rows, int64_t cols, const int64_t* rowptr, const int32_t* col,
                const double* val, double alpha, const double* v1, double beta, const double* w,
                double* dest) {
    CSR A;
    A.n_rows = rows;
    A.n_cols = cols;
    A.rowptr.assign(rowptr, rowptr + rows + 1);
    A.col.assign(col, col + rowptr[rows]);
    A.val.assign(val, val + rowptr[rows]);
    const int64_t nd = op == 2 ? cols : rows;
    const int64_t nw = op == 2 ? rows : cols;
    std::vector<double> d(dest, dest + nd), vv(v1 ? v1 : dest, (v1 ? v1 : dest) + nd), ww(w, w + nw);
    if (op == 0) axpy(A, d, alpha, vv, beta, ww);
    else if (op == 1) apply_ignore_zero_rows(A, d, beta, ww);
    else axpy_transposed(A, d, alpha, vv, beta, ww);
    std::memcpy(dest, d.data(), sizeof(double) * nd);
}

// element matrices of n Q2 hexahedra: X [n][8][3] corners in the reference
// order, k [n] (or NULL: 1), Ke [n][27][27], fe [n][27] (or NULL); mi [27][3]
// receives the local multi-indices when not NULL
void orc_q2_hex_elements(int64_t n, const double* X, const double* k, double f_const, double* Ke, double* fe,
                         int32_t* mi) {
    if (mi) {
        int t[27][3];
        orc::q2_hex_multi_index(t);
        for (int a = 0; a < 27; ++a)
            for (int d = 0; d < 3; ++d) mi[3 * a + d] = t[a][d];
    }
    for (int64_t e = 0; e < n; ++e) {
        double Xe[8][3], K[27][27], F[27];
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) Xe[a][d] = X[(e * 8 + a) * 3 + d];
        orc::q2_hex_element(Xe, k ? k[e] : 1.0, f_const, K, F);
        for (int a = 0; a < 27; ++a) {
            for (int b = 0; b < 27; ++b) Ke[(e * 27 + a) * 27 + b] = K[a][b];
            if (fe) fe[e * 27 + a] = F[a];
        }
    }
}

}  // extern "C"
