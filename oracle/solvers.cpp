// Oracle (test infrastructure only): Jacobi, dense LU, geometric multigrid
// V-cycle, CG and BiCGStab restated from ugcore.
#include <cmath>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

// ---- StdConvCheck (convergence_check_impl.h:134-230) --------------------------

bool ConvCheck::valid() const { return std::isfinite(current); }

bool ConvCheck::iteration_ended() const {
    if (!valid()) return true;
    if (step >= max_steps) return true;
    if (current < min_defect) return true;
    if (current / initial < rel_reduction) return true;
    return false;
}

bool ConvCheck::post() const {
    return current < min_defect || current / initial < rel_reduction;
}

// ---- Jacobi (jacobi.h:155-249) ------------------------------------------------

bool Jacobi::init(const CSR& A) {
    diag_inv.assign(A.n_rows, 0.0);
    for (int64_t i = 0; i < A.n_rows; ++i) {
        double d = 0.0;
        for (int64_t k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k)
            if (A.col[k] == i) d = A.val[k];
        double m = d;
        m *= 1.0 / omega;
        if (m == 0.0) return false;      // GetInverse of a zero diagonal
        diag_inv[i] = 1.0 / m;
    }
    return true;
}

void Jacobi::apply(std::vector<double>& c, const std::vector<double>& d) const {
    for (size_t i = 0; i < diag_inv.size(); ++i) c[i] = diag_inv[i] * d[i];
}

// ---- dense LU (lu.h:122-143, no_lapack/lu_decomp.h:47-200) ---------------------

bool DenseLU::init(const CSR& A) {
    n = (int)A.n_rows;
    a.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int64_t k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k) a[(size_t)i * n + A.col[k]] = A.val[k];
    piv.assign(n, 0);
    auto M = [&](int r, int c) -> double& { return a[(size_t)r * n + c]; };
    for (int k = 0; k < n; ++k) {
        int big = k;
        for (int j = k + 1; j < n; ++j)
            if (std::fabs(M(big, k)) < std::fabs(M(j, k))) big = j;
        if (big != k)
            for (int j = 0; j < n; ++j) std::swap(M(k, j), M(big, j));
        piv[k] = big;
        if (std::fabs(M(k, k)) < 1e-10) return false;
        for (int i = k + 1; i < n; ++i) {
            M(i, k) = M(i, k) / M(k, k);
            for (int j = k + 1; j < n; ++j) M(i, j) = M(i, j) - M(i, k) * M(k, j);
        }
    }
    return true;
}

void DenseLU::apply(std::vector<double>& x, const std::vector<double>& b) const {
    x = b;
    auto M = [&](int r, int c) { return a[(size_t)r * n + c]; };
    for (int i = 0; i < n; ++i)
        if ((size_t)i < piv[i]) std::swap(x[i], x[piv[i]]);
    for (int i = 0; i < n; ++i) {
        double s = x[i];
        for (int k = 0; k < i; ++k) s -= M(i, k) * x[k];
        x[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int k = i + 1; k < n; ++k) s -= M(i, k) * x[k];
        x[i] = s / M(i, i);
    }
}

// ---- AssembledMultiGridCycle (mg_solver_impl.hpp) ------------------------------
// Full refinement, LEVEL grid view: the surface is the top level with the
// identity surface<->level map, m_LocalFullRefLevel == top level, no shadowing
// DoFs (init_index_mappings, mg_solver_impl.hpp:1280-1370).

void GMG::init() {
    // set_cycle_type accepts V (1), W (2), F (-1) and gamma > 2; 0 and other
    // negative values would run no coarse correction and are refused
    if (cycle != -1 && cycle < 1) throw std::runtime_error("oracle GMG: cycle must be 1, 2, gamma > 2 or -1 (F)");
    if (nu1 < 0 || nu2 < 0) throw std::runtime_error("oracle GMG: smoothing step counts must be >= 0");
    if (base_lev < 0 || base_lev > top_lev) throw std::runtime_error("oracle GMG: base level out of range");
    const int nl = top_lev + 1;
    sc.assign(nl, {});
    sd.assign(nl, {});
    st.assign(nl, {});
    smoother.assign(nl, Jacobi());
    for (int l = base_lev; l <= top_lev; ++l) {
        sc[l].assign(A[l].n_rows, 0.0);
        sd[l].assign(A[l].n_rows, 0.0);
        st[l].assign(A[l].n_rows, 0.0);
        if (l > base_lev) {
            smoother[l].omega = omega;
            if (!smoother[l].init(A[l])) throw std::runtime_error("oracle GMG: Jacobi init failed");
        }
    }
    if (!base.init(A[base_lev])) throw std::runtime_error("oracle GMG: base LU singular");
}

void GMG::apply(std::vector<double>& c, const std::vector<double>& d) {
    // mg_solver_impl.hpp:175-275
    sd[top_lev] = d;
    std::fill(c.begin(), c.end(), 0.0);
    std::fill(sc[top_lev].begin(), sc[top_lev].end(), 0.0);
    lmgc(top_lev, cycle);
    for (size_t i = 0; i < c.size(); ++i) c[i] += sc[top_lev][i];
}

void GMG::lmgc(int lev, int type) {
    // mg_solver_impl.hpp:2090-2136
    if (lev == base_lev) { base_solve(top_lev); return; }
    presmooth_and_restriction(lev);
    if (lev - 1 == base_lev) {
        base_solve(lev - 1);                       // on the base level invert only once
    } else if (type == -1) {
        lmgc(lev - 1, -1);                         // F-cycle: one F-cycle on lev-1,
        lmgc(lev - 1, 1);                          // followed by a V-cycle
    } else {
        for (int i = 0; i < type; ++i) lmgc(lev - 1, type);
    }
    prolongation_and_postsmooth(lev);
}

void GMG::presmooth_and_restriction(int lev) {
    // mg_solver_impl.hpp:1686-1816
    auto& lsc = sc[lev];
    auto& lsd = sd[lev];
    auto& lst = st[lev];
    for (int nu = 0; nu < nu1; ++nu) {
        smoother[lev].apply(lst, lsd);
        apply_sub(A[lev], lsd, lst);
        if (nu < nu1 - 1)
            for (size_t i = 0; i < lsc.size(); ++i) lsc[i] += lst[i];
    }
    std::fill(sc[lev - 1].begin(), sc[lev - 1].end(), 0.0);
    if (nu1 > 0)
        for (size_t i = 0; i < lsc.size(); ++i) lsc[i] += lst[i];
    // StdTransfer::do_restrict (std_transfer_impl.h:774-806)
    apply_ignore_zero_rows(R[lev], sd[lev - 1], 1.0, lsd);
}

void GMG::prolongation_and_postsmooth(int lev) {
    // mg_solver_impl.hpp:1819-1964
    auto& lsc = sc[lev];
    auto& lsd = sd[lev];
    auto& lst = st[lev];
    // StdTransfer::prolongate (std_transfer_impl.h:719-772): axpy with alpha=0
    axpy(P[lev], lst, 0.0, lst, 1.0, sc[lev - 1]);
    for (size_t i = 0; i < lsc.size(); ++i) lsc[i] += lst[i];
    for (int nu = 0; nu < nu2; ++nu) {
        apply_sub(A[lev], lsd, lst);
        smoother[lev].apply(lst, lsd);
        for (size_t i = 0; i < lsc.size(); ++i) lsc[i] += lst[i];
    }
    if (lev >= full_ref_level()) apply_sub(A[lev], lsd, lst);
}

void GMG::base_solve(int lev) {
    // mg_solver_impl.hpp:1968-2086 (distributed/serial branch a)
    base.apply(sc[lev], sd[lev]);
    if (lev >= full_ref_level()) apply_sub(A[lev], sd[lev], sc[lev]);
}

// ---- CG (cg.h:103-242) ----------------------------------------------------------

SolveResult cg_solve(const CSR& A, GMG* precond, std::vector<double>& x, const std::vector<double>& b,
                     ConvCheck cc, bool record_x) {
    SolveResult res;
    res.record_x = record_x;
    auto rec = [&]() { if (record_x) res.xs.push_back(x); };
    std::vector<double> r = b;
    apply_sub(A, r, x);
    std::vector<double> q(r.size(), 0.0), z(r.size(), 0.0), p(r.size(), 0.0);
    if (precond) precond->apply(z, r); else z = r;
    cc.start(norm(r));
    rec();
    p = z;
    double rho_old = dot(z, r), rho;
    while (!cc.iteration_ended()) {
        axpy(A, q, 0.0, q, 1.0, p);
        double lambda = dot(q, p);
        if (lambda == 0.0) { res.ok = false; res.steps = cc.step; res.defects = cc.defects; return res; }
        const double alpha = rho_old / lambda;
        for (size_t i = 0; i < x.size(); ++i) x[i] = 1.0 * x[i] + alpha * p[i];
        for (size_t i = 0; i < r.size(); ++i) r[i] = 1.0 * r[i] + (-alpha) * q[i];
        cc.update(norm(r));
        rec();
        if (cc.iteration_ended()) break;
        if (precond) precond->apply(z, r); else z = r;
        rho = dot(z, r);
        const double beta = rho / rho_old;
        for (size_t i = 0; i < p.size(); ++i) p[i] = beta * p[i] + 1.0 * z[i];
        rho_old = rho;
    }
    res.ok = cc.post();
    res.steps = cc.step;
    res.defects = cc.defects;
    return res;
}

// ---- BiCGStab (bicgstab.h:112-385); restarts every opt.num_restarts steps and
// when |(r0, r)| <= opt.min_ortho ||r|| ||r0|| (bicgstab.h:157-215) -------------------

SolveResult bicgstab_solve(const CSR& A, GMG* precond, std::vector<double>& x, const std::vector<double>& b,
                           ConvCheck cc, BiCGStabOptions opt, bool record_x) {
    SolveResult res;
    res.record_x = record_x;
    auto rec = [&]() { if (record_x) res.xs.push_back(x); };
    const size_t n = b.size();
    std::vector<double> r = b;
    apply_sub(A, r, x);
    std::vector<double> r0(n), p(n), v(n), t(n), s(n), q(n);
    cc.start(norm(r));
    rec();
    double rho = 1, alpha = 1, omega = 1, norm_r0 = 0.0;
    bool restart = true;
    auto fail = [&]() { res.ok = false; res.steps = cc.step; res.defects = cc.defects; return res; };
    while (!cc.iteration_ended()) {
        if (opt.num_restarts > 0 && cc.step % opt.num_restarts == 0) restart = true;
        if (restart) {
            r0 = r;
            std::fill(p.begin(), p.end(), 0.0);
            alpha = 0.0;
            std::fill(v.begin(), v.end(), 0.0);
            omega = 1.0;
            rho = 1.0;
            norm_r0 = cc.current;
            restart = false;
        }
        const double rho_old = rho;
        rho = dot(r0, r);
        const double norm_r = cc.current;
        if (std::fabs(rho) / (norm_r * norm_r0) <= opt.min_ortho) restart = true;
        if (rho_old == 0.0) return fail();
        const double beta = (rho / rho_old) * (alpha / omega);
        for (size_t i = 0; i < n; ++i) p[i] = 1.0 * r[i] + beta * p[i] + (-beta * omega) * v[i];
        if (precond) precond->apply(q, p); else q = p;
        axpy(A, v, 0.0, v, 1.0, q);
        alpha = dot(v, r0);
        if (alpha == 0.0) return fail();
        alpha = rho / alpha;
        for (size_t i = 0; i < n; ++i) x[i] = 1.0 * x[i] + alpha * q[i];
        for (size_t i = 0; i < n; ++i) s[i] = 1.0 * r[i] + (-alpha) * v[i];
        cc.update(norm(s));
        rec();
        if (cc.iteration_ended()) { r = s; break; }
        if (precond) precond->apply(q, s); else q = s;
        axpy(A, t, 0.0, t, 1.0, q);
        const double tt = dot(t, t);
        omega = dot(s, t);
        if (tt == 0.0) return fail();
        omega = omega / tt;
        for (size_t i = 0; i < n; ++i) x[i] = 1.0 * x[i] + omega * q[i];
        for (size_t i = 0; i < n; ++i) r[i] = 1.0 * s[i] + (-omega) * t[i];
        cc.update(norm(r));
        rec();
        if (omega == 0.0) return fail();
    }
    res.ok = cc.post();
    res.steps = cc.step;
    res.defects = cc.defects;
    return res;
}

}  // namespace orc
