// Oracle (test infrastructure only): element discretisations, Dirichlet
// constraints and grid transfer.
//
// The ConvectionDiffusion element discretisation is an external plugin (not
// part of ugcore); the element matrices here follow its documented operator
// built from ugcore's building blocks:
//  * Q1 FE: LagrangeP1 shape functions on the reference hexahedron
//    (local_finite_element/lagrange/lagrangep1.h:52-96,
//    reference_element/reference_element.h ReferenceHexahedron corners),
//    trilinear reference mapping (reference_mapping.h:1099), 2x2x2 Gauss rule
//    (quadrature/gauss/gauss_quad_hexahedron.cpp, order 3).
//  * FV1: box method on P1 triangles; scvf from edge midpoint to barycentre
//    (disc_util/fv1_geom.cpp:258-463), normal oriented from `from` to `to`
//    (fv_util.h:194-262); diffusion flux k grad(u).n subtracted from `from`
//    and added to `to`; source f(corner)*|scv| (fv1_geom.h:311-314).
//  * Scatter: AddLocalMatrixToGlobal (lib_disc/common/local_algebra.h:766-789)
//    through SparseMatrix::operator() insertion semantics.
//  * Dirichlet: SetDirichletRow (algebra_common/sparsematrix_util.h:850-861),
//    rhs value 0 (lagrange_dirichlet_boundary_impl.h:753-1035).
#include <algorithm>
#include <climits>
#include <cmath>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

namespace {

const double REF_HEX[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

void q1_shape(const double* xi, double* phi, double (*dphi)[3]) {
    for (int i = 0; i < 8; ++i) {
        double f[3], g[3];
        for (int d = 0; d < 3; ++d) {
            f[d] = REF_HEX[i][d] > 0.5 ? xi[d] : 1.0 - xi[d];
            g[d] = REF_HEX[i][d] > 0.5 ? 1.0 : -1.0;
        }
        phi[i] = f[0] * f[1] * f[2];
        dphi[i][0] = g[0] * f[1] * f[2];
        dphi[i][1] = f[0] * g[1] * f[2];
        dphi[i][2] = f[0] * f[1] * g[2];
    }
}

double inv3(const double J[3][3], double Ji[3][3]) {
    double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                 J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                 J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double id = 1.0 / det;
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    return det;
}

}  // namespace

std::vector<uint8_t> dirichlet_mask(const MultiGrid& mg, int lev) {
    const Level& L = mg.levels[lev];
    std::vector<uint8_t> m(L.n_dof, 0);
    for (size_t v = 0; v < L.vpos.size(); ++v)
        if (L.vsub[v] == 1) m[L.dof[v]] = 1;
    return m;
}

// Q1 element matrix and load vector of hexahedron `e` on level `lev`
// (2x2x2 Gauss, trilinear geometry, k at each quadrature point)
static void q1_element(const MultiGrid& mg, const Level& L, size_t e, const Coefficient& k, double f_const,
                       double Ke[8][8], double fe[8], int ind[8]) {
    const double scale = 1.0 / (double)mg.finest_scale;
    const double g = 0.5 / std::sqrt(3.0);
    const double qp1[2] = {0.5 - g, 0.5 + g};
    double X[8][3];
    for (int i = 0; i < 8; ++i) {
        int v = L.vols[e][i];
        ind[i] = L.dof[v];
        for (int d = 0; d < 3; ++d) X[i][d] = (double)L.vpos[v][d] * scale;
    }
    for (int i = 0; i < 8; ++i) {
        fe[i] = 0.0;
        for (int j = 0; j < 8; ++j) Ke[i][j] = 0.0;
    }
    for (int qz = 0; qz < 2; ++qz)
        for (int qy = 0; qy < 2; ++qy)
            for (int qx = 0; qx < 2; ++qx) {
                const double xi[3] = {qp1[qx], qp1[qy], qp1[qz]};
                double phi[8], dphi[8][3];
                q1_shape(xi, phi, dphi);
                double J[3][3] = {{0}}, Ji[3][3], xq[3] = {0, 0, 0};
                for (int i = 0; i < 8; ++i)
                    for (int r = 0; r < 3; ++r) {
                        xq[r] += phi[i] * X[i][r];
                        for (int c = 0; c < 3; ++c) J[r][c] += X[i][r] * dphi[i][c];
                    }
                const double det = inv3(J, Ji);
                const double w = 0.125 * std::fabs(det);
                const double kq = k.eval(xq);
                double grad[8][3];
                for (int i = 0; i < 8; ++i)
                    for (int r = 0; r < 3; ++r)
                        grad[i][r] = Ji[0][r] * dphi[i][0] + Ji[1][r] * dphi[i][1] + Ji[2][r] * dphi[i][2];
                for (int i = 0; i < 8; ++i) {
                    for (int j = 0; j < 8; ++j)
                        Ke[i][j] += w * kq * (grad[i][0] * grad[j][0] + grad[i][1] * grad[j][1] + grad[i][2] * grad[j][2]);
                    fe[i] += w * f_const * phi[i];
                }
            }
}

void assemble_q1_hex(const MultiGrid& mg, int lev, const Coefficient& k, double f_const,
                     CSR& A, std::vector<double>& b) {
    const Level& L = mg.levels[lev];
    RowBuilder B;
    B.resize_and_clear(L.n_dof, L.n_dof);
    b.assign(L.n_dof, 0.0);
    for (size_t e = 0; e < L.vols.size(); ++e) {
        if (L.volsub[e] != 0) continue;   // element disc defined on subset "Inner"
        double Ke[8][8], fe[8];
        int ind[8];
        q1_element(mg, L, e, k, f_const, Ke, fe, ind);
        for (int i = 0; i < 8; ++i) {
            for (int j = 0; j < 8; ++j) B.at(ind[i], ind[j]) += Ke[i][j];
            b[ind[i]] += fe[i];
        }
    }
    A = defragment(B);
    auto dm = dirichlet_mask(mg, lev);
    for (int64_t r = 0; r < A.n_rows; ++r)
        if (dm[r]) {
            set_row(A, r, 0.0);
            double* d = find_entry(A, r, r);
            if (!d) throw std::runtime_error("oracle: Dirichlet row without diagonal");
            *d = 1.0;
            b[r] = 0.0;
        }
}

void adjoint_gradient_q1_hex(const MultiGrid& mg, const Coefficient& k, const std::vector<double>& u,
                             const std::vector<double>& lambda, std::vector<double>& g) {
    // dJ/dtheta_e = -lambda^T (dA/dtheta_e) u. On the fine level k is
    // exp(theta_e) at every quadrature point of element e, so dA_e/dtheta_e =
    // A_e; Dirichlet rows of A do not depend on theta.
    const int top = (int)mg.levels.size() - 1;
    const Level& L = mg.levels[top];
    const int64_t n = mg.finest_scale;
    g.assign(n * n * n, 0.0);
    auto dm = dirichlet_mask(mg, top);
    for (size_t e = 0; e < L.vols.size(); ++e) {
        double Ke[8][8], fe[8];
        int ind[8];
        q1_element(mg, L, e, k, 0.0, Ke, fe, ind);
        int64_t mn[3] = {INT64_MAX, INT64_MAX, INT64_MAX};
        for (int i = 0; i < 8; ++i)
            for (int d = 0; d < 3; ++d) mn[d] = std::min(mn[d], L.vpos[L.vols[e][i]][d]);
        double acc = 0.0;
        for (int a = 0; a < 8; ++a) {
            if (dm[ind[a]]) continue;
            double row = 0.0;
            for (int b2 = 0; b2 < 8; ++b2) row += Ke[a][b2] * u[ind[b2]];
            acc += lambda[ind[a]] * row;
        }
        g[(mn[2] * n + mn[1]) * n + mn[0]] = -acc;
    }
}

// ---------------------------------------------------------------------------
// Q2 (triquadratic Lagrange) hexahedron element matrices.
// Local shape order: lagrange.cpp:236-253 (SetLagrangeMultiIndex) applied to
// the reference hexahedron of reference_element.cpp (corners, EDGE->POINT and
// FACE->POINT tables): vertices p*corner (lagrange.cpp:40-62), edge DoFs
// i*co1 + (p-i)*co0 (:68-105), face DoFs p*co0 + i*(co1-co0) + j*(co3-co0)
// (:110-160), the interior (:167-230). Shape i is the product of the 1D
// equidistant Lagrange polynomials of its multi-index on the nodes {0, 1/2, 1}
// (lagrange.h:2106-2190, EquidistantLagrange1D). Geometry: the trilinear map
// of the 8 corners; 3x3x3 Gauss-Legendre points (exact for the Q2 stiffness
// of a parallelepiped); k constant per element.
// ---------------------------------------------------------------------------
void q2_hex_multi_index(int mi[27][3]) {
    static const int co[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                 {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    static const int edge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                                    {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
    static const int face[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                   {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};
    const int p = 2;
    int n = 0;
    for (int v = 0; v < 8; ++v, ++n)
        for (int d = 0; d < 3; ++d) mi[n][d] = p * co[v][d];
    for (int e = 0; e < 12; ++e)
        for (int i = 1; i < p; ++i, ++n)
            for (int d = 0; d < 3; ++d) mi[n][d] = i * co[edge[e][1]][d] + (p - i) * co[edge[e][0]][d];
    for (int f = 0; f < 6; ++f) {
        const int c0 = face[f][0], c1 = face[f][1], c2 = face[f][3];   // quadrilateral: co2 := corner 3
        for (int j = 1; j < p; ++j)
            for (int i = 1; i < p; ++i, ++n)
                for (int d = 0; d < 3; ++d)
                    mi[n][d] = p * co[c0][d] + i * (co[c1][d] - co[c0][d]) + j * (co[c2][d] - co[c0][d]);
    }
    for (int m2 = 1; m2 < p; ++m2)
        for (int m1 = 1; m1 < p; ++m1)
            for (int m0 = 1; m0 < p; ++m0, ++n) {
                mi[n][0] = m0;
                mi[n][1] = m1;
                mi[n][2] = m2;
            }
}

static double lagrange1d(int i, double x) {
    // Lagrange polynomial of node i/2 on the nodes {0, 1/2, 1}
    double v = 1.0;
    for (int m = 0; m < 3; ++m)
        if (m != i) v *= (x - 0.5 * m) / (0.5 * i - 0.5 * m);
    return v;
}
static double dlagrange1d(int i, double x) {
    double s = 0.0;
    for (int a = 0; a < 3; ++a) {
        if (a == i) continue;
        double t = 1.0 / (0.5 * i - 0.5 * a);
        for (int m = 0; m < 3; ++m)
            if (m != i && m != a) t *= (x - 0.5 * m) / (0.5 * i - 0.5 * m);
        s += t;
    }
    return s;
}

// the element loop of both variants: kq(xq) is the coefficient at the
// quadrature point xq
template <typename KFn>
static void q2_hex_element_impl(const double X[8][3], KFn kq_at, double f_const, double Ke[27][27], double fe[27]) {
    int mi[27][3];
    q2_hex_multi_index(mi);
    const double r = 0.5 * std::sqrt(0.6);
    const double gx[3] = {0.5 - r, 0.5, 0.5 + r}, gw[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
    for (int a = 0; a < 27; ++a) {
        fe[a] = 0.0;
        for (int b = 0; b < 27; ++b) Ke[a][b] = 0.0;
    }
    for (int qz = 0; qz < 3; ++qz)
        for (int qy = 0; qy < 3; ++qy)
            for (int qx = 0; qx < 3; ++qx) {
                const double xi[3] = {gx[qx], gx[qy], gx[qz]};
                const double w = gw[qx] * gw[qy] * gw[qz];
                double phi1[8], dphi1[8][3];
                q1_shape(xi, phi1, dphi1);
                double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (int a = 0; a < 8; ++a)
                    for (int rr = 0; rr < 3; ++rr)
                        for (int c = 0; c < 3; ++c) J[rr][c] += X[a][rr] * dphi1[a][c];
                double Ji[3][3];
                const double det = inv3(J, Ji);
                const double wd = w * std::fabs(det);
                double xq[3] = {0.0, 0.0, 0.0};
                for (int a = 0; a < 8; ++a)
                    for (int rr = 0; rr < 3; ++rr) xq[rr] += phi1[a] * X[a][rr];
                const double k = kq_at(xq);
                double g[27][3], ph[27];
                for (int a = 0; a < 27; ++a) {
                    double dref[3];
                    for (int d = 0; d < 3; ++d) {
                        dref[d] = dlagrange1d(mi[a][d], xi[d]);
                        for (int d2 = 0; d2 < 3; ++d2)
                            if (d2 != d) dref[d] *= lagrange1d(mi[a][d2], xi[d2]);
                    }
                    ph[a] = lagrange1d(mi[a][0], xi[0]) * lagrange1d(mi[a][1], xi[1]) * lagrange1d(mi[a][2], xi[2]);
                    for (int rr = 0; rr < 3; ++rr) g[a][rr] = Ji[0][rr] * dref[0] + Ji[1][rr] * dref[1] + Ji[2][rr] * dref[2];
                }
                for (int a = 0; a < 27; ++a) {
                    for (int b = 0; b < 27; ++b)
                        Ke[a][b] += wd * k * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
                    fe[a] += wd * f_const * ph[a];
                }
            }
}

void q2_hex_element(const double X[8][3], double k, double f_const, double Ke[27][27], double fe[27]) {
    q2_hex_element_impl(X, [k](const double*) { return k; }, f_const, Ke, fe);
}

// The 3-point Gauss rule puts its middle point on a fine-element boundary
// whenever an element spans an even number of fine elements, where a
// piecewise-constant coefficient has two values. Such a point takes the fine
// element on its upper side: xq is snapped to a 2^-40 grid first (the
// boundaries are dyadic, the other Gauss points lie far from them), so that
// the rounding of the trilinear map cannot decide it.
void q2_hex_element_coef(const double X[8][3], const Coefficient& k, double f_const, double Ke[27][27],
                         double fe[27]) {
    q2_hex_element_impl(
        X,
        [&k](const double* xq) {
            const double g = 1099511627776.0;   // 2^40
            const double xs[3] = {std::round(xq[0] * g) / g, std::round(xq[1] * g) / g, std::round(xq[2] * g) / g};
            return k.eval(xs);
        },
        f_const, Ke, fe);
}

void assemble_fv1_tri(const MultiGrid& mg, int lev, const Coefficient& k,
                      const std::vector<double>& f_at_dof, CSR& A, std::vector<double>& b) {
    const Level& L = mg.levels[lev];
    const double scale = 1.0 / (double)mg.finest_scale;
    RowBuilder B;
    B.resize_and_clear(L.n_dof, L.n_dof);
    b.assign(L.n_dof, 0.0);
    for (size_t f = 0; f < L.faces.size(); ++f) {
        if (L.fsub[f] != 0 || L.fnv[f] != 3) continue;
        double X[3][2];
        int ind[3];
        for (int i = 0; i < 3; ++i) {
            int v = L.faces[f][i];
            ind[i] = L.dof[v];
            X[i][0] = (double)L.vpos[v][0] * scale;
            X[i][1] = (double)L.vpos[v][1] * scale;
        }
        // P1 gradients: grad phi_i = J^{-T} grad_ref phi_i
        const double J00 = X[1][0] - X[0][0], J01 = X[2][0] - X[0][0];
        const double J10 = X[1][1] - X[0][1], J11 = X[2][1] - X[0][1];
        const double det = J00 * J11 - J01 * J10;
        const double area = 0.5 * std::fabs(det);
        const double gr[3][2] = {{-1, -1}, {1, 0}, {0, 1}};
        double grad[3][2];
        for (int i = 0; i < 3; ++i) {
            grad[i][0] = (J11 * gr[i][0] - J10 * gr[i][1]) / det;
            grad[i][1] = (-J01 * gr[i][0] + J00 * gr[i][1]) / det;
        }
        const double bary[2] = {(X[0][0] + X[1][0] + X[2][0]) / 3.0, (X[0][1] + X[1][1] + X[2][1]) / 3.0};
        double Ke[3][3] = {{0}};
        for (int e = 0; e < 3; ++e) {
            const int from = e, to = (e + 1) % 3;
            const double m[2] = {0.5 * (X[from][0] + X[to][0]), 0.5 * (X[from][1] + X[to][1])};
            // scvf segment midpoint->barycentre, normal = rotated segment, oriented from->to
            double n[2] = {bary[1] - m[1], -(bary[0] - m[0])};
            const double dir[2] = {X[to][0] - X[from][0], X[to][1] - X[from][1]};
            if (n[0] * dir[0] + n[1] * dir[1] < 0) { n[0] = -n[0]; n[1] = -n[1]; }
            const double ip[2] = {0.5 * (m[0] + bary[0]), 0.5 * (m[1] + bary[1])};
            const double kq = k.eval(ip);
            for (int sh = 0; sh < 3; ++sh) {
                const double flux = kq * (grad[sh][0] * n[0] + grad[sh][1] * n[1]);
                Ke[from][sh] -= flux;
                Ke[to][sh] += flux;
            }
        }
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) B.at(ind[i], ind[j]) += Ke[i][j];
            b[ind[i]] += f_at_dof[ind[i]] * (area / 3.0);
        }
    }
    A = defragment(B);
    auto dm = dirichlet_mask(mg, lev);
    for (int64_t r = 0; r < A.n_rows; ++r)
        if (dm[r]) {
            set_row(A, r, 0.0);
            double* d = find_entry(A, r, r);
            if (!d) throw std::runtime_error("oracle: Dirichlet row without diagonal");
            *d = 1.0;
            b[r] = 0.0;
        }
}

// ---------------------------------------------------------------------------
// FV1 on tetrahedra (fv1_geom.cpp:258-463 with the ReferenceTetrahedron tables
// of reference_element.cpp:547-760): one scvf per edge i with corners
// [edge midpoint, centre of face id(1,i,2,0), element centre, centre of face
// id(1,i,2,1)] (fv1_geom.cpp:118-150), normal = ElementNormal of that
// quadrilateral, 0.5*(c2-c0)x(c3-c1) (geometry_util.h:579-586), oriented from
// edge vertex 0 ("from") to vertex 1 ("to"); integration point = average of
// the four corners. scv of corner i: hexahedron [corner, mid(edge id(0,i,1,1)),
// face id(0,i,2,0), mid(edge id(0,i,1,0)), mid(edge id(0,i,1,2)), face
// id(0,i,2,2), centre, face id(0,i,2,1)] (fv1_geom.cpp:156-198), volume by
// ElementSize<ReferenceHexahedron> (geometry_util.h, two prisms).
// Diffusion flux k grad(phi_j).n subtracted from `from`, added to `to`;
// full upwind convection (conv_shape.h:278-330): up = flux >= 0 ? from : to,
// shape(up) = v.n, added to `from`, subtracted from `to`.
// ---------------------------------------------------------------------------
namespace {

const int RT_EDGE[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
const int RT_FACE[4][3] = {{0, 2, 1}, {1, 2, 3}, {0, 3, 2}, {0, 1, 3}};
const int RT_EDGE_FACE[6][2] = {{0, 3}, {0, 1}, {0, 2}, {3, 2}, {1, 3}, {2, 1}};
const int RT_POINT_EDGE[4][3] = {{2, 0, 3}, {0, 1, 4}, {1, 2, 5}, {3, 4, 5}};
const int RT_POINT_FACE[4][3] = {{0, 2, 3}, {0, 3, 1}, {0, 1, 2}, {3, 2, 1}};

void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

double hex_volume(const double (*x)[3]) {
    double x50[3], x14[3], x10[3], x20[3], x60[3], x24[3], x30[3], x70[3], m1[3], n1[3], m2[3], n2[3];
    for (int d = 0; d < 3; ++d) {
        x50[d] = x[5][d] - x[0][d];
        x14[d] = x[1][d] - x[4][d];
        x10[d] = x[1][d] - x[0][d];
        x20[d] = x[2][d] - x[0][d];
        x60[d] = x[6][d] - x[0][d];
        x24[d] = x[2][d] - x[4][d];
        x30[d] = x[3][d] - x[0][d];
        x70[d] = x[7][d] - x[0][d];
    }
    cross3(x50, x14, m1);
    cross3(x10, x20, n1);
    cross3(x60, x24, m2);
    cross3(x20, x30, n2);
    double s1 = 0, s2 = 0;
    for (int d = 0; d < 3; ++d) {
        s1 += (n1[d] + m1[d]) * x60[d];
        s2 += (n2[d] + m2[d]) * x70[d];
    }
    return (1.0 / 6.0) * (s1 + s2);
}

// element matrix (diffusion with kip[scvf] + convection vel; conv=false drops
// the convection) and scv volumes of a tetrahedron with corners X
void fv1_tet_element(const double (*X)[3], const double* kip, const double* vel, bool conv, double Ke[4][4],
                     double vol[4]) {
    double M[6][3], F[4][3], C[3];
    for (int e = 0; e < 6; ++e)
        for (int d = 0; d < 3; ++d) M[e][d] = (X[RT_EDGE[e][0]][d] + X[RT_EDGE[e][1]][d]) * 0.5;
    for (int f = 0; f < 4; ++f)
        for (int d = 0; d < 3; ++d) F[f][d] = (X[RT_FACE[f][0]][d] + X[RT_FACE[f][1]][d] + X[RT_FACE[f][2]][d]) / 3.0;
    for (int d = 0; d < 3; ++d) C[d] = (X[0][d] + X[1][d] + X[2][d] + X[3][d]) * 0.25;
    // P1 gradients: grad phi_c = row c-1 of J^{-1} (c = 1..3), grad phi_0 = -sum
    double J[3][3], Ji[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) J[r][c] = X[c + 1][r] - X[0][r];
    inv3(J, Ji);
    double grad[4][3];
    for (int c = 1; c < 4; ++c)
        for (int d = 0; d < 3; ++d) grad[c][d] = Ji[c - 1][d];
    for (int d = 0; d < 3; ++d) grad[0][d] = -(grad[1][d] + grad[2][d] + grad[3][d]);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Ke[i][j] = 0.0;
    for (int i = 0; i < 6; ++i) {
        const double* c0 = M[i];
        const double* c1 = F[RT_EDGE_FACE[i][0]];
        const double* c3 = F[RT_EDGE_FACE[i][1]];
        double a[3], b[3], n[3];
        for (int d = 0; d < 3; ++d) {
            a[d] = C[d] - c0[d];
            b[d] = c3[d] - c1[d];
        }
        cross3(a, b, n);
        for (int d = 0; d < 3; ++d) n[d] *= 0.5;
        const int from = RT_EDGE[i][0], to = RT_EDGE[i][1];
        for (int sh = 0; sh < 4; ++sh) {
            const double flux = kip[i] * (grad[sh][0] * n[0] + grad[sh][1] * n[1] + grad[sh][2] * n[2]);
            Ke[from][sh] -= flux;
            Ke[to][sh] += flux;
        }
        if (conv) {
            const double q = vel[0] * n[0] + vel[1] * n[1] + vel[2] * n[2];
            const int up = q >= 0.0 ? from : to;
            Ke[from][up] += q;
            Ke[to][up] -= q;
        }
    }
    for (int i = 0; i < 4; ++i) {
        double H[8][3];
        const int* pe = RT_POINT_EDGE[i];
        const int* pf = RT_POINT_FACE[i];
        const double* src[8] = {X[i], M[pe[1]], F[pf[0]], M[pe[0]], M[pe[2]], F[pf[2]], C, F[pf[1]]};
        for (int q = 0; q < 8; ++q)
            for (int d = 0; d < 3; ++d) H[q][d] = src[q][d];
        vol[i] = hex_volume(H);
    }
}

// fine cell containing the integration point of scvf i of tetrahedron t on
// level L (lattice positions, exact integer arithmetic: 48 * ip)
int64_t tet_ip_cell(const Level& L, size_t t, int i, int64_t n_cells) {
    int64_t c[3];
    for (int d = 0; d < 3; ++d) {
        auto P = [&](int q) { return L.vpos[L.vols[t][q]][d]; };
        const int64_t sall = P(0) + P(1) + P(2) + P(3);
        const int* f0 = RT_FACE[RT_EDGE_FACE[i][0]];
        const int* f1 = RT_FACE[RT_EDGE_FACE[i][1]];
        const int64_t s48 = 6 * (P(RT_EDGE[i][0]) + P(RT_EDGE[i][1])) + 4 * (P(f0[0]) + P(f0[1]) + P(f0[2])) +
                            3 * sall + 4 * (P(f1[0]) + P(f1[1]) + P(f1[2]));
        int64_t cc = s48 / 48;
        if (cc >= n_cells) cc = n_cells - 1;
        c[d] = cc;
    }
    return (c[2] * n_cells + c[1]) * n_cells + c[0];
}

}  // namespace

void tet_vertex_position(const MultiGrid& mg, int lev, int v, const TetParams& tp, double x[3]) {
    const Level& L = mg.levels[lev];
    const double scale = 1.0 / (double)mg.finest_scale;
    for (int d = 0; d < 3; ++d) x[d] = (double)L.vpos[v][d] * scale;
    const int top = (int)mg.levels.size() - 1;
    if (lev != top || tp.jitter_amp == 0.0 || L.vsub[v] != 0) return;
    const int64_t W = mg.finest_scale + 1;
    const int64_t l = (L.vpos[v][2] * W + L.vpos[v][1]) * W + L.vpos[v][0];
    for (int d = 0; d < 3; ++d) {
        const double u = uniform_from_counter(tp.jitter_seed, (uint64_t)(3 * l + d));
        x[d] += tp.jitter_amp * scale * (2.0 * u - 1.0);
    }
}

void assemble_fv1_tet(const MultiGrid& mg, int lev, const std::vector<double>& theta_cells, const TetParams& tp,
                      double f_const, CSR& A, std::vector<double>& b) {
    const Level& L = mg.levels[lev];
    const int64_t n_cells = mg.finest_scale;
    RowBuilder B;
    B.resize_and_clear(L.n_dof, L.n_dof);
    b.assign(L.n_dof, 0.0);
    for (size_t t = 0; t < L.vols.size(); ++t) {
        if (L.volsub[t] != 0 || L.volnv[t] != 4) continue;
        double X[4][3], kip[6], Ke[4][4], vol[4];
        int ind[4];
        for (int q = 0; q < 4; ++q) {
            tet_vertex_position(mg, lev, L.vols[t][q], tp, X[q]);
            ind[q] = L.dof[L.vols[t][q]];
        }
        for (int i = 0; i < 6; ++i)
            kip[i] = theta_cells.empty() ? 1.0 : std::exp(theta_cells[tet_ip_cell(L, t, i, n_cells)]);
        fv1_tet_element(X, kip, tp.vel, true, Ke, vol);
        for (int i = 0; i < 4; ++i) {
            for (int j = 0; j < 4; ++j) B.at(ind[i], ind[j]) += Ke[i][j];
            b[ind[i]] += f_const * vol[i];
        }
    }
    A = defragment(B);
    auto dm = dirichlet_mask(mg, lev);
    for (int64_t r = 0; r < A.n_rows; ++r)
        if (dm[r]) {
            set_row(A, r, 0.0);
            double* d = find_entry(A, r, r);
            if (!d) throw std::runtime_error("oracle: Dirichlet row without diagonal");
            *d = 1.0;
            b[r] = 0.0;
        }
}

void adjoint_gradient_fv1_tet(const MultiGrid& mg, const std::vector<double>& theta_cells, const TetParams& tp,
                              const std::vector<double>& u, const std::vector<double>& lambda,
                              std::vector<double>& g) {
    // on the top level every scvf integration point of a tetrahedron lies in
    // its own fine cell, so dA_e/dtheta_cell = the diffusion part of A_e
    const int top = (int)mg.levels.size() - 1;
    const Level& L = mg.levels[top];
    const int64_t n = mg.finest_scale;
    g.assign(n * n * n, 0.0);
    auto dm = dirichlet_mask(mg, top);
    for (size_t t = 0; t < L.vols.size(); ++t) {
        double X[4][3], kip[6], Ke[4][4], vol[4];
        int ind[4];
        for (int q = 0; q < 4; ++q) {
            tet_vertex_position(mg, top, L.vols[t][q], tp, X[q]);
            ind[q] = L.dof[L.vols[t][q]];
        }
        const int64_t cell = tet_ip_cell(L, t, 0, n);
        const double k = theta_cells.empty() ? 1.0 : std::exp(theta_cells[cell]);
        for (int i = 0; i < 6; ++i) kip[i] = k;
        fv1_tet_element(X, kip, tp.vel, false, Ke, vol);
        double acc = 0.0;
        for (int a = 0; a < 4; ++a) {
            if (dm[ind[a]]) continue;
            double row = 0.0;
            for (int b2 = 0; b2 < 4; ++b2) row += Ke[a][b2] * u[ind[b2]];
            acc += lambda[ind[a]] * row;
        }
        g[cell] -= acc;
    }
}

CSR prolongation_p1(const MultiGrid& mg, int fine_lev) {
    const Level& F = mg.levels[fine_lev];
    const Level& C = mg.levels[fine_lev - 1];
    RowBuilder B;
    B.resize_and_clear(F.n_dof, C.n_dof);
    for (int si = 0; si < mg.num_subsets; ++si)
        for (size_t v = 0; v < F.vpos.size(); ++v) {
            if (F.vsub[v] != si) continue;
            const int row = F.dof[v];
            const int p = F.vparent[v];
            switch (F.vparent_type[v]) {
                case OT_VERTEX: B.at(row, C.dof[p]) = 1.0; break;
                case OT_EDGE:
                    for (int i = 0; i < 2; ++i) B.at(row, C.dof[C.edges[p][i]]) = 0.5;
                    break;
                case OT_FACE:
                    if (C.fnv[p] != 4) throw std::runtime_error("oracle: face parent must be a quadrilateral");
                    for (int i = 0; i < 4; ++i) B.at(row, C.dof[C.faces[p][i]]) = 0.25;
                    break;
                case OT_VOLUME:
                    for (int i = 0; i < 8; ++i) B.at(row, C.dof[C.vols[p][i]]) = 0.125;
                    break;
                default: throw std::runtime_error("oracle: vertex without parent");
            }
        }
    CSR P = defragment(B);
    // adjust_prolongation: fine Dirichlet rows zeroed, vertex-copy children keep
    // the injection 1 at their parent
    for (size_t v = 0; v < F.vpos.size(); ++v) {
        if (F.vsub[v] != 1) continue;
        const int row = F.dof[v];
        set_row(P, row, 0.0);
        if (F.vparent_type[v] == OT_VERTEX) *find_entry(P, row, C.dof[F.vparent[v]]) = 1.0;
    }
    return P;
}

CSR restriction_from(const MultiGrid& mg, int fine_lev, const CSR& P) {
    const Level& F = mg.levels[fine_lev];
    const Level& C = mg.levels[fine_lev - 1];
    CSR R = transpose_keep_zeros(P);
    for (size_t v = 0; v < F.vpos.size(); ++v) {
        if (F.vsub[v] != 1 || F.vparent_type[v] != OT_VERTEX) continue;
        const int crow = C.dof[F.vparent[v]];
        set_row(R, crow, 0.0);
        *find_entry(R, crow, F.dof[v]) = 1.0;
    }
    return R;
}

}  // namespace orc
