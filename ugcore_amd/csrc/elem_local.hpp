// Local (element) matrices shared by the lattice assembly kernels and the
// element-loop assembly from caller-supplied connectivity (device code,
// included by the .hip files that need it).
//
//   * q1_hex_element: Q1 Lagrange FE on a hexahedron with trilinear geometry,
//     2x2x2 Gauss (LagrangeP1 on the reference hexahedron, lagrangep1.h:52-96;
//     FEGeometry, fe_geom.h:49-474), -div(k grad u) = f;
//   * fv1_tri_element: the FV1 box method on a P1 triangle (FV1Geometry::update,
//     fv1_geom.cpp:329-463; NormalOnSCVF, fv_util.h:194-262);
//   * fv1_tet_element: the FV1 box method on a P1 tetrahedron with optional
//     full-upwind convection (ConvectionShapesFullUpwind::update,
//     conv_shape.h:278-330).
// The coefficient k is given per integration point by the caller.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ugamd {
namespace {

// Q1 element matrix and load vector; local vertex order of the reference
// hexahedron (x fastest, bottom then top); kq(xq) is the coefficient at the
// quadrature point xq; fe = f_const * int phi_i
template <typename KFn>
__device__ inline void q1_hex_element_k(const double X[8][3], KFn kq_at, double f_const, double Ke[8][8],
                                        double fe[8]) {
    const double g = 0.28867513459481287;   // 0.5/sqrt(3)
    for (int a = 0; a < 8; ++a) {
        fe[a] = 0.0;
        for (int b = 0; b < 8; ++b) Ke[a][b] = 0.0;
    }
    for (int q = 0; q < 8; ++q) {
        const double xi[3] = {(q & 1) ? 0.5 + g : 0.5 - g, (q & 2) ? 0.5 + g : 0.5 - g, (q & 4) ? 0.5 + g : 0.5 - g};
        double phi[8], dphi[8][3];
        for (int a = 0; a < 8; ++a) {
            const int bx = (a == 1 || a == 2 || a == 5 || a == 6), by = (a == 2 || a == 3 || a == 6 || a == 7), bz = (a >= 4);
            const double fx = bx ? xi[0] : 1.0 - xi[0], fy = by ? xi[1] : 1.0 - xi[1], fz = bz ? xi[2] : 1.0 - xi[2];
            const double gx = bx ? 1.0 : -1.0, gy = by ? 1.0 : -1.0, gz = bz ? 1.0 : -1.0;
            phi[a] = fx * fy * fz;
            dphi[a][0] = gx * fy * fz;
            dphi[a][1] = fx * gy * fz;
            dphi[a][2] = fx * fy * gz;
        }
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, xq[3] = {0, 0, 0};
        for (int a = 0; a < 8; ++a)
            for (int r = 0; r < 3; ++r) {
                xq[r] += phi[a] * X[a][r];
                for (int c = 0; c < 3; ++c) J[r][c] += X[a][r] * dphi[a][c];
            }
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        const double id = 1.0 / det;
        double Ji[3][3];
        Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
        Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
        Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
        Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
        Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
        Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
        Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
        Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
        Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
        const double w = 0.125 * fabs(det);
        const double kq = kq_at(xq);
        double grad[8][3];
        for (int a = 0; a < 8; ++a)
            for (int r = 0; r < 3; ++r)
                grad[a][r] = Ji[0][r] * dphi[a][0] + Ji[1][r] * dphi[a][1] + Ji[2][r] * dphi[a][2];
        const double wk = w * kq;
        for (int a = 0; a < 8; ++a) {
            for (int b = 0; b < 8; ++b)
                Ke[a][b] += wk * (grad[a][0] * grad[b][0] + grad[a][1] * grad[b][1] + grad[a][2] * grad[b][2]);
            fe[a] += w * f_const * phi[a];
        }
    }
}

// FV1 on a P1 triangle: scvf e joins the midpoint of edge (e, e+1) to the
// barycentre; k grad(phi).n is subtracted from `from`, added to `to`;
// kq(ip) is the coefficient at the scvf's integration point (the midpoint of
// its two corners); area: the triangle's area (each scv holds a third)
template <typename KFn>
__device__ inline void fv1_tri_element_k(const double X[3][2], KFn kq_at, double Ke[3][3], double& area) {
    const double J00 = X[1][0] - X[0][0], J01 = X[2][0] - X[0][0];
    const double J10 = X[1][1] - X[0][1], J11 = X[2][1] - X[0][1];
    const double det = J00 * J11 - J01 * J10;
    area = 0.5 * fabs(det);
    const double gr[3][2] = {{-1, -1}, {1, 0}, {0, 1}};
    double grad[3][2];
    for (int i = 0; i < 3; ++i) {
        grad[i][0] = (J11 * gr[i][0] - J10 * gr[i][1]) / det;
        grad[i][1] = (-J01 * gr[i][0] + J00 * gr[i][1]) / det;
    }
    const double bx = (X[0][0] + X[1][0] + X[2][0]) / 3.0, by = (X[0][1] + X[1][1] + X[2][1]) / 3.0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Ke[a][b] = 0.0;
    for (int e = 0; e < 3; ++e) {
        const int fr = e, to = (e + 1) % 3;
        const double mx = 0.5 * (X[fr][0] + X[to][0]), my = 0.5 * (X[fr][1] + X[to][1]);
        double nx = by - my, ny = -(bx - mx);
        if (nx * (X[to][0] - X[fr][0]) + ny * (X[to][1] - X[fr][1]) < 0) { nx = -nx; ny = -ny; }
        const double ip[2] = {0.5 * (mx + bx), 0.5 * (my + by)};
        const double kq = kq_at(ip);
        for (int sh = 0; sh < 3; ++sh) {
            const double flux = kq * (grad[sh][0] * nx + grad[sh][1] * ny);
            Ke[fr][sh] -= flux;
            Ke[to][sh] += flux;
        }
    }
}

// ReferenceTetrahedron tables (reference_element.cpp:547-760)
__constant__ int8_t kTetEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
__constant__ int8_t kTetFace[4][3] = {{0, 2, 1}, {1, 2, 3}, {0, 3, 2}, {0, 1, 3}};
__constant__ int8_t kTetEdgeFace[6][2] = {{0, 3}, {0, 1}, {0, 2}, {3, 2}, {1, 3}, {2, 1}};
__constant__ int8_t kTetPointEdge[4][3] = {{2, 0, 3}, {0, 1, 4}, {1, 2, 5}, {3, 4, 5}};
__constant__ int8_t kTetPointFace[4][3] = {{0, 2, 3}, {0, 3, 1}, {0, 1, 2}, {3, 2, 1}};

__device__ __forceinline__ void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// volume of a hexahedron by its two-prism split (the scv volumes of FV1)
__device__ __forceinline__ double hex_volume(const double (*x)[3]) {
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

// FV1 element of a tetrahedron (fv1_geom.cpp:258-463): scvf i per edge with
// corners [edge mid, centre of face kTetEdgeFace[i][0], element centre, centre
// of face kTetEdgeFace[i][1]], normal 0.5 (c2-c0) x (c3-c1) from edge vertex 0
// to 1; diffusion kip[i] grad(phi).n subtracted from `from`, added to `to`;
// full upwind convection; scv volumes by the hexahedron formula (vol may be
// null)
__device__ inline void fv1_tet_element(const double (*X)[3], const double* kip, const double* vel, bool conv,
                                       double Ke[4][4], double vol[4]) {
    double M[6][3], F[4][3], C[3];
    for (int e = 0; e < 6; ++e)
        for (int d = 0; d < 3; ++d) M[e][d] = (X[kTetEdge[e][0]][d] + X[kTetEdge[e][1]][d]) * 0.5;
    for (int f = 0; f < 4; ++f)
        for (int d = 0; d < 3; ++d)
            F[f][d] = (X[kTetFace[f][0]][d] + X[kTetFace[f][1]][d] + X[kTetFace[f][2]][d]) / 3.0;
    for (int d = 0; d < 3; ++d) C[d] = (X[0][d] + X[1][d] + X[2][d] + X[3][d]) * 0.25;
    // P1 gradients: rows of J^{-1}, J = [X1-X0, X2-X0, X3-X0]
    double J[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) J[r][c] = X[c + 1][r] - X[0][r];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const double id = 1.0 / det;
    double g[4][3];
    g[1][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    g[1][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    g[1][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    g[2][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    g[2][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    g[2][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    g[3][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    g[3][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    g[3][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    for (int d = 0; d < 3; ++d) g[0][d] = -(g[1][d] + g[2][d] + g[3][d]);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Ke[i][j] = 0.0;
    for (int i = 0; i < 6; ++i) {
        const double* c0 = M[i];
        const double* c1 = F[kTetEdgeFace[i][0]];
        const double* c3 = F[kTetEdgeFace[i][1]];
        double a[3], b[3], nrm[3];
        for (int d = 0; d < 3; ++d) {
            a[d] = C[d] - c0[d];
            b[d] = c3[d] - c1[d];
        }
        cross3(a, b, nrm);
        for (int d = 0; d < 3; ++d) nrm[d] *= 0.5;
        const int from = kTetEdge[i][0], to = kTetEdge[i][1];
        for (int sh = 0; sh < 4; ++sh) {
            const double flux = kip[i] * (g[sh][0] * nrm[0] + g[sh][1] * nrm[1] + g[sh][2] * nrm[2]);
            Ke[from][sh] -= flux;
            Ke[to][sh] += flux;
        }
        if (conv) {
            const double q = vel[0] * nrm[0] + vel[1] * nrm[1] + vel[2] * nrm[2];
            const int up = q >= 0.0 ? from : to;
            Ke[from][up] += q;
            Ke[to][up] -= q;
        }
    }
    if (!vol) return;
    for (int i = 0; i < 4; ++i) {
        double H[8][3];
        const int8_t* pe = kTetPointEdge[i];
        const int8_t* pf = kTetPointFace[i];
        for (int d = 0; d < 3; ++d) {
            H[0][d] = X[i][d];
            H[1][d] = M[pe[1]][d];
            H[2][d] = F[pf[0]][d];
            H[3][d] = M[pe[0]][d];
            H[4][d] = M[pe[2]][d];
            H[5][d] = F[pf[2]][d];
            H[6][d] = C[d];
            H[7][d] = F[pf[1]][d];
        }
        vol[i] = hex_volume(H);
    }
}

}  // namespace
}  // namespace ugamd
