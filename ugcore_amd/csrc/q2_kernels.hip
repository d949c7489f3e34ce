// Q2 (triquadratic Lagrange) hexahedron element matrices on gfx950: one
// element per wavefront, the B^T D B contraction on the FP64 matrix cores.
//
// Per element (corners X in the reference hexahedron order, coefficient k,
// constant source f): Ke[a][b] = sum_q w_q |det J_q| k grad phi_a . grad phi_b
// and fe[a] = sum_q w_q |det J_q| f phi_a over the 3x3x3 Gauss-Legendre
// points, with the trilinear geometry of the 8 corners. Written as
// Ke = B^T C B with B (81 x 27) the physical gradients, one row per point and
// component, and C = diag(w_q |det J_q| k): each lane forms the two entries
// of B (padded to 84 x 32) it feeds to each v_mfma_f64_16x16x4f64 step in
// registers, and the wave accumulates three 16 x 16 tiles of the symmetric Ke
// over 21 steps of 4 rows (only the Jacobians sit in LDS).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ugx_internal.hpp"

namespace ugamd {
namespace {

constexpr int kQ2Rows = 84;   // 27 points x 3 components, padded to a multiple of 4

__device__ __forceinline__ double q2_gauss(int i) {
    // 3-point Gauss-Legendre on [0, 1]: 1/2 -+ sqrt(3/5)/2
    return i == 0 ? 0.11270166537925831 : i == 1 ? 0.5 : 0.8872983346207417;
}
__device__ __forceinline__ double q2_gauss_w(int i) { return i == 1 ? 8.0 / 18.0 : 5.0 / 18.0; }

typedef double q2d4 __attribute__((ext_vector_type(4)));

// grad_xi phi_a and phi_a at Gauss point q (q = qx + 3 qy + 9 qz), evaluated
// at compile time from the 1D Lagrange polynomials of the nodes 0, 1/2, 1 and
// the local order
struct Q2RefGrad {
    double g[27][27][3];
    double phi[27][27];   // phi_a at point q
};
// Local shape order of LagrangeLSFS<ReferenceHexahedron, 2> (lagrange.cpp:
// 236-253, lagrange.h:2106-2190): multi-indices of the vertices (2 x corner),
// the edge midpoints of edges (0,1) (1,2) (2,3) (3,0) (0,4) (1,5) (2,6) (3,7)
// (4,5) (5,6) (6,7) (7,4), the centres of faces {0,3,2,1} {0,1,5,4} {1,2,6,5}
// {2,3,7,6} {3,0,4,7} {4,5,6,7} (reference_element.cpp), then the interior.
constexpr int8_t kQ2Mi[27][3] = {
    {0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2}, {0, 2, 2}, {1, 0, 0},
    {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1}, {1, 0, 2}, {2, 1, 2},
    {1, 2, 2}, {0, 1, 2}, {1, 1, 0}, {1, 0, 1}, {2, 1, 1}, {1, 2, 1}, {0, 1, 1}, {1, 1, 2}, {1, 1, 1}};
constexpr double q2c_x(int i) { return i == 0 ? 0.11270166537925831 : i == 1 ? 0.5 : 0.8872983346207417; }
constexpr double q2c_l(int i, double x) {
    return i == 0 ? (1.0 - x) * (1.0 - 2.0 * x) : i == 1 ? 4.0 * x * (1.0 - x) : x * (2.0 * x - 1.0);
}
constexpr double q2c_dl(int i, double x) { return i == 0 ? 4.0 * x - 3.0 : i == 1 ? 4.0 - 8.0 * x : 4.0 * x - 1.0; }
constexpr Q2RefGrad q2_ref_grad() {
    Q2RefGrad t{};
    for (int q = 0; q < 27; ++q)
        for (int a = 0; a < 27; ++a) {
            const double x0 = q2c_x(q % 3), x1 = q2c_x((q / 3) % 3), x2 = q2c_x(q / 9);
            const int m0 = kQ2Mi[a][0], m1 = kQ2Mi[a][1], m2 = kQ2Mi[a][2];
            const double l0 = q2c_l(m0, x0), l1 = q2c_l(m1, x1), l2 = q2c_l(m2, x2);
            t.g[q][a][0] = q2c_dl(m0, x0) * l1 * l2;
            t.g[q][a][1] = l0 * q2c_dl(m1, x1) * l2;
            t.g[q][a][2] = l0 * l1 * q2c_dl(m2, x2);
            t.phi[q][a] = l0 * l1 * l2;
        }
    return t;
}
__constant__ Q2RefGrad kQ2Grad = q2_ref_grad();

constexpr int kQ2Waves = 4;   // elements (wavefronts) per workgroup

__global__ __launch_bounds__(64 * kQ2Waves) void k_q2_hex_elements(int64_t n, const double* __restrict__ X,
                                                                   const double* __restrict__ kcoef, double f_const,
                                                                   double* __restrict__ Ke, double* __restrict__ fe) {
    const int w = threadIdx.x >> 6;
    const int64_t e = (int64_t)blockIdx.x * kQ2Waves + w;
    const bool valid = e < n;   // the last workgroup may hold fewer elements; every wave reaches the barriers
    __shared__ double Xs_all[kQ2Waves][24];
    __shared__ double Jis_all[kQ2Waves][27][9];
    __shared__ double cs_all[kQ2Waves][28], fw_all[kQ2Waves][27];
    double* Xs = Xs_all[w];
    double(*Jis)[9] = Jis_all[w];
    double* cs = cs_all[w];
    double* fw = fw_all[w];
    const int l = threadIdx.x & 63;
    if (l < 24) {
        // a wave without an element works on the unit cube and writes nothing
        const int c = l / 3, d = l - 3 * c;
        const int bit = d == 0 ? (c == 1 || c == 2 || c == 5 || c == 6) : d == 1 ? (c == 2 || c == 3 || c == 6 || c == 7)
                                                                         : (c >= 4);
        Xs[l] = valid ? X[e * 24 + l] : (double)bit;
    }
    if (l == 27) cs[27] = 0.0;   // the padding rows 81..83 of B
    __syncthreads();
    // 1. Jacobian of the trilinear map at point q = lane (x fastest)
    if (l < 27) {
        const int qx = l % 3, qy = (l / 3) % 3, qz = l / 9;
        const double xi[3] = {q2_gauss(qx), q2_gauss(qy), q2_gauss(qz)};
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int a = 0; a < 8; ++a) {
            const int cx = (a == 1 || a == 2 || a == 5 || a == 6), cy = (a == 2 || a == 3 || a == 6 || a == 7),
                      cz = a >= 4;
            const double fx = cx ? xi[0] : 1.0 - xi[0], fy = cy ? xi[1] : 1.0 - xi[1], fz = cz ? xi[2] : 1.0 - xi[2];
            const double d[3] = {(cx ? 1.0 : -1.0) * fy * fz, fx * (cy ? 1.0 : -1.0) * fz, fx * fy * (cz ? 1.0 : -1.0)};
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) J[r][c] += Xs[3 * a + r] * d[c];
        }
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        const double id = 1.0 / det;
        double* ji = Jis[l];
        ji[0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
        ji[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
        ji[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
        ji[3] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
        ji[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
        ji[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
        ji[6] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
        ji[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
        ji[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
        const double wd = q2_gauss_w(qx) * q2_gauss_w(qy) * q2_gauss_w(qz) * fabs(det);
        cs[l] = wd * (kcoef ? kcoef[e] : 1.0);
        fw[l] = wd * f_const;
    }
    __syncthreads();
    // load vector: fe[a] = sum_q w_q |det J_q| f phi_a(xi_q)
    if (fe && valid && l < 27) {
        double s = 0.0;
        for (int q = 0; q < 27; ++q) s += fw[q] * kQ2Grad.phi[q][l];
        fe[e * 27 + l] = s;
    }
    // 2. Ke = (C B)^T B in the 16 x 16 tiles (0,0), (0,1), (1,1); (1,0) is the
    //    transpose of (0,1). Step kk covers the rows k = 4 kk .. 4 kk + 3 of B
    //    (point k / 3, component k % 3; rows 81..83 are zero): lane
    //    l = 16 lk + li forms B[4 kk + lk][a] = (J_q^-T grad_xi phi_a(xi_q))_(k % 3)
    //    for the shapes a = li and 16 + li (zero beyond 26) and supplies them
    //    as A[li][lk] (scaled by C) and B[lk][li] of the 16x4 / 4x16 step.
    //    Accumulator element v of lane l is row l / 16 + 4 v, column l % 16 of
    //    the tile (the f64 layout, measured on gfx950).
    const int li = l & 15, lk = l >> 4;
    const int a1 = 16 + li < 27 ? 16 + li : 26;
    const bool has1 = 16 + li < 27;
    q2d4 acc00 = {0.0, 0.0, 0.0, 0.0}, acc01 = acc00, acc11 = acc00;
    for (int kk = 0; kk < kQ2Rows / 4; ++kk) {
        const int k = 4 * kk + lk;
        const int q = k / 3, r = k - 3 * q;
        double v0 = 0.0, v1 = 0.0, c = 0.0;
        if (k < 81) {
            const double j0 = Jis[q][r], j1 = Jis[q][3 + r], j2 = Jis[q][6 + r];
            const double* g0 = kQ2Grad.g[q][li];
            const double* g1 = kQ2Grad.g[q][a1];
            c = cs[q];
            v0 = j0 * g0[0] + j1 * g0[1] + j2 * g0[2];
            if (has1) v1 = j0 * g1[0] + j1 * g1[1] + j2 * g1[2];
        }
        acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(c * v0, v0, acc00, 0, 0, 0);
        acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(c * v0, v1, acc01, 0, 0, 0);
        acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(c * v1, v1, acc11, 0, 0, 0);
    }
    if (!valid) return;
    double* out = Ke + e * 729;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = lk + 4 * v;
        out[i * 27 + li] = acc00[v];
        if (has1) {
            out[i * 27 + 16 + li] = acc01[v];
            out[(16 + li) * 27 + i] = acc01[v];
        }
        if (16 + i < 27 && has1) out[(16 + i) * 27 + 16 + li] = acc11[v];
    }
}

}  // namespace

int q2_hex_elements(int64_t n, const double* X, const double* k, double f_const, double* Ke, double* fe,
                    hipStream_t s) {
    if (n <= 0) return 0;
    if (!X || !Ke || n > (int64_t)INT32_MAX) return -1;
    hipLaunchKernelGGL(k_q2_hex_elements, dim3((unsigned)((n + kQ2Waves - 1) / kQ2Waves)), dim3(64 * kQ2Waves), 0, s,
                       n, X, k, f_const, Ke, fe);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugamd
