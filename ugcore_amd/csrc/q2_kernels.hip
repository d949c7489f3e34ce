// Q2 (triquadratic Lagrange) hexahedron element matrices on gfx950: one
// element per wavefront, the B^T D B contraction on the FP64 matrix cores.
//
// Per element (corners X in the reference hexahedron order, coefficient k,
// constant source f): Ke[a][b] = sum_q w_q |det J_q| k grad phi_a . grad phi_b
// and fe[a] = sum_q w_q |det J_q| f phi_a over the 3x3x3 Gauss-Legendre
// points, with the trilinear geometry of the 8 corners. Written as
// Ke = B^T C B with B (81 x 27) the physical gradients, one row per point and
// component, and C = diag(w_q |det J_q| k): the wave stages B (padded to
// 84 x 32) in LDS and forms the four 16 x 16 tiles of Ke with
// v_mfma_f64_16x16x4f64, 21 steps of 4 rows each.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ugx_internal.hpp"

namespace ugamd {
namespace {

// Local shape order of LagrangeLSFS<ReferenceHexahedron, 2> (lagrange.cpp:
// 236-253, lagrange.h:2106-2190): multi-indices of the vertices (2 x corner),
// the edge midpoints of edges (0,1) (1,2) (2,3) (3,0) (0,4) (1,5) (2,6) (3,7)
// (4,5) (5,6) (6,7) (7,4), the centres of faces {0,3,2,1} {0,1,5,4} {1,2,6,5}
// {2,3,7,6} {3,0,4,7} {4,5,6,7} (reference_element.cpp), then the interior.
__constant__ int8_t kQ2Mi[27][3] = {
    {0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2}, {0, 2, 2}, {1, 0, 0},
    {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1}, {1, 0, 2}, {2, 1, 2},
    {1, 2, 2}, {0, 1, 2}, {1, 1, 0}, {1, 0, 1}, {2, 1, 1}, {1, 2, 1}, {0, 1, 1}, {1, 1, 2}, {1, 1, 1}};

constexpr int kQ2Rows = 84;   // 27 points x 3 components, padded to a multiple of 4
constexpr int kQ2Ld = 33;     // LDS row stride of B (32 columns + 1)

// 1D Lagrange polynomials of the nodes 0, 1/2, 1 and their derivatives
__device__ __forceinline__ double q2_l(int i, double x) {
    return i == 0 ? (1.0 - x) * (1.0 - 2.0 * x) : i == 1 ? 4.0 * x * (1.0 - x) : x * (2.0 * x - 1.0);
}
__device__ __forceinline__ double q2_dl(int i, double x) {
    return i == 0 ? 4.0 * x - 3.0 : i == 1 ? 4.0 - 8.0 * x : 4.0 * x - 1.0;
}
__device__ __forceinline__ double q2_gauss(int i) {
    // 3-point Gauss-Legendre on [0, 1]: 1/2 -+ sqrt(3/5)/2
    return i == 0 ? 0.11270166537925831 : i == 1 ? 0.5 : 0.8872983346207417;
}
__device__ __forceinline__ double q2_gauss_w(int i) { return i == 1 ? 8.0 / 18.0 : 5.0 / 18.0; }

typedef double q2d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64) void k_q2_hex_elements(int64_t n, const double* __restrict__ X,
                                                        const double* __restrict__ kcoef, double f_const,
                                                        double* __restrict__ Ke, double* __restrict__ fe) {
    const int64_t e = blockIdx.x;
    if (e >= n) return;
    __shared__ double Bs[kQ2Rows * kQ2Ld];
    __shared__ double Xs[24];
    __shared__ double Jis[27][9];
    __shared__ double cs[27], fw[27];
    const int l = threadIdx.x;
    if (l < 24) Xs[l] = X[e * 24 + l];
    for (int t = l; t < kQ2Rows * kQ2Ld; t += 64) {
        const int r = t / kQ2Ld, c = t - r * kQ2Ld;
        if (r >= 81 || c >= 27) Bs[t] = 0.0;
    }
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
    // 2. B[3q + r][a] = (J_q^-T grad_xi phi_a(xi_q))_r for the 729 (q, a) pairs
    for (int t = l; t < 729; t += 64) {
        const int q = t / 27, a = t - q * 27;
        const double xi[3] = {q2_gauss(q % 3), q2_gauss((q / 3) % 3), q2_gauss(q / 9)};
        const int m0 = kQ2Mi[a][0], m1 = kQ2Mi[a][1], m2 = kQ2Mi[a][2];
        const double l0 = q2_l(m0, xi[0]), l1 = q2_l(m1, xi[1]), l2 = q2_l(m2, xi[2]);
        const double d0 = q2_dl(m0, xi[0]) * l1 * l2, d1 = l0 * q2_dl(m1, xi[1]) * l2, d2 = l0 * l1 * q2_dl(m2, xi[2]);
        const double* ji = Jis[q];
#pragma unroll
        for (int r = 0; r < 3; ++r) Bs[(3 * q + r) * kQ2Ld + a] = ji[r] * d0 + ji[3 + r] * d1 + ji[6 + r] * d2;
    }
    // load vector: fe[a] = sum_q w_q |det J_q| f phi_a(xi_q)
    if (fe && l < 27) {
        const int m0 = kQ2Mi[l][0], m1 = kQ2Mi[l][1], m2 = kQ2Mi[l][2];
        double s = 0.0;
        for (int q = 0; q < 27; ++q)
            s += fw[q] * q2_l(m0, q2_gauss(q % 3)) * q2_l(m1, q2_gauss((q / 3) % 3)) * q2_l(m2, q2_gauss(q / 9));
        fe[e * 27 + l] = s;
    }
    __syncthreads();
    // 3. Ke = (C B)^T B in 16 x 16 tiles: lane l supplies A[l % 16][l / 16] and
    //    B[l / 16][l % 16] of each 16x4 / 4x16 step; accumulator element v of
    //    lane l is row l / 16 + 4 v, column l % 16 of the tile (the f64 layout,
    //    measured on gfx950)
    const int li = l & 15, lk = l >> 4;
    double* out = Ke + e * 729;
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
            q2d4 acc = {0.0, 0.0, 0.0, 0.0};
            for (int kk = 0; kk < kQ2Rows / 4; ++kk) {
                const int k = 4 * kk + lk;
                const double c = k < 81 ? cs[k / 3] : 0.0;
                const double av = c * Bs[k * kQ2Ld + 16 * ti + li];
                const double bv = Bs[k * kQ2Ld + 16 * tj + li];
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = 16 * ti + lk + 4 * v, j = 16 * tj + li;
                if (i < 27 && j < 27) out[i * 27 + j] = acc[v];
            }
        }
}

}  // namespace

int q2_hex_elements(int64_t n, const double* X, const double* k, double f_const, double* Ke, double* fe,
                    hipStream_t s) {
    if (n <= 0) return 0;
    if (!X || !Ke || n > (int64_t)INT32_MAX) return -1;
    hipLaunchKernelGGL(k_q2_hex_elements, dim3((unsigned)n), dim3(64), 0, s, n, X, k, f_const, Ke, fe);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugamd
