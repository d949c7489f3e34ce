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

#include <algorithm>
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

constexpr int kQ2Waves = 8;   // elements (wavefronts) per workgroup

// the reference gradient and shape tables copied into the workgroup's LDS
// (their per-lane reads in the matrix steps then cost an LDS access instead
// of a cached global one); visible after the element routine's first barrier
__device__ __forceinline__ void q2_stage_table(Q2RefGrad& tab) {
    const double* src = reinterpret_cast<const double*>(&kQ2Grad);
    double* dst = reinterpret_cast<double*>(&tab);
    for (int i = threadIdx.x; i < (int)(sizeof(Q2RefGrad) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
}

// The element computation of one wavefront (lane l): corners from src.corner(l)
// for l < 24 (valid elements) and the coefficient src.k(q) at Gauss point q
// (x fastest), then sink.ke(i, j, v) for the 729
// entries and sink.fe(a, v) for the 27 load entries. Every wave of the
// workgroup calls it once (it holds two barriers); invalid waves work on the
// unit cube and emit nothing.
template <class Src, class Sink>
__device__ __forceinline__ void q2_element_wave(int l, bool valid, const Src& src, Sink& sink, double f_const,
                                                double* Xs, double (*Jis)[9], double* cs, double* fw, bool want_fe,
                                                const Q2RefGrad& tab) {
    if (l < 24) {
        const int c = l / 3, d = l - 3 * c;
        const int bit = d == 0 ? (c == 1 || c == 2 || c == 5 || c == 6) : d == 1 ? (c == 2 || c == 3 || c == 6 || c == 7)
                                                                         : (c >= 4);
        Xs[l] = valid ? src.corner(l) : (double)bit;
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
        cs[l] = wd * (valid ? src.k(l) : 1.0);
        fw[l] = wd * f_const;
    }
    __syncthreads();
    // load vector: fe[a] = sum_q w_q |det J_q| f phi_a(xi_q)
    if (want_fe && valid && l < 27) {
        double s = 0.0;
        for (int q = 0; q < 27; ++q) s += fw[q] * tab.phi[q][l];
        sink.fe(l, s);
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
            const double* g0 = tab.g[q][li];
            const double* g1 = tab.g[q][a1];
            c = cs[q];
            v0 = j0 * g0[0] + j1 * g0[1] + j2 * g0[2];
            if (has1) v1 = j0 * g1[0] + j1 * g1[1] + j2 * g1[2];
        }
        acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(c * v0, v0, acc00, 0, 0, 0);
        acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(c * v0, v1, acc01, 0, 0, 0);
        acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(c * v1, v1, acc11, 0, 0, 0);
    }
    if (!valid) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = lk + 4 * v;
        sink.ke(i, li, acc00[v]);
        if (has1) {
            sink.ke(i, 16 + li, acc01[v]);
            sink.ke(16 + li, i, acc01[v]);
        }
        if (16 + i < 27 && has1) sink.ke(16 + i, 16 + li, acc11[v]);
    }
}

struct Q2ArraySrc {
    const double* X;
    const double* kc;
    int64_t e;
    __device__ double corner(int l) const { return X[e * 24 + l]; }
    __device__ double k(int) const { return kc ? kc[e] : 1.0; }
};
struct Q2ArraySink {
    double* Kout;
    double* fout;
    __device__ void ke(int i, int j, double v) { Kout[i * 27 + j] = v; }
    __device__ void fe(int a, double v) { fout[a] = v; }
};

__global__ __launch_bounds__(64 * kQ2Waves) void k_q2_hex_elements(int64_t n, const double* __restrict__ X,
                                                                   const double* __restrict__ kcoef, double f_const,
                                                                   double* __restrict__ Ke, double* __restrict__ fe) {
    const int w = threadIdx.x >> 6;
    const int64_t e = (int64_t)blockIdx.x * kQ2Waves + w;
    const bool valid = e < n;   // the last workgroup may hold fewer elements; every wave reaches the barriers
    __shared__ double Xs_all[kQ2Waves][24];
    __shared__ double Jis_all[kQ2Waves][27][9];
    __shared__ double cs_all[kQ2Waves][28], fw_all[kQ2Waves][27];
    __shared__ Q2RefGrad tab;
    q2_stage_table(tab);
    Q2ArraySrc src{X, kcoef, e};
    Q2ArraySink sink{Ke + (valid ? e : 0) * 729, fe ? fe + (valid ? e : 0) * 27 : nullptr};
    q2_element_wave(threadIdx.x & 63, valid, src, sink, f_const, Xs_all[w], Jis_all[w], cs_all[w], fw_all[w],
                    fe != nullptr, tab);
}

// ---- assembly on the unit-cube lattice of n^3 elements ----
// Q2 nodes on the (2n+1)^3 lattice, numbered x fastest; element (ex, ey, ez)
// holds the nodes 2 e + m for its local multi-indices m. Two nodes are coupled
// when they share an element, which factorises per direction: an even
// coordinate c (a lattice plane of vertices) reaches c-2 .. c+2, an odd one
// (inside one element) c-1 .. c+1.
__device__ __forceinline__ void q2_range(int c, int W, int& lo, int& hi) {
    const int r = (c & 1) ? 1 : 2;
    lo = c - r < 0 ? 0 : c - r;
    hi = c + r > W - 1 ? W - 1 : c + r;
}

__global__ void k_q2_lattice_count(int W, int64_t N, int32_t* cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % W), y = (int)((i / W) % W), z = (int)(i / ((int64_t)W * W));
        int a, b, c, d, e, f;
        q2_range(x, W, a, b);
        q2_range(y, W, c, d);
        q2_range(z, W, e, f);
        cnt[i] = (b - a + 1) * (d - c + 1) * (f - e + 1);
    }
}

// one wavefront per row: lane l writes entries l, l + 64 of the row's
// (z, y, x)-ordered box of columns, so the row's writes are contiguous
__global__ __launch_bounds__(256) void k_q2_lattice_fill(int W, int64_t N, const int32_t* __restrict__ rowptr,
                                                         int32_t* __restrict__ col) {
    const int64_t i = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= N) return;
    const int lane = threadIdx.x & 63;
    const int x = (int)(i % W), y = (int)((i / W) % W), z = (int)(i / ((int64_t)W * W));
    int x0, x1, y0, y1, z0, z1;
    q2_range(x, W, x0, x1);
    q2_range(y, W, y0, y1);
    q2_range(z, W, z0, z1);
    const int nx = x1 - x0 + 1, ny = y1 - y0 + 1, len = nx * ny * (z1 - z0 + 1);
    const int32_t p = rowptr[i];
    for (int e = lane; e < len; e += 64) {
        const int xx = x0 + e % nx, yy = y0 + (e / nx) % ny, zz = z0 + e / (nx * ny);
        col[p + e] = (int32_t)(((int64_t)zz * W + yy) * W + xx);
    }
}

// the CSR slot of (r, c): the row's columns are its box [x0,x1] x [y0,y1] x
// [z0,z1] in (z, y, x) order, so the position follows from the offsets
__device__ __forceinline__ int32_t q2_slot(int W, const int32_t* __restrict__ rowptr, int32_t r, int32_t c) {
    const int x = r % W, y = (r / W) % W, z = r / (W * W);
    const int cx = c % W, cy = (c / W) % W, cz = c / (W * W);
    int x0, x1, y0, y1, z0, z1;
    q2_range(x, W, x0, x1);
    q2_range(y, W, y0, y1);
    q2_range(z, W, z0, z1);
    const int nx = x1 - x0 + 1, ny = y1 - y0 + 1;
    return rowptr[r] + ((cz - z0) * ny + (cy - y0)) * nx + (cx - x0);
}

struct Q2LatticeSrc {
    int n;
    double h;
    int ex, ey, ez;
    const double* theta;
    __device__ double corner(int l) const {
        const int c = l / 3, d = l - 3 * c;
        const int bit = d == 0 ? (c == 1 || c == 2 || c == 5 || c == 6) : d == 1 ? (c == 2 || c == 3 || c == 6 || c == 7)
                                                                         : (c >= 4);
        const int e = d == 0 ? ex : d == 1 ? ey : ez;
        return (double)(e + bit) * h;
    }
    __device__ double k(int) const { return theta ? exp(theta[((int64_t)ez * n + ey) * n + ex]) : 1.0; }
};
struct Q2CsrSink {
    int W, ex, ey, ez;
    const int32_t* rowptr;
    const int32_t* col;
    double* val;
    double* rhs;
    __device__ int32_t node(int a) const {
        return (int32_t)(((int64_t)(2 * ez + kQ2Mi[a][2]) * W + (2 * ey + kQ2Mi[a][1])) * W + (2 * ex + kQ2Mi[a][0]));
    }
    __device__ void ke(int i, int j, double v) { val[q2_slot(W, rowptr, node(i), node(j))] += v; }
    __device__ void fe(int a, double v) { rhs[node(a)] += v; }
};

// one colour (element parities px, py, pz): no two of its elements share a node
__global__ __launch_bounds__(64 * kQ2Waves) void k_q2_lattice_assemble(int n, int color, const double* theta,
                                                                       double f_const, const int32_t* rowptr,
                                                                       const int32_t* col, double* val, double* rhs) {
    const int hn = (n + 1) / 2;
    const int w = threadIdx.x >> 6;
    const int64_t t = (int64_t)blockIdx.x * kQ2Waves + w;
    const int px = color & 1, py = (color >> 1) & 1, pz = color >> 2;
    const int tx = (int)(t % hn), ty = (int)((t / hn) % hn), tz = (int)(t / ((int64_t)hn * hn));
    const int ex = 2 * tx + px, ey = 2 * ty + py, ez = 2 * tz + pz;
    const bool valid = t < (int64_t)hn * hn * hn && ex < n && ey < n && ez < n;
    __shared__ double Xs_all[kQ2Waves][24];
    __shared__ double Jis_all[kQ2Waves][27][9];
    __shared__ double cs_all[kQ2Waves][28], fw_all[kQ2Waves][27];
    __shared__ Q2RefGrad tab;
    q2_stage_table(tab);
    Q2LatticeSrc src{n, 1.0 / n, ex, ey, ez, theta};
    Q2CsrSink sink{2 * n + 1, ex, ey, ez, rowptr, col, val, rhs};
    q2_element_wave(threadIdx.x & 63, valid, src, sink, f_const, Xs_all[w], Jis_all[w], cs_all[w], fw_all[w],
                    rhs != nullptr, tab);
}

// Dirichlet rows (u = 0 on the boundary of the cube): row := unit row on its
// pattern, rhs := 0
__global__ void k_q2_lattice_dirichlet(int W, int64_t N, const int32_t* __restrict__ rowptr,
                                       const int32_t* __restrict__ col, double* val, double* rhs) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % W), y = (int)((i / W) % W), z = (int)(i / ((int64_t)W * W));
        if (x > 0 && y > 0 && z > 0 && x < W - 1 && y < W - 1 && z < W - 1) continue;
        for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) val[p] = col[p] == (int32_t)i ? 1.0 : 0.0;
        rhs[i] = 0.0;
    }
}

// ---- Q2 hexahedra of a caller-supplied connectivity (ElementDiscretization) ----
struct Q2ElemSrc {
    const double* coords;   // DoF coordinates [N][3]
    const int32_t* dofs;    // [ne][27] in the local shape order
    const double* kel;      // per element or null
    const double* kq;       // per element and Gauss point [ne][27] or null
    int64_t e;
    __device__ double corner(int l) const {
        const int c = l / 3, d = l - 3 * c;
        return coords[(int64_t)dofs[e * 27 + c] * 3 + d];
    }
    __device__ double k(int q) const { return kq ? kq[e * 27 + q] : (kel ? kel[e] : 1.0); }
};
// matrix entries through the per-entry slots (slot[(27 i + j) * ne + pos] for
// the element at colour-sorted position pos), load entries into rhs
struct Q2SlotSink {
    const int32_t* slot;
    const int32_t* dofs;
    int64_t ne, pos, e;
    double* val;
    double* rhs;
    __device__ void ke(int i, int j, double v) {
        if (val) val[slot[(int64_t)(27 * i + j) * ne + pos]] += v;
    }
    __device__ void fe(int a, double v) { rhs[dofs[e * 27 + a]] += v; }
};

// one colour's elements (positions pos0 .. pos0+cnt-1 of the colour order),
// one element per wavefront
__global__ __launch_bounds__(64 * kQ2Waves) void k_elem_hex_q2(int64_t ne, const int32_t* __restrict__ order,
                                                               const int32_t* __restrict__ dofs,
                                                               const double* __restrict__ coords,
                                                               const double* __restrict__ kel,
                                                               const double* __restrict__ kq, double f_const,
                                                               const int32_t* __restrict__ slot, double* val,
                                                               double* rhs, int64_t pos0, int64_t cnt) {
    const int w = threadIdx.x >> 6;
    const int64_t t = (int64_t)blockIdx.x * kQ2Waves + w;
    const bool valid = t < cnt;   // every wave reaches the barriers of the element routine
    const int64_t pos = pos0 + (valid ? t : 0);
    const int64_t e = order[pos];
    __shared__ double Xs_all[kQ2Waves][24];
    __shared__ double Jis_all[kQ2Waves][27][9];
    __shared__ double cs_all[kQ2Waves][28], fw_all[kQ2Waves][27];
    __shared__ Q2RefGrad tab;
    q2_stage_table(tab);
    Q2ElemSrc src{coords, dofs, kel, kq, e};
    Q2SlotSink sink{slot, dofs, ne, pos, e, val, rhs};
    q2_element_wave(threadIdx.x & 63, valid, src, sink, f_const, Xs_all[w], Jis_all[w], cs_all[w], fw_all[w],
                    rhs != nullptr, tab);
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

int q2_lattice_count(int n, int32_t* cnt, hipStream_t s) {
    const int W = 2 * n + 1;
    const int64_t N = (int64_t)W * W * W;
    hipLaunchKernelGGL(k_q2_lattice_count, dim3((unsigned)std::min<int64_t>((N + 255) / 256, 65536)), dim3(256), 0, s,
                       W, N, cnt);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int q2_lattice_fill(int n, const int32_t* rowptr, int32_t* col, hipStream_t s) {
    const int W = 2 * n + 1;
    const int64_t N = (int64_t)W * W * W;
    hipLaunchKernelGGL(k_q2_lattice_fill, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, s, W, N, rowptr, col);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int q2_lattice_assemble(int n, const double* theta, double f_const, const int32_t* rowptr, const int32_t* col,
                        double* val, double* rhs, hipStream_t s) {
    const int hn = (n + 1) / 2;
    const int64_t per = (int64_t)hn * hn * hn;
    for (int color = 0; color < 8; ++color)
        hipLaunchKernelGGL(k_q2_lattice_assemble, dim3((unsigned)((per + kQ2Waves - 1) / kQ2Waves)),
                           dim3(64 * kQ2Waves), 0, s, n, color, theta, f_const, rowptr, col, val, rhs);
    const int W = 2 * n + 1;
    const int64_t N = (int64_t)W * W * W;
    hipLaunchKernelGGL(k_q2_lattice_dirichlet, dim3((unsigned)std::min<int64_t>((N + 255) / 256, 65536)), dim3(256),
                       0, s, W, N, rowptr, col, val, rhs);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int elem_assemble_q2_colour(int64_t ne, const int32_t* order, const int32_t* dofs, const double* coords,
                            const double* kel, const double* kq, double f_const, const int32_t* slot, double* val,
                            double* rhs, int64_t pos0, int64_t cnt, hipStream_t s) {
    if (cnt <= 0) return 0;
    hipLaunchKernelGGL(k_elem_hex_q2, dim3((unsigned)((cnt + kQ2Waves - 1) / kQ2Waves)), dim3(64 * kQ2Waves), 0, s,
                       ne, order, dofs, coords, kel, kq, f_const, slot, val, rhs, pos0, cnt);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugamd
