// Q2 Lagrange on the hexahedral lattice hierarchy (unit cube from one
// hexahedron): LEVEL DoF numbering, element connectivity, coefficients at the
// Gauss points, and the grid transfers of StdTransfer.
//
// Numbering (DoFDistribution::reinit, dof_distribution.cpp:1401-1444): Q2
// puts one DoF on every vertex, edge, face and hexahedron of a level, numbered
// by reinit<Vertex>, <Edge>, <Face>, <Volume> in turn, each subset ("Inner",
// then "Boundary") in list order. The objects of level l are in one-to-one
// correspondence with the vertices of level l+1 (copies of the vertices, then
// one midpoint per edge, centre per face, centre per hexahedron, created in
// the order of their parents; global_multi_grid_refiner.cpp:242-445), and a
// level-(l+1) vertex inherits its parent's subset. The level-(l+1) P1
// numbering lists the vertices of each subset in that creation order, so the
// Q2 order of level l is the stable sort of the level-(l+1) vertices by
// (type of parent object, subset) with the P1 number as the tie-break; the
// type is the number of odd coordinates of the vertex on the level-(l+1)
// lattice (= the doubled lattice of level l, where the Q2 nodes sit).
//
// Transfers (StdTransfer::assemble_prolongation, LAGRANGE branch,
// std_transfer_impl.h:214-388): a fine DoF takes the DoFs of its parent's
// closure with the parent's Q2 shape values at the DoF position. Along an
// axis a fine node at a multiple of 4 (fine doubled units) lies on a coarse
// vertex plane (one coarse node, weight 1); otherwise it lies inside coarse
// element e = i/4 at t = (i-4e)/4 in {1/4, 1/2, 3/4}, which the 3 coarse nodes
// 2e, 2e+1, 2e+2 carry with the 1D quadratic Lagrange values (3/8, 3/4, -1/8),
// (0, 1, 0) or (-1/8, 3/4, 3/8). The parent is the coarse object spanned by
// the axes of the second kind, its closure the product of the node sets.
// adjust_prolongation (lagrange_dirichlet_boundary_impl.h:485-614): a fine
// Dirichlet row keeps its pattern with the value 1 at the parent's own DoF
// (the middle node along every spanned axis) and 0 elsewhere;
// adjust_restriction (:617-740) on R = P^T: the row of a coarse Dirichlet DoF
// keeps only the entry of its last Dirichlet child in the order the reference
// visits them (vertices, edges, faces; list order), i.e. the child with the
// largest fine DoF number.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <climits>
#include <cmath>
#include <cstdint>

#include "ugx_internal.hpp"

namespace ugamd {

namespace {

inline unsigned nblk(int64_t n, int t) {
    const int64_t b = (n + t - 1) / t;
    return (unsigned)(b < 1 ? 1 : (b > 1048576 ? 1048576 : b));
}

// sort keys of the level-(l+1) vertices: (2 * parent type + subset) above the
// P1 DoF number; values: the vertex's lattice index
__global__ void k_q2_node_keys(int64_t N1, int64_t n_inner1, int W1, const int32_t* __restrict__ lat_of_dof1,
                               uint64_t* __restrict__ key, int32_t* __restrict__ node) {
    for (int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; d < N1; d += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = lat_of_dof1[d];
        const uint32_t vu = (uint32_t)v, Wu = (uint32_t)W1;
        const int x = (int)(vu % Wu), y = (int)((vu / Wu) % Wu), z = (int)(vu / (Wu * Wu));
        const uint64_t type = (uint64_t)((x & 1) + (y & 1) + (z & 1));
        const uint64_t sub = d >= n_inner1 ? 1 : 0;
        key[d] = ((2 * type + sub) << 32) | (uint64_t)d;
        node[d] = v;
    }
}

__global__ void k_q2_invert(int64_t N, int W2, const int32_t* __restrict__ node_of_dof, int32_t* __restrict__ dof_of_node,
                            uint8_t* __restrict__ dir) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < N; k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = node_of_dof[k];
        dof_of_node[v] = (int32_t)k;
        const uint32_t vu = (uint32_t)v, Wu = (uint32_t)W2;
        const int x = (int)(vu % Wu), y = (int)((vu / Wu) % Wu), z = (int)(vu / (Wu * Wu));
        dir[k] = (x == 0 || y == 0 || z == 0 || x == W2 - 1 || y == W2 - 1 || z == W2 - 1) ? 1 : 0;
    }
}

// local multi-indices of the Q2 shapes (q2_kernels.hip, lagrange.cpp:236-253)
__constant__ int8_t kMi[27][3] = {
    {0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2}, {0, 2, 2}, {1, 0, 0},
    {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1}, {1, 0, 2}, {2, 1, 2},
    {1, 2, 2}, {0, 1, 2}, {1, 1, 0}, {1, 0, 1}, {2, 1, 1}, {1, 2, 1}, {0, 1, 1}, {1, 1, 2}, {1, 1, 1}};

__global__ void k_q2_connectivity(int n, const int32_t* __restrict__ dof_of_node, int32_t* __restrict__ conn) {
    const int64_t ne = (int64_t)n * n * n;
    const int W2 = 2 * n + 1;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ne * 27;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = i / 27;
        const int a = (int)(i - 27 * e);
        const int ex = (int)(e % n), ey = (int)((e / n) % n), ez = (int)(e / ((int64_t)n * n));
        const int64_t v = ((int64_t)(2 * ez + kMi[a][2]) * W2 + (2 * ey + kMi[a][1])) * W2 + (2 * ex + kMi[a][0]);
        conn[i] = dof_of_node[v];
    }
}

__global__ void k_q2_coords(int64_t N, int n, const int32_t* __restrict__ node_of_dof, double* __restrict__ X) {
    const int W2 = 2 * n + 1;
    const double s = 1.0 / (2.0 * n);
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < N; k += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t vu = (uint32_t)node_of_dof[k], Wu = (uint32_t)W2;
        X[3 * k] = (double)(vu % Wu) * s;
        X[3 * k + 1] = (double)((vu / Wu) % Wu) * s;
        X[3 * k + 2] = (double)(vu / (Wu * Wu)) * s;
    }
}

// kq[e][q] = exp(theta) of the fine element holding Gauss point q (x fastest)
// of element e; off[i] = floor(xi_i * sub): the fine-cell offset inside an
// element of the Gauss point i along an axis. The middle point (xi = 1/2) of
// an element spanning an even number of fine elements lies on a fine-element
// boundary and takes the element on its upper side.
struct GaussCells {
    int off[3];
};
__global__ void k_q2_coef_points(int n, int sub, GaussCells gc, const double* __restrict__ theta,
                                 double* __restrict__ kq) {
    const int64_t ne = (int64_t)n * n * n;
    const int64_t nf = (int64_t)n * sub;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ne * 27;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = i / 27;
        const int q = (int)(i - 27 * e);
        const int ex = (int)(e % n), ey = (int)((e / n) % n), ez = (int)(e / ((int64_t)n * n));
        const int64_t cx = (int64_t)ex * sub + gc.off[q % 3], cy = (int64_t)ey * sub + gc.off[(q / 3) % 3],
                      cz = (int64_t)ez * sub + gc.off[q / 9];
        kq[i] = theta ? exp(theta[(cz * nf + cy) * nf + cx]) : 1.0;
    }
}

// along one axis: the coarse nodes (coarse doubled units) and 1D weights of
// fine coordinate i (fine doubled units); returns their count (1 or 3) and
// the parent's own node
__device__ __forceinline__ int axis_nodes(int i, int* c, double* w, int& own) {
    if ((i & 3) == 0) {
        c[0] = i >> 1;
        w[0] = 1.0;
        own = c[0];
        return 1;
    }
    const int e = i >> 2, r = i & 3;
    c[0] = 2 * e;
    c[1] = 2 * e + 1;
    c[2] = 2 * e + 2;
    if (r == 1) {
        w[0] = 0.375;
        w[1] = 0.75;
        w[2] = -0.125;
    } else if (r == 2) {
        w[0] = 0.0;
        w[1] = 1.0;
        w[2] = 0.0;
    } else {
        w[0] = -0.125;
        w[1] = 0.75;
        w[2] = 0.375;
    }
    own = 2 * e + 1;
    return 3;
}

__global__ void k_q2_prolong_count(int64_t Nf, int W2f, const int32_t* __restrict__ node_of_dof_f,
                                   int32_t* __restrict__ cnt) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < Nf; r += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t vu = (uint32_t)node_of_dof_f[r], Wu = (uint32_t)W2f;
        const int x = (int)(vu % Wu), y = (int)((vu / Wu) % Wu), z = (int)(vu / (Wu * Wu));
        cnt[r] = ((x & 3) ? 3 : 1) * ((y & 3) ? 3 : 1) * ((z & 3) ? 3 : 1);
    }
}

__global__ void k_q2_prolong_fill(int64_t Nf, int W2f, int W2c, const int32_t* __restrict__ node_of_dof_f,
                                  const uint8_t* __restrict__ dir_f, const int32_t* __restrict__ dof_of_node_c,
                                  const int32_t* __restrict__ rowptr, int32_t* __restrict__ col,
                                  double* __restrict__ val) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < Nf; r += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t vu = (uint32_t)node_of_dof_f[r], Wu = (uint32_t)W2f;
        const int x = (int)(vu % Wu), y = (int)((vu / Wu) % Wu), z = (int)(vu / (Wu * Wu));
        int cx[3], cy[3], cz[3], ox, oy, oz;
        double wx[3], wy[3], wz[3];
        const int nx = axis_nodes(x, cx, wx, ox), ny = axis_nodes(y, cy, wy, oy), nz = axis_nodes(z, cz, wz, oz);
        const bool dirichlet = dir_f[r] != 0;
        const int32_t own = dof_of_node_c[((int64_t)oz * W2c + oy) * W2c + ox];
        int32_t cols[27];
        double vals[27];
        int m = 0;
        for (int c = 0; c < nz; ++c)
            for (int b = 0; b < ny; ++b)
                for (int a = 0; a < nx; ++a) {
                    const int32_t d = dof_of_node_c[((int64_t)cz[c] * W2c + cy[b]) * W2c + cx[a]];
                    cols[m] = d;
                    vals[m] = dirichlet ? (d == own ? 1.0 : 0.0) : wz[c] * wy[b] * wx[a];
                    ++m;
                }
        // ascending columns (insertion sort of at most 27)
        for (int i = 1; i < m; ++i) {
            const int32_t ck = cols[i];
            const double vk = vals[i];
            int j = i - 1;
            while (j >= 0 && cols[j] > ck) {
                cols[j + 1] = cols[j];
                vals[j + 1] = vals[j];
                --j;
            }
            cols[j + 1] = ck;
            vals[j + 1] = vk;
        }
        const int32_t base = rowptr[r];
        for (int i = 0; i < m; ++i) {
            col[base + i] = cols[i];
            val[base + i] = vals[i];
        }
    }
}

// last[p] = the largest fine Dirichlet DoF whose parent's own DoF is p
__global__ void k_q2_last_child(int64_t Nf, int W2f, int W2c, const int32_t* __restrict__ node_of_dof_f,
                                const uint8_t* __restrict__ dir_f, const int32_t* __restrict__ dof_of_node_c,
                                int32_t* __restrict__ last) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < Nf; r += (int64_t)gridDim.x * blockDim.x) {
        if (!dir_f[r]) continue;
        const uint32_t vu = (uint32_t)node_of_dof_f[r], Wu = (uint32_t)W2f;
        const int x = (int)(vu % Wu), y = (int)((vu / Wu) % Wu), z = (int)(vu / (Wu * Wu));
        const int ox = (x & 3) ? 2 * (x >> 2) + 1 : x >> 1, oy = (y & 3) ? 2 * (y >> 2) + 1 : y >> 1,
                  oz = (z & 3) ? 2 * (z >> 2) + 1 : z >> 1;
        atomicMax(&last[dof_of_node_c[((int64_t)oz * W2c + oy) * W2c + ox]], (int32_t)r);
    }
}

__global__ void k_q2_restrict_rows(int64_t Nc, const int32_t* __restrict__ last, const int32_t* __restrict__ rowptr,
                                   const int32_t* __restrict__ col, double* __restrict__ val) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < Nc; p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t l = last[p];
        if (l < 0) continue;
        for (int32_t k = rowptr[p]; k < rowptr[p + 1]; ++k) val[k] = col[k] == l ? 1.0 : 0.0;
    }
}

}  // namespace

int q2_numbering_temp_bytes(int64_t N1, size_t* bytes) {
    *bytes = 0;
    if (N1 <= 0) return 0;
    return rocprim::radix_sort_pairs(nullptr, *bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                     (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)N1, 0, 64) == hipSuccess
               ? 0
               : -1;
}

int q2_numbering(int64_t N1, int64_t n_inner1, int W1, const int32_t* lat_of_dof1, uint64_t* key, uint64_t* key2,
                 int32_t* node, int32_t* node_of_dof, void* temp, size_t temp_bytes, int32_t* dof_of_node,
                 uint8_t* dir, hipStream_t s) {
    if (N1 <= 0) return 0;
    hipLaunchKernelGGL(k_q2_node_keys, dim3(nblk(N1, 256)), dim3(256), 0, s, N1, n_inner1, W1, lat_of_dof1, key, node);
    size_t b = temp_bytes;
    if (rocprim::radix_sort_pairs(temp, b, key, key2, node, node_of_dof, (size_t)N1, 0, 64, s) != hipSuccess)
        return -1;
    hipLaunchKernelGGL(k_q2_invert, dim3(nblk(N1, 256)), dim3(256), 0, s, N1, W1, node_of_dof, dof_of_node, dir);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int q2_connectivity(int n, const int32_t* dof_of_node, int32_t* conn, hipStream_t s) {
    const int64_t m = (int64_t)n * n * n * 27;
    if (m <= 0) return 0;
    hipLaunchKernelGGL(k_q2_connectivity, dim3(nblk(m, 256)), dim3(256), 0, s, n, dof_of_node, conn);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int q2_coords(int64_t N, int n, const int32_t* node_of_dof, double* X, hipStream_t s) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(k_q2_coords, dim3(nblk(N, 256)), dim3(256), 0, s, N, n, node_of_dof, X);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int q2_coef_points(int n, int nfine, const double* theta, double* kq, hipStream_t s) {
    if (n <= 0 || nfine % n != 0) return -1;
    const int sub = nfine / n;
    GaussCells gc;
    const double r = 0.5 * std::sqrt(0.6);
    const double xi[3] = {0.5 - r, 0.5, 0.5 + r};
    for (int i = 0; i < 3; ++i) gc.off[i] = (int)std::floor(xi[i] * sub);
    const int64_t m = (int64_t)n * n * n * 27;
    hipLaunchKernelGGL(k_q2_coef_points, dim3(nblk(m, 256)), dim3(256), 0, s, n, sub, gc, theta, kq);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int q2_prolong_count(int64_t Nf, int nf, const int32_t* node_of_dof_f, int32_t* cnt, hipStream_t s) {
    if (Nf <= 0) return 0;
    hipLaunchKernelGGL(k_q2_prolong_count, dim3(nblk(Nf, 256)), dim3(256), 0, s, Nf, 2 * nf + 1, node_of_dof_f, cnt);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int q2_prolong_fill(int64_t Nf, int nf, const int32_t* node_of_dof_f, const uint8_t* dir_f,
                    const int32_t* dof_of_node_c, const int32_t* rowptr, int32_t* col, double* val, hipStream_t s) {
    if (Nf <= 0) return 0;
    hipLaunchKernelGGL(k_q2_prolong_fill, dim3(nblk(Nf, 256)), dim3(256), 0, s, Nf, 2 * nf + 1, nf + 1,
                       node_of_dof_f, dir_f, dof_of_node_c, rowptr, col, val);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int q2_restriction_adjust(int64_t Nf, int nf, int64_t Nc, const int32_t* node_of_dof_f, const uint8_t* dir_f,
                          const int32_t* dof_of_node_c, int32_t* last, const int32_t* rowptr, const int32_t* col,
                          double* val, hipStream_t s) {
    if (Nf <= 0 || Nc <= 0) return 0;
    hipLaunchKernelGGL(k_q2_last_child, dim3(nblk(Nf, 256)), dim3(256), 0, s, Nf, 2 * nf + 1, nf + 1, node_of_dof_f,
                       dir_f, dof_of_node_c, last);
    hipLaunchKernelGGL(k_q2_restrict_rows, dim3(nblk(Nc, 256)), dim3(256), 0, s, Nc, last, rowptr, col, val);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugamd
