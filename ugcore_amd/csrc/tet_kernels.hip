// HIP kernels of the tetrahedral (kind 2) hierarchy: FV1 convection-diffusion
// assembly, the CSR pattern from the tetrahedra, edge-based grid transfer, the
// transposed (adjoint) operators and the adjoint gradient.
//
// A level is a lattice of n^3 cells, W = n+1 points per side, each cell split
// into 6 tetrahedra given per cell as packed corner codes (4 x 3 bits, bit 0 x,
// bit 1 y, bit 2 z of the corner within the cell) in the tetrahedron's local
// vertex order. Level operators use the 27-diagonal lattice storage of the
// other kernels (diagonal d = (dx+1) + 3(dy+1) + 9(dz+1), wavefront tiles).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "elem_local.hpp"
#include "ugx_internal.hpp"

namespace ugamd {

namespace {

__device__ __forceinline__ int64_t tile_index(int d, int64_t i, int nd) {
    return ((i >> 6) * nd + d) * 64 + (i & 63);
}

__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ double uniform_open(uint64_t seed, uint64_t counter) {
    const uint64_t h = splitmix(splitmix(seed) ^ counter);
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// corner c of cell (i,j,k): lattice offsets
__device__ __forceinline__ void corner_off(int c, int& dx, int& dy, int& dz) {
    dx = c & 1;
    dy = (c >> 1) & 1;
    dz = (c >> 2) & 1;
}

// vertex position (coordinates) of level lattice point (i,j,k)
__device__ __forceinline__ void vertex_position(const TetDisc& P, int i, int j, int k, double* x) {
    const int64_t fi = (int64_t)i << P.shift, fj = (int64_t)j << P.shift, fk = (int64_t)k << P.shift;
    x[0] = (double)fi * P.scale;
    x[1] = (double)fj * P.scale;
    x[2] = (double)fk * P.scale;
    if (!P.jitter) return;
    if (i == 0 || j == 0 || k == 0 || i == P.n || j == P.n || k == P.n) return;
    const int64_t Wf = (int64_t)P.n_fine + 1;
    const int64_t l = (fk * Wf + fj) * Wf + fi;
    for (int d = 0; d < 3; ++d) {
        const double u = uniform_open(P.jseed, (uint64_t)(3 * l + d));
        x[d] += P.jamp * P.scale * (2.0 * u - 1.0);
    }
}

// the tetrahedron `code` of cell (ci,cj,ck): lattice coordinates of its corners
__device__ __forceinline__ void tet_corners(uint16_t code, int ci, int cj, int ck, int (*c)[3]) {
    for (int q = 0; q < 4; ++q) {
        int dx, dy, dz;
        corner_off((code >> (3 * q)) & 7, dx, dy, dz);
        c[q][0] = ci + dx;
        c[q][1] = cj + dy;
        c[q][2] = ck + dz;
    }
}

// diffusion coefficient at the integration point of scvf i: exp(theta) of the
// fine cell containing it, from the exact integer position 48*ip (fine units)
__device__ __forceinline__ double tet_kip(const TetDisc& P, const int (*c)[3], int i) {
    if (!P.theta) return 1.0;
    int64_t cell[3];
    for (int d = 0; d < 3; ++d) {
        auto Q = [&](int q) { return (int64_t)c[q][d] << P.shift; };
        const int8_t* f0 = kTetFace[kTetEdgeFace[i][0]];
        const int8_t* f1 = kTetFace[kTetEdgeFace[i][1]];
        const int64_t s48 = 6 * (Q(kTetEdge[i][0]) + Q(kTetEdge[i][1])) + 4 * (Q(f0[0]) + Q(f0[1]) + Q(f0[2])) +
                            3 * (Q(0) + Q(1) + Q(2) + Q(3)) + 4 * (Q(f1[0]) + Q(f1[1]) + Q(f1[2]));
        int64_t cc = s48 / 48;
        if (cc >= P.n_fine) cc = P.n_fine - 1;
        cell[d] = cc;
    }
    return exp(P.theta[(cell[2] * P.n_fine + cell[1]) * P.n_fine + cell[0]]);
}

__device__ __forceinline__ int dir27(const int* a, const int* b) {
    return (b[0] - a[0] + 1) + 3 * (b[1] - a[1] + 1) + 9 * (b[2] - a[2] + 1);
}

__device__ __forceinline__ int32_t slot(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, int32_t r,
                                        int32_t c) {
    int32_t lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo < hi) {
        const int32_t m = (lo + hi) >> 1;
        if (col[m] < c) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// cells of one colour (cell index parity per axis): no two share a vertex
__device__ __forceinline__ bool colour_cell(int n, int color, int64_t t, int& ci, int& cj, int& ck) {
    const int ox = color & 1, oy = (color >> 1) & 1, oz = (color >> 2) & 1;
    const int64_t nx = (n - ox + 1) / 2, ny = (n - oy + 1) / 2, nz = (n - oz + 1) / 2;
    if (t >= nx * ny * nz) return false;
    ci = 2 * (int)(t % nx) + ox;
    cj = 2 * (int)((t / nx) % ny) + oy;
    ck = 2 * (int)(t / (nx * ny)) + oz;
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// neighbour masks and CSR pattern (all vertex pairs of a tetrahedron couple,
// the diagonal always exists: nnz = V + 2E)
// ---------------------------------------------------------------------------
__global__ void k_tet_mask(int n, const uint16_t* __restrict__ cell_tets, uint32_t* mask) {
    const int64_t cells = (int64_t)n * n * n;
    const int64_t W = n + 1;
    for (int64_t cell = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; cell < cells;
         cell += (int64_t)gridDim.x * blockDim.x) {
        const int ci = (int)(cell % n), cj = (int)((cell / n) % n), ck = (int)(cell / ((int64_t)n * n));
        for (int t = 0; t < 6; ++t) {
            int c[4][3];
            tet_corners(cell_tets[cell * 6 + t], ci, cj, ck, c);
            for (int a = 0; a < 4; ++a) {
                uint32_t bits = 1u << 13;
                for (int b = 0; b < 4; ++b) bits |= 1u << dir27(c[a], c[b]);
                atomicOr(&mask[((int64_t)c[a][2] * W + c[a][1]) * W + c[a][0]], bits);
            }
        }
    }
}

__global__ void k_pattern_count_mask(int64_t N, const int32_t* __restrict__ lat_of_dof, const uint32_t* __restrict__ mask,
                                     int32_t* cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        cnt[i] = __popc(mask[lat_of_dof[i]] | (1u << 13));
}

__global__ void k_pattern_fill_mask(int W, int64_t N, const int32_t* __restrict__ lat_of_dof,
                                    const int32_t* __restrict__ dof_of_lat, const uint32_t* __restrict__ mask,
                                    const int32_t* __restrict__ rowptr, int32_t* __restrict__ col) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t l = lat_of_dof[i];
        const uint32_t m = mask[l] | (1u << 13);
        int32_t c[27];
        int k = 0;
        for (int d = 0; d < 27; ++d) {
            if (!(m & (1u << d))) continue;
            const int64_t off = (d % 3 - 1) + (int64_t)((d / 3) % 3 - 1) * W + (int64_t)(d / 9 - 1) * W * W;
            int32_t v = dof_of_lat[l + off];
            // insertion sort by DoF
            int q = k++;
            while (q > 0 && c[q - 1] > v) {
                c[q] = c[q - 1];
                --q;
            }
            c[q] = v;
        }
        const int32_t base = rowptr[i];
        for (int q = 0; q < k; ++q) col[base + q] = c[q];
    }
}

// ---------------------------------------------------------------------------
// assembly: one thread per cell of one colour, its 6 tetrahedra in order;
// targets: CSR in DoF numbering (+ rhs in DoF numbering) and/or the 27
// lattice diagonals (tiled, lattice rows)
// ---------------------------------------------------------------------------
template <typename TM>
__global__ __launch_bounds__(128) void k_assemble_fv1_tet(TetDisc P, int color, const uint16_t* __restrict__ cell_tets,
                                                          const int32_t* __restrict__ dof_of_lat,
                                                          const int32_t* __restrict__ rowptr,
                                                          const int32_t* __restrict__ col, TM* csr_val, TM* dia,
                                                          double* rhs) {
    int ci, cj, ck;
    if (!colour_cell(P.n, color, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, ci, cj, ck)) return;
    const int64_t W = P.n + 1;
    const int64_t cell = ((int64_t)ck * P.n + cj) * P.n + ci;
    for (int t = 0; t < 6; ++t) {
        int c[4][3];
        tet_corners(cell_tets[cell * 6 + t], ci, cj, ck, c);
        double X[4][3], kip[6], Ke[4][4], vol[4];
        int64_t lat[4];
        for (int q = 0; q < 4; ++q) {
            vertex_position(P, c[q][0], c[q][1], c[q][2], X[q]);
            lat[q] = ((int64_t)c[q][2] * W + c[q][1]) * W + c[q][0];
        }
        for (int i = 0; i < 6; ++i) kip[i] = tet_kip(P, c, i);
        fv1_tet_element(X, kip, P.vel, true, Ke, vol);
        for (int a = 0; a < 4; ++a) {
            if (dia) {
                for (int b = 0; b < 4; ++b) {
                    const int64_t k = tile_index(dir27(c[a], c[b]), lat[a], 27);
                    dia[k] = (TM)((double)dia[k] + Ke[a][b]);
                }
            }
            if (csr_val) {
                const int32_t ra = dof_of_lat[lat[a]];
                for (int b = 0; b < 4; ++b) {
                    const int32_t k = slot(rowptr, col, ra, dof_of_lat[lat[b]]);
                    csr_val[k] = (TM)((double)csr_val[k] + Ke[a][b]);
                }
                if (rhs) rhs[ra] += P.f * vol[a];
            }
        }
    }
}

// dJ/dtheta per fine cell: -sum over its tetrahedra of lambda_e^T D_e u_e,
// D_e the diffusion part (all integration points of a top-level tetrahedron
// lie in its own cell), interior rows only; u, lambda in DoF numbering
__global__ __launch_bounds__(128) void k_adjoint_grad_fv1_tet(TetDisc P, const uint16_t* __restrict__ cell_tets,
                                                              const int32_t* __restrict__ dof_of_lat, int64_t n_inner,
                                                              const double* __restrict__ u,
                                                              const double* __restrict__ lam, double* grad) {
    const int64_t cells = (int64_t)P.n * P.n * P.n;
    const int64_t cell = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    const int ci = (int)(cell % P.n), cj = (int)((cell / P.n) % P.n), ck = (int)(cell / ((int64_t)P.n * P.n));
    const int64_t W = P.n + 1;
    const double k = P.theta ? exp(P.theta[cell]) : 1.0;
    double g = 0.0;
    for (int t = 0; t < 6; ++t) {
        int c[4][3];
        tet_corners(cell_tets[cell * 6 + t], ci, cj, ck, c);
        double X[4][3], kip[6], Ke[4][4];
        int32_t dof[4];
        for (int q = 0; q < 4; ++q) {
            vertex_position(P, c[q][0], c[q][1], c[q][2], X[q]);
            dof[q] = dof_of_lat[((int64_t)c[q][2] * W + c[q][1]) * W + c[q][0]];
        }
        for (int i = 0; i < 6; ++i) kip[i] = k;
        fv1_tet_element(X, kip, P.vel, false, Ke, nullptr);
        double acc = 0.0;
        for (int a = 0; a < 4; ++a) {
            if (dof[a] >= n_inner) continue;
            double row = 0.0;
            for (int b = 0; b < 4; ++b) row += Ke[a][b] * u[dof[b]];
            acc += lam[dof[a]] * row;
        }
        g -= acc;
    }
    grad[cell] = g;
}

// ---------------------------------------------------------------------------
// grid transfer (StdTransfer P1, std_transfer_impl.h:45-168, with the
// Dirichlet adjustments of lagrange_dirichlet_boundary_impl.h:485-740):
// a fine vertex copies its coarse parent or averages the two endpoints of its
// parent edge; fine Dirichlet rows of P are zero except vertex copies; coarse
// Dirichlet rows of R = P^T are the injection from the copy child
// ---------------------------------------------------------------------------
template <typename TV>
__global__ __launch_bounds__(256) void k_prolong_add_tet(int Wc, int Wf, int64_t n_f, const TV* __restrict__ scc,
                                                         const uint8_t* __restrict__ pcode,
                                                         const uint8_t* __restrict__ dir_f, TV* __restrict__ st,
                                                         TV* __restrict__ sc) {
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i >= n_f) return;
    const int x = (int)(i % Wf), y = (int)((i / Wf) % Wf), z = (int)(i / ((int64_t)Wf * Wf));
    const int code = pcode[i];
    double v = 0.0;
    if (code == 0) {
        v = (double)scc[((int64_t)(z >> 1) * Wc + (y >> 1)) * Wc + (x >> 1)];
    } else if (!dir_f[i]) {
        const int d = code + 13;
        const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
        const int64_t a = ((int64_t)((z - dz) >> 1) * Wc + ((y - dy) >> 1)) * Wc + ((x - dx) >> 1);
        const int64_t b = ((int64_t)((z + dz) >> 1) * Wc + ((y + dy) >> 1)) * Wc + ((x + dx) >> 1);
        v = 0.5 * (double)scc[a] + 0.5 * (double)scc[b];
    }
    // c += t in the vector precision (mg_solver_impl.hpp:1905)
    const TV t = (TV)v;
    st[i] = t;
    sc[i] = sc[i] + t;
}

template <typename TV>
__global__ __launch_bounds__(256) void k_restrict_tet(int Wc, int Wf, int64_t n_c, const TV* __restrict__ sdf,
                                                      const uint8_t* __restrict__ pcode,
                                                      const uint8_t* __restrict__ dir_c,
                                                      const uint8_t* __restrict__ dir_f, TV* __restrict__ sdc) {
    const int64_t c = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (c >= n_c) return;
    const int x = (int)(c % Wc), y = (int)((c / Wc) % Wc), z = (int)(c / ((int64_t)Wc * Wc));
    const int64_t f0 = ((int64_t)(2 * z) * Wf + 2 * y) * Wf + 2 * x;
    if (dir_c[c]) {
        sdc[c] = sdf[f0];
        return;
    }
    double acc = (double)sdf[f0];
    for (int d = 0; d < 27; ++d) {
        if (d == 13) continue;
        const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
        const int fx = 2 * x + dx, fy = 2 * y + dy, fz = 2 * z + dz;
        if (fx < 0 || fy < 0 || fz < 0 || fx >= Wf || fy >= Wf || fz >= Wf) continue;
        const int64_t f = ((int64_t)fz * Wf + fy) * Wf + fx;
        const int code = d > 13 ? d - 13 : 13 - d;
        if (pcode[f] != code || dir_f[f]) continue;
        acc += 0.5 * (double)sdf[f];
    }
    sdc[c] = (TV)acc;
}

// ---------------------------------------------------------------------------
// adjoint operators
// ---------------------------------------------------------------------------
// transpose of the Dirichlet-eliminated 27-diagonal lattice operator:
// AT(i, i+off_d) = A_elim(i+off_d, i) = A[26-d][i+off_d] for interior i and j,
// 0 for interior i and Dirichlet j, identity rows for Dirichlet i
template <typename T>
__global__ void k_dia_transpose_elim(int W, int64_t n, const uint8_t* __restrict__ dir, const T* __restrict__ A,
                                     T* __restrict__ AT) {
    const int64_t W2 = (int64_t)W * W;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        for (int d = 0; d < 27; ++d) {
            const int64_t off = (d % 3 - 1) + (int64_t)((d / 3) % 3 - 1) * W + (int64_t)(d / 9 - 1) * W2;
            const int64_t j = i + off;
            T v = (T)0;
            if (dir[i]) {
                v = d == 13 ? (T)1 : (T)0;
            } else if (j >= 0 && j < n && !dir[j]) {
                v = A[tile_index(26 - d, j, 27)];
            }
            AT[tile_index(d, i, 27)] = v;
        }
    }
}

// for a structurally symmetric CSR pattern: tp[p] = position of (c, r) for
// the entry p = (r, c); bad is set when an entry has no mirror
__global__ void k_csr_transpose_map(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                    int32_t* __restrict__ tp, int* bad) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        for (int32_t p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            const int32_t c = col[p];
            const int32_t q = slot(rowptr, col, c, (int32_t)r);
            if (q >= rowptr[c + 1] || col[q] != (int32_t)r) {
                atomicOr(bad, 1);
                tp[p] = p;
            } else {
                tp[p] = q;
            }
        }
    }
}

// values of A_elim^T on the same pattern: DoFs >= n_inner are Dirichlet
__global__ void k_csr_adjoint_values(int64_t n, int64_t n_inner, const int32_t* __restrict__ rowptr,
                                     const int32_t* __restrict__ col, const int32_t* __restrict__ tp,
                                     const double* __restrict__ val, double* __restrict__ valT) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        for (int32_t p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            const int64_t c = col[p];
            double v;
            if (c >= n_inner) v = (c == r) ? 1.0 : 0.0;
            else if (r >= n_inner) v = 0.0;
            else v = val[tp[p]];
            valT[p] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    return (unsigned)(g < 1 ? 1 : g);
}
#define UGX_TET_CHECK() return hipGetLastError() == hipSuccess ? 0 : -1

int tet_mask(int n, const uint16_t* cell_tets, uint32_t* mask, hipStream_t s) {
    const int64_t cells = (int64_t)n * n * n;
    hipLaunchKernelGGL(k_tet_mask, dim3(grid_for(cells, 256)), dim3(256), 0, s, n, cell_tets, mask);
    UGX_TET_CHECK();
}

int pattern_count_mask(int64_t N, const int32_t* lat_of_dof, const uint32_t* mask, int32_t* cnt, hipStream_t s) {
    hipLaunchKernelGGL(k_pattern_count_mask, dim3(grid_for(N, 256)), dim3(256), 0, s, N, lat_of_dof, mask, cnt);
    UGX_TET_CHECK();
}

int pattern_fill_mask(int W, int64_t N, const int32_t* lat_of_dof, const int32_t* dof_of_lat, const uint32_t* mask,
                      const int32_t* rowptr, int32_t* col, hipStream_t s) {
    hipLaunchKernelGGL(k_pattern_fill_mask, dim3(grid_for(N, 256)), dim3(256), 0, s, W, N, lat_of_dof, dof_of_lat,
                       mask, rowptr, col);
    UGX_TET_CHECK();
}

template <typename TM>
int assemble_fv1_tet(const TetDisc& P, const uint16_t* cell_tets, const int32_t* dof_of_lat, const int32_t* rowptr,
                     const int32_t* col, TM* csr_val, TM* dia, double* rhs, hipStream_t s) {
    for (int color = 0; color < 8; ++color) {
        const int ox = color & 1, oy = (color >> 1) & 1, oz = (color >> 2) & 1;
        const int64_t per = (int64_t)((P.n - ox + 1) / 2) * ((P.n - oy + 1) / 2) * ((P.n - oz + 1) / 2);
        if (per == 0) continue;
        hipLaunchKernelGGL(k_assemble_fv1_tet<TM>, dim3((unsigned)((per + 127) / 128)), dim3(128), 0, s, P, color,
                           cell_tets, dof_of_lat, rowptr, col, csr_val, dia, rhs);
    }
    UGX_TET_CHECK();
}
template int assemble_fv1_tet<double>(const TetDisc&, const uint16_t*, const int32_t*, const int32_t*, const int32_t*,
                                      double*, double*, double*, hipStream_t);

int adjoint_grad_fv1_tet(const TetDisc& P, const uint16_t* cell_tets, const int32_t* dof_of_lat, int64_t n_inner,
                         const double* u, const double* lambda, double* grad, hipStream_t s) {
    const int64_t cells = (int64_t)P.n * P.n * P.n;
    hipLaunchKernelGGL(k_adjoint_grad_fv1_tet, dim3((unsigned)((cells + 127) / 128)), dim3(128), 0, s, P, cell_tets,
                       dof_of_lat, n_inner, u, lambda, grad);
    UGX_TET_CHECK();
}

template <typename TV>
int prolong_add_tet(int Wc, int Wf, int64_t n_f, const TV* scc, const uint8_t* pcode, const uint8_t* dir_f, TV* st,
                    TV* sc, hipStream_t s) {
    if (n_f <= 0) return 0;
    hipLaunchKernelGGL(k_prolong_add_tet<TV>, dim3((unsigned)((n_f + 255) / 256)), dim3(256), 0, s, Wc, Wf, n_f, scc,
                       pcode, dir_f, st, sc);
    UGX_TET_CHECK();
}
template int prolong_add_tet<double>(int, int, int64_t, const double*, const uint8_t*, const uint8_t*, double*,
                                     double*, hipStream_t);
template int prolong_add_tet<float>(int, int, int64_t, const float*, const uint8_t*, const uint8_t*, float*, float*,
                                    hipStream_t);

template <typename TV>
int restrict_tet(int Wc, int Wf, int64_t n_c, const TV* sdf, const uint8_t* pcode, const uint8_t* dir_c,
                 const uint8_t* dir_f, TV* sdc, hipStream_t s) {
    if (n_c <= 0) return 0;
    hipLaunchKernelGGL(k_restrict_tet<TV>, dim3((unsigned)((n_c + 255) / 256)), dim3(256), 0, s, Wc, Wf, n_c, sdf,
                       pcode, dir_c, dir_f, sdc);
    UGX_TET_CHECK();
}
template int restrict_tet<double>(int, int, int64_t, const double*, const uint8_t*, const uint8_t*, const uint8_t*,
                                  double*, hipStream_t);
template int restrict_tet<float>(int, int, int64_t, const float*, const uint8_t*, const uint8_t*, const uint8_t*,
                                 float*, hipStream_t);

template <typename T>
int dia_transpose_elim(int W, int64_t n, const uint8_t* dir, const T* A, T* AT, hipStream_t s) {
    hipLaunchKernelGGL(k_dia_transpose_elim<T>, dim3(grid_for(n, 256)), dim3(256), 0, s, W, n, dir, A, AT);
    UGX_TET_CHECK();
}
template int dia_transpose_elim<double>(int, int64_t, const uint8_t*, const double*, double*, hipStream_t);

int csr_transpose_map(int64_t n, const int32_t* rowptr, const int32_t* col, int32_t* tp, int* bad, hipStream_t s) {
    hipLaunchKernelGGL(k_csr_transpose_map, dim3(grid_for(n, 256)), dim3(256), 0, s, n, rowptr, col, tp, bad);
    UGX_TET_CHECK();
}

int csr_adjoint_values(int64_t n, int64_t n_inner, const int32_t* rowptr, const int32_t* col, const int32_t* tp,
                       const double* val, double* valT, hipStream_t s) {
    hipLaunchKernelGGL(k_csr_adjoint_values, dim3(grid_for(n, 256)), dim3(256), 0, s, n, n_inner, rowptr, col, tp,
                       val, valT);
    UGX_TET_CHECK();
}

// coordinates of the DoFs of a level, in DoF order (dim per DoF): the
// lattice points of the unit square/cube, for the tetrahedral hierarchy the
// vertex positions the assembly uses (jittered on the top level)
__global__ void k_dof_positions(TetDisc P, int kind, int dim, int W, int64_t N, const int32_t* __restrict__ lat_of_dof,
                                double* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t l = lat_of_dof[i];
    const int x = (int)(l % W), y = (int)((l / W) % W), z = dim == 3 ? (int)(l / ((int64_t)W * W)) : 0;
    if (kind == 2) {
        double X[3];
        vertex_position(P, x, y, z, X);
        for (int d = 0; d < 3; ++d) out[3 * i + d] = X[d];
        return;
    }
    const double h = 1.0 / (double)(W - 1);
    out[dim * i] = x * h;
    out[dim * i + 1] = y * h;
    if (dim == 3) out[dim * i + 2] = z * h;
}

int dof_positions(const TetDisc& P, int kind, int dim, int W, int64_t N, const int32_t* lat_of_dof, double* out,
                  hipStream_t s) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(k_dof_positions, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, P, kind, dim, W, N,
                       lat_of_dof, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugamd
