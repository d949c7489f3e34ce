// Galerkin coarse-level operators A_c = R A_f P on the lattice hierarchies.
//
// The reference forms them with a sparse triple product per level,
// AddMultiplyOf(*lc.A, *R, *A_f, *P) in AssembledMultiGridCycle::
// init_rap_operator (lib_disc/operator/linear_operator/multi_grid_solver/
// mg_solver_impl.hpp:829-1013, selected by set_rap, mg_solver.h:149), starting
// from the surface matrix on the top level. The transfer operators are the P1
// ones of StdTransfer (std_transfer_impl.h:45-168) with the Dirichlet
// adjustments of DirichletBoundary (lagrange_dirichlet_boundary_impl.h:485-740):
//   * P(G, J) for a fine vertex G that is not Dirichlet: the P1 weight of coarse
//     vertex J (hexahedra: trilinear, prod_k (1 or 1/2); triangles split along
//     (1,1): 1 for the copy of J, 1/2 for the midpoints of the edges at J);
//   * a fine Dirichlet vertex keeps only the injection P(2J, J) = 1 when it is
//     the copy of coarse vertex J, its row is zero otherwise;
//   * R = P^T, with the row of a coarse Dirichlet DoF replaced by the injection
//     from its fine copy.
// With these, a coarse Dirichlet row of R A_f P is the fine Dirichlet row
// (identity) times P, i.e. the identity row, and an interior coarse row i
// collects, over the fine neighbours F = 2i + d (all interior), R(i, F)
// A_f(F, G) P(G, J) for the fine stencil G of F. Every coarse column J lies
// within one lattice step of i, so the result fits the same 27- (3D) or
// 9-point (2D) diagonal storage as the assembled level operators.
//
// One thread per coarse lattice point; the sums run over d, then over the fine
// stencil offsets e that can reach J, in a fixed order.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ugx_internal.hpp"

namespace ugamd {

namespace {

// P1 weight for a fine-minus-twice-coarse offset diff (each component in
// {-1, 0, 1}); the triangles of the 2D lattice are split along (1,1)
template <int DIM>
__device__ __forceinline__ double p1_weight(int dx, int dy, int dz) {
    if (DIM == 2) {
        if (dx * dy < 0) return 0.0;                  // no edge along (1,-1)
        return (dx == 0 && dy == 0) ? 1.0 : 0.5;      // the copy, or an edge midpoint at J
    }
    return (dx != 0 ? 0.5 : 1.0) * (dy != 0 ? 0.5 : 1.0) * (dz != 0 ? 0.5 : 1.0);
}

template <int DIM>
__global__ __launch_bounds__(256) void k_rap_dia(int Wf, int Wc, const uint8_t* __restrict__ dir_f,
                                                 const uint8_t* __restrict__ dir_c, const double* __restrict__ Af,
                                                 double* __restrict__ Ac) {
    constexpr int ND = DIM == 3 ? 27 : 9;
    const int64_t nc = DIM == 3 ? (int64_t)Wc * Wc * Wc : (int64_t)Wc * Wc;
    const int64_t ci = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (ci >= nc) return;
    const int cx = (int)(ci % Wc), cy = (int)((ci / Wc) % Wc), cz = DIM == 3 ? (int)(ci / ((int64_t)Wc * Wc)) : 0;
    // the lattice boundary is the Dirichlet boundary of these hierarchies
    const bool boundary = cx == 0 || cx == Wc - 1 || cy == 0 || cy == Wc - 1 || (DIM == 3 && (cz == 0 || cz == Wc - 1));
    if (dir_c[ci] || boundary) {
        for (int o = 0; o < ND; ++o) Ac[dia_tiled_index(o, ci, ND)] = o == ND / 2 ? 1.0 : 0.0;
        return;
    }
    const int64_t Wf2 = (int64_t)Wf * Wf;
    for (int o = 0; o < ND; ++o) {
        const int ox = o % 3 - 1, oy = (o / 3) % 3 - 1, oz = DIM == 3 ? o / 9 - 1 : 0;
        const int jx = cx + ox, jy = cy + oy, jz = cz + oz;
        const bool inside = jx >= 0 && jx < Wc && jy >= 0 && jy < Wc && (DIM == 2 || (jz >= 0 && jz < Wc));
        double acc = 0.0;
        if (inside) {
            for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const double wf = p1_weight<DIM>(dx, dy, dz);   // R(i, F) = P(F, i)
                        if (wf == 0.0) continue;
                        const int fx = 2 * cx + dx, fy = 2 * cy + dy, fz = 2 * cz + dz;
                        const int64_t fi = DIM == 3 ? (int64_t)fz * Wf2 + (int64_t)fy * Wf + fx : (int64_t)fy * Wf + fx;
                        // fine stencil offsets e with |d + e - 2o| <= 1 in every component
                        const int ex0 = max(-1, 2 * ox - dx - 1), ex1 = min(1, 2 * ox - dx + 1);
                        const int ey0 = max(-1, 2 * oy - dy - 1), ey1 = min(1, 2 * oy - dy + 1);
                        const int ez0 = DIM == 3 ? max(-1, 2 * oz - dz - 1) : 0;
                        const int ez1 = DIM == 3 ? min(1, 2 * oz - dz + 1) : 0;
                        for (int ez = ez0; ez <= ez1; ++ez)
                            for (int ey = ey0; ey <= ey1; ++ey)
                                for (int ex = ex0; ex <= ex1; ++ex) {
                                    const int e = DIM == 3 ? (ez + 1) * 9 + (ey + 1) * 3 + (ex + 1)
                                                           : (ey + 1) * 3 + (ex + 1);
                                    const double a = Af[dia_tiled_index(e, fi, ND)];
                                    if (a == 0.0) continue;
                                    const int64_t gi = DIM == 3 ? fi + (int64_t)ez * Wf2 + (int64_t)ey * Wf + ex
                                                                : fi + (int64_t)ey * Wf + ex;
                                    const int qx = dx + ex - 2 * ox, qy = dy + ey - 2 * oy, qz = dz + ez - 2 * oz;
                                    double wp;
                                    if (dir_f[gi])
                                        wp = (qx == 0 && qy == 0 && qz == 0) ? 1.0 : 0.0;   // injection only
                                    else
                                        wp = p1_weight<DIM>(qx, qy, qz);
                                    if (wp != 0.0) acc += wf * a * wp;
                                }
                    }
        }
        Ac[dia_tiled_index(o, ci, ND)] = acc;
    }
}

}  // namespace

int rap_dia(int dim, int Wf, int Wc, const uint8_t* dir_f, const uint8_t* dir_c, const double* Af, double* Ac,
            hipStream_t s) {
    if (Wf != 2 * Wc - 1 || Wc < 2) return -1;
    const int64_t nc = dim == 3 ? (int64_t)Wc * Wc * Wc : (int64_t)Wc * Wc;
    const dim3 grid((unsigned)((nc + 255) / 256));
    if (dim == 3)
        hipLaunchKernelGGL(k_rap_dia<3>, grid, dim3(256), 0, s, Wf, Wc, dir_f, dir_c, Af, Ac);
    else if (dim == 2)
        hipLaunchKernelGGL(k_rap_dia<2>, grid, dim3(256), 0, s, Wf, Wc, dir_f, dir_c, Af, Ac);
    else
        return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugamd
