// Row-by-row assembly of the lattice operators: the surface operator into
// its CSR, the level operators into their tiled diagonals.
//
// The reference assembles element by element and adds each local matrix into
// the global one (elem_disc_assemble_util.h:326-413, AddLocalMatrixToGlobal in
// local_algebra.h:766-789). On the structured lattice every entry (p, q) is the
// sum of the local entries of the elements that contain both points, so a
// group of 32 lanes per lattice point p forms row p directly: lane j takes
// the neighbour q = p + (j%3-1, j/3%3-1, j/9-1), sums the contributions of the
// (at most 8) common elements and writes the entry once. Nothing is
// accumulated in memory (no colouring, no read-modify-write, no clearing
// pass), and the CSR position of the entry is its rank among the row's
// neighbour DoFs, because a row of the lattice pattern holds exactly those
// DoFs sorted by number; the column indices are not read.
//
// Q1 on the unit cube: every element is the cube of side h, so its local
// matrix is k_e h K with K = sum_q w_q grad(phi_a) . grad(phi_b) on the
// reference cube (2x2x2 Gauss points; lagrange.h Q1 shape functions) and its
// load vector f h^3 F with F_a = sum_q w_q phi_a; K and F are computed on the
// host by that quadrature. k_e = exp(theta_e) per fine element. On a coarser
// level an element spans sub^3 fine elements and the coefficient is taken at
// each quadrature point: the local matrix is h sum_q k(x_q) G_q with
// G_q = w_q grad(phi_a) . grad(phi_b) at that point.
//
// Table layout (kQ1Tab doubles): G_q[a][b] at 64q + 8a + b, K at 512 + 8a + b,
// F at 576 + a.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>

#include "ugx_internal.hpp"

namespace ugamd {

namespace {

constexpr int kG = 0, kK = 512, kF = 576;

inline unsigned nblk(int64_t n, int t) {
    const int64_t b = (n + t - 1) / t;
    return (unsigned)(b < 1 ? 1 : (b > 1048576 ? 1048576 : b));
}

// local index of corner offset (ox, oy, oz) in the reference hexahedron
__device__ __forceinline__ int hex_local(int ox, int oy, int oz) { return 4 * oz + (oy ? 3 - ox : ox); }

// rank of v among the 27 direction lanes of the lane's half-wave (INT_MAX
// never counts)
__device__ __forceinline__ int half_rank(int32_t v) {
    int rank = 0;
#pragma unroll
    for (int i = 0; i < 27; ++i) rank += __shfl(v, i, 32) < v ? 1 : 0;
    return rank;
}

// The kernels below run one workgroup per lattice line (py, pz) (2D: one per
// py) and eight points of it at a time, one per 32-lane half-wave; lane j < 27
// of a half-wave owns direction (j%3-1, j/3%3-1, j/9-1), the centre is lane 13.
struct LineLane {
    int py, pz, dx, qy, qz;
    bool yz;   // the lane's neighbour line exists (and, 2D, is not an anti-diagonal)
    __device__ LineLane(int dim, int W, int line, int j) {
        py = dim == 3 ? line % W : line;
        pz = dim == 3 ? line / W : 0;
        dx = j % 3 - 1;
        const int dy = (j / 3) % 3 - 1, dz = j / 9 - 1;
        qy = py + dy;
        qz = pz + dz;
        yz = j < 27 && qy >= 0 && qy < W && (dim == 3 ? (qz >= 0 && qz < W) : (dz == 0 && dx * dy != -1));
    }
    __device__ bool valid(int W, int px) const { return yz && px + dx >= 0 && px + dx < W; }
};

// rp_lat[p] = rowptr[dof_of_lat[p]]: the row start of every lattice point
__global__ void k_rowptr_of_lattice(int64_t npts, const int32_t* __restrict__ dof_of_lat,
                                    const int32_t* __restrict__ rowptr, int32_t* __restrict__ rp_lat) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npts; p += (int64_t)gridDim.x * blockDim.x)
        rp_lat[p] = rowptr[dof_of_lat[p]];
}

// The 9 lattice lines around the workgroup's line (line c = (dz+1)*3 + dy+1,
// INT_MAX where it is outside the lattice or, 2D, off the plane) and the row
// starts of the line's points, staged in LDS (STAGED) or read from memory
template <bool STAGED>
struct LineData {
    const int32_t* dof;   // STAGED: 9 lines of W; else dof_of_lat
    const int32_t* rp;    // STAGED: W; else rp_lat
    int W;
    int64_t pline;
    __device__ LineData(int dim, int W_, int py, int pz, const int32_t* __restrict__ dof_of_lat,
                        const int32_t* __restrict__ rp_lat, int32_t* lds)
        : W(W_), pline(((int64_t)pz * W_ + py) * W_) {
        if (!STAGED) {
            dof = dof_of_lat;
            rp = rp_lat;
            return;
        }
        for (int i = threadIdx.x; i < 10 * W; i += blockDim.x) {
            const int c = i / W, x = i - c * W;
            if (c == 9) {
                lds[i] = rp_lat[pline + x];
                continue;
            }
            const int qy = py + c % 3 - 1, qz = pz + c / 3 - 1;
            const bool ok = qy >= 0 && qy < W && (dim == 3 ? (qz >= 0 && qz < W) : c / 3 == 1);
            lds[i] = ok ? dof_of_lat[((int64_t)qz * W + qy) * W + x] : INT_MAX;
        }
        dof = lds;
        rp = lds + 9 * W;
    }
    // DoF of the lane's neighbour of point px (the lane's line, qx = px + dx)
    __device__ int32_t nd(const LineLane& L, int j, int px) const {
        if (STAGED) return dof[(j / 3) * W + px + L.dx];
        return dof[((int64_t)L.qz * W + L.qy) * W + px + L.dx];
    }
    __device__ int32_t row_start(int px) const { return STAGED ? rp[px] : rp[pline + px]; }
};

// col of every row of the lattice pattern: the neighbour DoFs of the row's
// point in ascending order (position = rank). Dynamic LDS (STAGED): 10 W ints
template <bool STAGED>
__global__ __launch_bounds__(256) void k_lattice_pattern(int dim, int W, const int32_t* __restrict__ dof_of_lat,
                                                         const int32_t* __restrict__ rp_lat,
                                                         int32_t* __restrict__ col) {
    extern __shared__ int32_t lines[];
    const int j = threadIdx.x & 31, half = threadIdx.x >> 5;
    const LineLane L(dim, W, blockIdx.x, j);
    const LineData<STAGED> D(dim, W, L.py, L.pz, dof_of_lat, rp_lat, lines);
    if (STAGED) __syncthreads();
    for (int px0 = half; px0 < W; px0 += 16) {
        const int px1 = px0 + 8;
        // loads of both points first (the second point's latency overlaps)
        const bool ok0 = L.valid(W, px0), ok1 = px1 < W && L.valid(W, px1);
        const int32_t nd0 = ok0 ? D.nd(L, j, px0) : INT_MAX;
        const int32_t nd1 = ok1 ? D.nd(L, j, px1) : INT_MAX;
        const int32_t rp0 = D.row_start(px0), rp1 = px1 < W ? D.row_start(px1) : 0;
        const int r0 = half_rank(nd0);
        if (ok0) col[rp0 + r0] = nd0;
        if (px1 < W) {
            const int r1 = half_rank(nd1);
            if (ok1) col[rp1 + r1] = nd1;
        }
    }
}

// the sum over the (at most 8) elements around p = (px, L.py, L.pz) that also
// contain the lane's neighbour q of k_e * table(a, b) (a, b: local indices of
// p, q in e), kl[2*oz + oy][ex] = k of element (ex, py-1+oy, pz-1+oz); f: the
// sum of F_a over those elements (used for the centre lane)
__device__ __forceinline__ double q1_entry(int n, int px, const LineLane& L, const double* kl, const double* st,
                                           double& f) {
    const int qx = px + L.dx;
    double v = 0.0;
    f = 0.0;
#pragma unroll
    for (int oz = 0; oz < 2; ++oz)
#pragma unroll
        for (int oy = 0; oy < 2; ++oy)
#pragma unroll
            for (int ox = 0; ox < 2; ++ox) {
                const int ex = px - 1 + ox, ey = L.py - 1 + oy, ez = L.pz - 1 + oz;
                const int bx = qx - ex, by = L.qy - ey, bz = L.qz - ez;
                const bool use = ex >= 0 && ex < n && ey >= 0 && ey < n && ez >= 0 && ez < n && bx >= 0 && bx <= 1 &&
                                 by >= 0 && by <= 1 && bz >= 0 && bz <= 1;
                if (use) {
                    const int a = hex_local(1 - ox, 1 - oy, 1 - oz);
                    v = fma(kl[(2 * oz + oy) * n + ex], st[8 * a + hex_local(bx, by, bz)], v);
                    f += st[64 + a];
                }
            }
    return v;
}

// stage k of the element lines (py-1+oy, pz-1+oz) (0 outside the lattice)
__device__ __forceinline__ void stage_element_lines(int n, int py, int pz, const double* __restrict__ kel,
                                                    double* kl) {
    for (int i = threadIdx.x; i < 4 * n; i += blockDim.x) {
        const int c = i / n, ex = i - c * n, ey = py - 1 + (c & 1), ez = pz - 1 + (c >> 1);
        kl[i] = (ey >= 0 && ey < n && ez >= 0 && ez < n) ? (kel ? kel[((int64_t)ez * n + ey) * n + ex] : 1.0) : 0.0;
    }
}

// Q1 operator (and load vector) of the top level into the CSR of the lattice
// pattern; kel: exp(theta) per element or null (k=1). Dynamic LDS: 4n doubles
// (element coefficients) then, STAGED, 10 W ints (lines)
template <bool STAGED>
__global__ __launch_bounds__(256) void k_q1_lattice_csr(int n, double h, double fh3, const double* __restrict__ tab,
                                                        const double* __restrict__ kel,
                                                        const int32_t* __restrict__ dof_of_lat,
                                                        const int32_t* __restrict__ rp_lat, double* __restrict__ val,
                                                        double* __restrict__ rhs) {
    extern __shared__ double kl[];
    __shared__ double st[72];
    const int W = n + 1;
    const int j = threadIdx.x & 31, half = threadIdx.x >> 5;
    const LineLane L(3, W, blockIdx.x, j);
    for (int i = threadIdx.x; i < 72; i += blockDim.x) st[i] = tab[kK + i];
    stage_element_lines(n, L.py, L.pz, kel, kl);
    const LineData<STAGED> D(3, W, L.py, L.pz, dof_of_lat, rp_lat, reinterpret_cast<int32_t*>(kl + 4 * n));
    __syncthreads();
    for (int px0 = half; px0 < W; px0 += 16) {
        const int px1 = px0 + 8;
        const bool ok0 = L.valid(W, px0), ok1 = px1 < W && L.valid(W, px1);
        const int32_t nd0 = ok0 ? D.nd(L, j, px0) : INT_MAX;
        const int32_t nd1 = ok1 ? D.nd(L, j, px1) : INT_MAX;
        const int32_t rp0 = D.row_start(px0), rp1 = px1 < W ? D.row_start(px1) : 0;
        double f;
        const int r0 = half_rank(nd0);
        if (ok0) {
            val[rp0 + r0] = h * q1_entry(n, px0, L, kl, st, f);
            if (j == 13 && rhs) rhs[nd0] = fh3 * f;
        }
        if (px1 < W) {
            const int r1 = half_rank(nd1);
            if (ok1) {
                val[rp1 + r1] = h * q1_entry(n, px1, L, kl, st, f);
                if (j == 13 && rhs) rhs[nd1] = fh3 * f;
            }
        }
    }
}

// Q1 operator of a level (n elements per axis, sub = fine elements per element
// and axis) into the tiled diagonals (lattice numbering, 27 diagonals,
// direction (dx,dy,dz) = diagonal (dz+1)*9 + (dy+1)*3 + dx+1; 0 where the
// neighbour is outside). kfine: exp(theta) per fine element or null (k=1).
// With sub > 1 the coefficient is taken at each quadrature point:
// kq[(2*oz + oy) * 8n + 8 ex + q] holds it for quadrature point q of element
// (ex, py-1+oy, pz-1+oz), staged per workgroup (o0/o1: the fine-cell offset
// inside an element of the points at 0.5 -/+ 0.5/sqrt(3) along an axis).
// Dynamic LDS: 4n doubles (one coefficient per element) or 32n (per point);
// direct: per-point coefficients read from kfine instead (levels too wide for
// the staged lines)
__global__ __launch_bounds__(256) void k_q1_lattice_dia(int n, int sub, int o0, int o1, double h,
                                                        const double* __restrict__ tab,
                                                        const double* __restrict__ kfine, double* __restrict__ dia,
                                                        int direct) {
    extern __shared__ double kl[];
    __shared__ double st[576];
    const int W = n + 1, nf = n * sub;
    const bool per_point = kfine && sub > 1;
    const int py = blockIdx.x % W, pz = blockIdx.x / W;
    for (int i = threadIdx.x; i < 576; i += blockDim.x) st[i] = tab[i];
    if (!per_point) {
        stage_element_lines(n, py, pz, kfine, kl);
    } else if (!direct) {
        for (int i = threadIdx.x; i < 32 * n; i += blockDim.x) {
            const int c = i / (8 * n), ex = (i >> 3) % n, q = i & 7;
            const int ey = py - 1 + (c & 1), ez = pz - 1 + (c >> 1);
            double k = 0.0;
            if (ey >= 0 && ey < n && ez >= 0 && ez < n) {
                const int cx = ex * sub + ((q & 1) ? o1 : o0), cy = ey * sub + ((q & 2) ? o1 : o0),
                          cz = ez * sub + ((q & 4) ? o1 : o0);
                k = kfine[((int64_t)cz * nf + cy) * nf + cx];
            }
            kl[i] = k;
        }
    }
    __syncthreads();
    // half-wave per point, lane j < 27 per diagonal
    const int j = threadIdx.x & 31, half = threadIdx.x >> 5;
    if (j >= 27) return;
    const LineLane L(3, W, blockIdx.x, j);
    const int64_t pline = (int64_t)blockIdx.x * W;
    for (int px = half; px < W; px += 8) {
        double v = 0.0;
        if (L.valid(W, px)) {
            const int qx = px + L.dx;
#pragma unroll
            for (int oz = 0; oz < 2; ++oz)
#pragma unroll
                for (int oy = 0; oy < 2; ++oy)
#pragma unroll
                    for (int ox = 0; ox < 2; ++ox) {
                        const int ex = px - 1 + ox, ey = py - 1 + oy, ez = pz - 1 + oz;
                        const int bx = qx - ex, by = L.qy - ey, bz = L.qz - ez;
                        const bool use = ex >= 0 && ex < n && ey >= 0 && ey < n && ez >= 0 && ez < n && bx >= 0 &&
                                         bx <= 1 && by >= 0 && by <= 1 && bz >= 0 && bz <= 1;
                        if (!use) continue;
                        const int ab = 8 * hex_local(1 - ox, 1 - oy, 1 - oz) + hex_local(bx, by, bz);
                        const int c = 2 * oz + oy;
                        if (!per_point) {
                            v = fma(kl[c * n + ex], st[kK + ab], v);
                        } else if (!direct) {
                            const double* kq = kl + c * 8 * n + 8 * ex;
#pragma unroll
                            for (int q = 0; q < 8; ++q) v = fma(kq[q], st[kG + 64 * q + ab], v);
                        } else {
                            for (int q = 0; q < 8; ++q) {
                                const int cx = ex * sub + ((q & 1) ? o1 : o0), cy = ey * sub + ((q & 2) ? o1 : o0),
                                          cz = ez * sub + ((q & 4) ? o1 : o0);
                                v = fma(kfine[((int64_t)cz * nf + cy) * nf + cx], st[kG + 64 * q + ab], v);
                            }
                        }
                    }
            v *= h;
        }
        dia[dia_tiled_index(j, pline + px, 27)] = v;
    }
}

__global__ void k_exp_values(int64_t n, const double* __restrict__ theta, double* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = exp(theta[i]);
}

}  // namespace

void q1_lattice_tables(double* tab) {
    // G_q = w_q grad(phi_a) . grad(phi_b), K = sum_q G_q, F_a = sum_q w_q phi_a
    // on the unit cube, 2x2x2 Gauss points (weights 1/8)
    const double g = 0.5 / std::sqrt(3.0);
    for (int i = 0; i < kQ1Tab; ++i) tab[i] = 0.0;
    for (int q = 0; q < 8; ++q) {
        const double xi[3] = {(q & 1) ? 0.5 + g : 0.5 - g, (q & 2) ? 0.5 + g : 0.5 - g, (q & 4) ? 0.5 + g : 0.5 - g};
        double phi[8], d[8][3];
        for (int a = 0; a < 8; ++a) {
            const int bx = (a == 1 || a == 2 || a == 5 || a == 6), by = (a == 2 || a == 3 || a == 6 || a == 7),
                      bz = (a >= 4);
            const double fx = bx ? xi[0] : 1.0 - xi[0], fy = by ? xi[1] : 1.0 - xi[1], fz = bz ? xi[2] : 1.0 - xi[2];
            const double gx = bx ? 1.0 : -1.0, gy = by ? 1.0 : -1.0, gz = bz ? 1.0 : -1.0;
            phi[a] = fx * fy * fz;
            d[a][0] = gx * fy * fz;
            d[a][1] = fx * gy * fz;
            d[a][2] = fx * fy * gz;
        }
        for (int a = 0; a < 8; ++a) {
            for (int b = 0; b < 8; ++b) {
                const double gq = 0.125 * (d[a][0] * d[b][0] + d[a][1] * d[b][1] + d[a][2] * d[b][2]);
                tab[kG + 64 * q + 8 * a + b] = gq;
                tab[kK + 8 * a + b] += gq;
            }
            tab[kF + a] += 0.125 * phi[a];
        }
    }
}

int rowptr_of_lattice(int64_t npts, const int32_t* dof_of_lat, const int32_t* rowptr, int32_t* rp_lat, hipStream_t s) {
    if (npts <= 0) return 0;
    hipLaunchKernelGGL(k_rowptr_of_lattice, dim3(nblk(npts, 256)), dim3(256), 0, s, npts, dof_of_lat, rowptr, rp_lat);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// the lines of a workgroup are staged in LDS up to this lattice width
constexpr int kStageMaxW = 640;

int lattice_pattern(int dim, int W, const int32_t* dof_of_lat, const int32_t* rp_lat, int32_t* col, hipStream_t s) {
    const int64_t lines = dim == 3 ? (int64_t)W * W : W;
    if (lines <= 0) return 0;
    if (W <= kStageMaxW)
        hipLaunchKernelGGL(k_lattice_pattern<true>, dim3((unsigned)lines), dim3(256), sizeof(int32_t) * 10 * W, s,
                           dim, W, dof_of_lat, rp_lat, col);
    else
        hipLaunchKernelGGL(k_lattice_pattern<false>, dim3((unsigned)lines), dim3(256), 0, s, dim, W, dof_of_lat,
                           rp_lat, col);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int exp_values(int64_t n, const double* theta, double* out, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_exp_values, dim3(nblk(n, 256)), dim3(256), 0, s, n, theta, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int q1_lattice_csr(int n, double h, double f_const, const double* tab, const double* kel, const int32_t* dof_of_lat,
                   const int32_t* rp_lat, double* val, double* rhs, hipStream_t s) {
    const int64_t W = n + 1;
    const size_t lds = sizeof(double) * 4 * (size_t)n;
    if (n <= 0 || lds > 48 * 1024) return -1;
    const double fh3 = f_const * h * h * h;
    if (W <= kStageMaxW)
        hipLaunchKernelGGL(k_q1_lattice_csr<true>, dim3((unsigned)(W * W)), dim3(256),
                           lds + sizeof(int32_t) * 10 * W, s, n, h, fh3, tab, kel, dof_of_lat, rp_lat, val, rhs);
    else
        hipLaunchKernelGGL(k_q1_lattice_csr<false>, dim3((unsigned)(W * W)), dim3(256), lds, s, n, h, fh3, tab, kel,
                           dof_of_lat, rp_lat, val, rhs);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int q1_lattice_dia(int n, int nfine, double h, const double* tab, const double* kfine, double* dia, hipStream_t s) {
    if (n <= 0 || nfine % n != 0) return -1;
    const int64_t W = n + 1;
    const int sub = nfine / n;
    const double g = 0.5 / std::sqrt(3.0);
    const int o0 = (int)std::floor((0.5 - g) * sub), o1 = (int)std::floor((0.5 + g) * sub);
    size_t lds = sizeof(double) * (kfine && sub > 1 ? 32 : 4) * (size_t)n;
    int direct = 0;
    if (lds > 48 * 1024) {
        if (!(kfine && sub > 1) || W * W * W >= ((int64_t)1 << 32)) return -1;
        lds = 0;
        direct = 1;
    }
    hipLaunchKernelGGL(k_q1_lattice_dia, dim3((unsigned)(W * W)), dim3(256), lds, s, n, sub, o0, o1, h, tab, kfine,
                       dia, direct);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugamd
