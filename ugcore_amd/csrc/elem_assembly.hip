// Element-loop assembly from caller-supplied connectivity: the element loop of
// DomainDiscretization::assemble_linear / assemble_jacobian / assemble_defect
// (domain_disc_impl.h:469-948; StdGlobAssembler, elem_disc_assemble_util.h:
// 326-413, 577-687, 902-1003) over one element type, for an element -> DoF
// table such as DoFDistribution::indices gives per element
// (dof_distribution.cpp:1117-1206) and DoF coordinates.
//
// Set-up, once per connectivity, all on the device:
//   * the CSR pattern: every element coupling (row, col) is a 64-bit key
//     row << 32 | col; a radix sort groups equal keys, run heads and a scan
//     number the distinct entries, and the row pointers are read off the sorted
//     keys. Rows come out sorted by column and every coupling is stored, also
//     one whose value is zero, as AddLocalMatrixToGlobal's insertion
//     (local_algebra.h:766-789) leaves them;
//   * the slot of every local entry (e, a, b) in that CSR;
//   * a colouring of the elements such that no two elements of one colour
//     share a DoF (Jones-Plassmann rounds over the DoF -> element incidence,
//     each local maximum of a hashed priority taking the lowest colour its
//     coloured neighbours leave free), and the elements sorted by colour.
// Each assembly then runs one launch per colour, one element per thread: the
// local matrix in registers, added to its slots without atomics.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <cstdint>

#include "elem_local.hpp"
#include "ugx_internal.hpp"

namespace ugamd {

namespace {

inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t > 0 ? (n + t - 1) / t : 1); }

int bits_for(int64_t maxval) {
    int b = 1;
    while (b < 63 && ((int64_t)1 << b) <= maxval) ++b;
    return b;
}

__global__ void k_couple_keys(int64_t ne, int nv, const int32_t* __restrict__ dofs, uint64_t* __restrict__ key,
                              int32_t* __restrict__ idx) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nv2 = (int64_t)nv * nv;
    if (t >= ne * nv2) return;
    const int64_t e = t / nv2;
    const int ab = (int)(t - e * nv2);
    const uint32_t r = (uint32_t)dofs[e * nv + ab / nv], c = (uint32_t)dofs[e * nv + ab % nv];
    key[t] = ((uint64_t)r << 32) | c;
    idx[t] = (int32_t)t;
}

__global__ void k_run_heads(int64_t n, const uint64_t* __restrict__ key, int32_t* __restrict__ head) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) head[k] = (k == 0 || key[k] != key[k - 1]) ? 1 : 0;
}

// run_incl[k]: number of run heads in [0, k]; the entry of sorted position k
// is run_incl[k] - 1
__global__ void k_pattern_cols(int64_t n, const uint64_t* __restrict__ key, const int32_t* __restrict__ head,
                               const int32_t* __restrict__ run_incl, int32_t* __restrict__ col) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n && head[k]) col[run_incl[k] - 1] = (int32_t)(key[k] & 0xffffffffu);
}

// rowptr[r] = first entry of row r: thread k (a run head) fills the rows
// between the previous key's row and its own; thread n closes the rows after
// the last key
__global__ void k_pattern_rowptr(int64_t n, const uint64_t* __restrict__ key, const int32_t* __restrict__ head,
                                 const int32_t* __restrict__ run_incl, int64_t nrows, int32_t* __restrict__ rowptr) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k > n) return;
    if (k < n && !head[k]) return;
    const int64_t lo = k == 0 ? -1 : (int64_t)(key[k - 1] >> 32);
    const int64_t hi = k == n ? nrows : (int64_t)(key[k] >> 32);
    const int32_t at = k == n ? (n > 0 ? run_incl[n - 1] : 0) : run_incl[k] - 1;
    for (int64_t r = lo + 1; r <= hi && r <= nrows; ++r) rowptr[r] = at;
}

__global__ void k_slot_of_local(int64_t n, const int32_t* __restrict__ idx_sorted, const int32_t* __restrict__ run_incl,
                                int32_t* __restrict__ slot_local) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) slot_local[idx_sorted[k]] = run_incl[k] - 1;
}

// incidence keys: the DoF of local vertex t = e*nv + a
__global__ void k_incidence_keys(int64_t n, const int32_t* __restrict__ dofs, uint32_t* __restrict__ key,
                                 int32_t* __restrict__ idx) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n) return;
    key[t] = (uint32_t)dofs[t];
    idx[t] = (int32_t)t;
}

// ptr[v] = first sorted position whose key is >= v, v = 0..nkeys
__global__ void k_ptr_from_sorted(int64_t n, const uint32_t* __restrict__ key, int64_t nkeys,
                                  int32_t* __restrict__ ptr) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k > n) return;
    const int64_t lo = k == 0 ? -1 : (int64_t)key[k - 1];
    const int64_t hi = k == n ? nkeys : (int64_t)key[k];
    for (int64_t v = lo + 1; v <= hi && v <= nkeys; ++v) ptr[v] = (int32_t)k;
}

__device__ __forceinline__ uint64_t prio_hash(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// one Jones-Plassmann round: an uncoloured element whose priority beats every
// uncoloured neighbour takes the lowest colour no coloured neighbour holds.
// Two elements coloured in the same round are never neighbours, so the
// colours read here were all written in earlier rounds.
__global__ void k_colour_round(int64_t ne, int nv, const int32_t* __restrict__ dofs,
                               const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc_t,
                               int32_t* __restrict__ colour, int* __restrict__ coloured, int* __restrict__ overflow) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= ne || colour[e] >= 0) return;
    const uint64_t pe = prio_hash((uint64_t)e);
    uint64_t used = 0;
    for (int a = 0; a < nv; ++a) {
        const int32_t v = dofs[e * nv + a];
        for (int32_t k = inc_ptr[v]; k < inc_ptr[v + 1]; ++k) {
            const int64_t f = inc_t[k] / nv;
            if (f == e) continue;
            const int32_t cf = colour[f];
            if (cf < 0) {
                const uint64_t pf = prio_hash((uint64_t)f);
                if (pf > pe || (pf == pe && f > e)) return;
            } else {
                used |= (uint64_t)1 << cf;
            }
        }
    }
    if (used == ~(uint64_t)0) {
        *overflow = 1;
        return;
    }
    colour[e] = __ffsll((long long)~used) - 1;
    atomicAdd(coloured, 1);
}

__global__ void k_colour_keys(int64_t ne, const int32_t* __restrict__ colour, uint32_t* __restrict__ key,
                              int32_t* __restrict__ idx) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= ne) return;
    key[e] = (uint32_t)colour[e];
    idx[e] = (int32_t)e;
}

// slot[ab * ne + pos] = slot of local entry ab of the element at colour-sorted position pos
__global__ void k_slots_by_position(int64_t ne, int nv2, const int32_t* __restrict__ order,
                                    const int32_t* __restrict__ slot_local, int32_t* __restrict__ slot) {
    const int64_t pos = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (pos >= ne) return;
    const int64_t e = order[pos];
    for (int ab = 0; ab < nv2; ++ab) slot[(int64_t)ab * ne + pos] = slot_local[e * nv2 + ab];
}

struct ElemArgs {
    int64_t ne;
    const int32_t* order;
    const int32_t* dofs;
    const double* coords;
    const double* kel;      // diffusion per element (null: 1)
    double f_const;
    const double* f_dof;    // source per DoF (null: f_const)
    double vel[3];
    int conv;
    const int32_t* slot;
    double* val;            // matrix values (null: no matrix)
    double* rhs;            // right-hand side (null: none)
    const double* u;        // defect mode: d += A_e u_e - f_e
    double* d;
};

template <int NV, int DIM>
__device__ __forceinline__ void load_element(const ElemArgs& A, int64_t e, int32_t* dof, double (*X)[DIM]) {
    for (int a = 0; a < NV; ++a) {
        dof[a] = A.dofs[e * NV + a];
        for (int c = 0; c < DIM; ++c) X[a][c] = A.coords[(int64_t)dof[a] * DIM + c];
    }
}

template <int NV>
__device__ __forceinline__ void scatter(const ElemArgs& A, int64_t pos, const int32_t* dof, const double (*Ke)[NV],
                                        const double* fe) {
    if (A.val) {
        for (int a = 0; a < NV; ++a)
            for (int b = 0; b < NV; ++b) {
                const int32_t k = A.slot[(int64_t)(a * NV + b) * A.ne + pos];
                A.val[k] += Ke[a][b];
            }
    }
    if (A.rhs)
        for (int a = 0; a < NV; ++a) A.rhs[dof[a]] += fe[a];
    if (A.d) {
        for (int a = 0; a < NV; ++a) {
            double s = 0.0;
            for (int b = 0; b < NV; ++b) s += Ke[a][b] * A.u[dof[b]];
            A.d[dof[a]] += s - fe[a];
        }
    }
}

// P1 triangle, FV1; source per scv: f * area / 3 (f per DoF or constant)
__global__ __launch_bounds__(128) void k_elem_tri_fv1(ElemArgs A, int64_t pos0, int64_t cnt) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int64_t pos = pos0 + t;
    const int64_t e = A.order[pos];
    int32_t dof[3];
    double X[3][2];
    load_element<3, 2>(A, e, dof, X);
    const double k = A.kel ? A.kel[e] : 1.0;
    double Ke[3][3], area, fe[3];
    fv1_tri_element_k(X, [=](const double*) { return k; }, Ke, area);
    for (int a = 0; a < 3; ++a) fe[a] = (A.f_dof ? A.f_dof[dof[a]] : A.f_const) * (area / 3.0);
    scatter<3>(A, pos, dof, Ke, fe);
}

// P1 tetrahedron, FV1 with optional full-upwind convection; source f * |scv|
__global__ __launch_bounds__(128) void k_elem_tet_fv1(ElemArgs A, int64_t pos0, int64_t cnt) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int64_t pos = pos0 + t;
    const int64_t e = A.order[pos];
    int32_t dof[4];
    double X[4][3];
    load_element<4, 3>(A, e, dof, X);
    const double k = A.kel ? A.kel[e] : 1.0;
    const double kip[6] = {k, k, k, k, k, k};
    double Ke[4][4], vol[4], fe[4];
    fv1_tet_element(X, kip, A.vel, A.conv != 0, Ke, vol);
    for (int a = 0; a < 4; ++a) fe[a] = (A.f_dof ? A.f_dof[dof[a]] : A.f_const) * vol[a];
    scatter<4>(A, pos, dof, Ke, fe);
}

// Q1 hexahedron, FE (2x2x2 Gauss); source f_const * int phi
__global__ __launch_bounds__(128) void k_elem_hex_q1(ElemArgs A, int64_t pos0, int64_t cnt) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int64_t pos = pos0 + t;
    const int64_t e = A.order[pos];
    int32_t dof[8];
    double X[8][3];
    load_element<8, 3>(A, e, dof, X);
    const double k = A.kel ? A.kel[e] : 1.0;
    double Ke[8][8], fe[8];
    q1_hex_element_k(X, [=](const double*) { return k; }, A.f_const, Ke, fe);
    scatter<8>(A, pos, dof, Ke, fe);
}

__global__ void k_zero_listed(int64_t n, const int32_t* __restrict__ idx, double* __restrict__ x) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < n) x[idx[t]] = 0.0;
}

// ---- connectivity of the lattice domains (the top level of a hierarchy) ----

// kind 0: hexahedra (ex, ey, ez) x fastest, corners in the reference order
__global__ void k_lattice_hex_dofs(int n, const int32_t* __restrict__ dof_of_lat, const double* __restrict__ theta,
                                   int32_t* __restrict__ dofs, double* __restrict__ kel) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t ne = (int64_t)n * n * n;
    if (e >= ne) return;
    const int ex = (int)(e % n), ey = (int)((e / n) % n), ez = (int)(e / ((int64_t)n * n));
    const int64_t W = n + 1;
    for (int a = 0; a < 8; ++a) {
        const int dx = (a == 1 || a == 2 || a == 5 || a == 6), dy = (a == 2 || a == 3 || a == 6 || a == 7), dz = a >= 4;
        dofs[e * 8 + a] = dof_of_lat[((ez + dz) * W + (ey + dy)) * W + (ex + dx)];
    }
    kel[e] = theta ? exp(theta[e]) : 1.0;
}

// kind 1: squares (sx, sy) split along (1,1): lower (0,0),(1,0),(1,1) and
// upper (0,0),(1,1),(0,1); element 2*cell + upper
__global__ void k_lattice_tri_dofs(int n, const int32_t* __restrict__ dof_of_lat, const double* __restrict__ theta,
                                   int32_t* __restrict__ dofs, double* __restrict__ kel) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= 2 * (int64_t)n * n) return;
    const int64_t cell = e >> 1;
    const int upper = (int)(e & 1);
    const int sx = (int)(cell % n), sy = (int)(cell / n);
    const int cx[2][3] = {{0, 1, 1}, {0, 1, 0}}, cy[2][3] = {{0, 0, 1}, {0, 1, 1}};
    const int64_t W = n + 1;
    for (int a = 0; a < 3; ++a) dofs[e * 3 + a] = dof_of_lat[(sy + cy[upper][a]) * W + sx + cx[upper][a]];
    kel[e] = theta ? exp(theta[cell]) : 1.0;
}

// kind 2: the 6 tetrahedra per cell from their packed corner codes
__global__ void k_lattice_tet_dofs(int n, const uint16_t* __restrict__ cell_tets, const int32_t* __restrict__ dof_of_lat,
                                   const double* __restrict__ theta, int32_t* __restrict__ dofs,
                                   double* __restrict__ kel) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= 6 * (int64_t)n * n * n) return;
    const int64_t cell = e / 6;
    const int ci = (int)(cell % n), cj = (int)((cell / n) % n), ck = (int)(cell / ((int64_t)n * n));
    const uint16_t code = cell_tets[e];
    const int64_t W = n + 1;
    for (int q = 0; q < 4; ++q) {
        const int c = (code >> (3 * q)) & 7;
        dofs[e * 4 + q] = dof_of_lat[((ck + ((c >> 2) & 1)) * W + (cj + ((c >> 1) & 1))) * W + ci + (c & 1)];
    }
    kel[e] = theta ? exp(theta[cell]) : 1.0;
}

struct WidenI32 {
    __host__ __device__ int64_t operator()(int32_t v) const { return v; }
};

__global__ void k_rowptr_from_incl(int64_t n, const int64_t* __restrict__ incl, int32_t* __restrict__ rowptr) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i > n) return;
    rowptr[i] = i == 0 ? 0 : (int32_t)incl[i - 1];
}

}  // namespace

// ---- row pointers from row counts (scanned in 64 bits) -------------------------

int rowptr_scan_temp_bytes(int64_t n, size_t* bytes) {
    *bytes = 0;
    if (n <= 0) return 0;
    auto in = rocprim::make_transform_iterator((const int32_t*)nullptr, WidenI32());
    return rocprim::inclusive_scan(nullptr, *bytes, in, (int64_t*)nullptr, (size_t)n, rocprim::plus<int64_t>()) ==
                   hipSuccess
               ? 0
               : -1;
}

int rowptr_from_counts(int64_t n, const int32_t* cnt, int64_t* incl, void* temp, size_t bytes, int32_t* rowptr,
                       hipStream_t s) {
    if (n > 0) {
        auto in = rocprim::make_transform_iterator(cnt, WidenI32());
        size_t b = bytes;
        if (rocprim::inclusive_scan(temp, b, in, incl, (size_t)n, rocprim::plus<int64_t>(), s) != hipSuccess)
            return -1;
    }
    hipLaunchKernelGGL(k_rowptr_from_incl, dim3(blocks(n + 1, 256)), dim3(256), 0, s, n, incl, rowptr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---- pattern -------------------------------------------------------------------

int elem_pattern_temp_bytes(int64_t ne, int nv, int64_t ndofs, size_t* sort_bytes, size_t* scan_bytes) {
    const int64_t n = ne * nv * nv;
    *sort_bytes = *scan_bytes = 0;
    if (n == 0) return 0;
    const int bits = 32 + bits_for(ndofs);
    if (rocprim::radix_sort_pairs(nullptr, *sort_bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                  (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)n, 0, bits) != hipSuccess)
        return -1;
    if (rocprim::inclusive_scan(nullptr, *scan_bytes, (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)n,
                                rocprim::plus<int32_t>()) != hipSuccess)
        return -1;
    return 0;
}

int elem_pattern(int64_t ne, int nv, int64_t ndofs, const int32_t* dofs, uint64_t* key, uint64_t* key2, int32_t* idx,
                 int32_t* idx2, int32_t* head, int32_t* run_incl, void* temp, size_t sort_bytes, size_t scan_bytes,
                 int32_t* rowptr, int32_t* nnz_dev, hipStream_t s) {
    const int64_t n = ne * nv * nv;
    if (n > 0) {
        hipLaunchKernelGGL(k_couple_keys, dim3(blocks(n, 256)), dim3(256), 0, s, ne, nv, dofs, key, idx);
        size_t sb = sort_bytes;
        if (rocprim::radix_sort_pairs(temp, sb, key, key2, idx, idx2, (size_t)n, 0, 32 + bits_for(ndofs), s) !=
            hipSuccess)
            return -1;
        hipLaunchKernelGGL(k_run_heads, dim3(blocks(n, 256)), dim3(256), 0, s, n, key2, head);
        size_t cb = scan_bytes;
        if (rocprim::inclusive_scan(temp, cb, head, run_incl, (size_t)n, rocprim::plus<int32_t>(), s) != hipSuccess)
            return -1;
        if (hipMemcpyAsync(nnz_dev, run_incl + (n - 1), sizeof(int32_t), hipMemcpyDeviceToDevice, s) != hipSuccess)
            return -1;
    } else if (hipMemsetAsync(nnz_dev, 0, sizeof(int32_t), s) != hipSuccess) {
        return -1;
    }
    hipLaunchKernelGGL(k_pattern_rowptr, dim3(blocks(n + 1, 256)), dim3(256), 0, s, n, key2, head, run_incl, ndofs,
                       rowptr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int elem_pattern_fill(int64_t ne, int nv, const uint64_t* key2, const int32_t* idx2, const int32_t* head,
                      const int32_t* run_incl, int32_t* col, int32_t* slot_local, hipStream_t s) {
    const int64_t n = ne * nv * nv;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_pattern_cols, dim3(blocks(n, 256)), dim3(256), 0, s, n, key2, head, run_incl, col);
    hipLaunchKernelGGL(k_slot_of_local, dim3(blocks(n, 256)), dim3(256), 0, s, n, idx2, run_incl, slot_local);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---- incidence and colouring ---------------------------------------------------------

int elem_incidence_temp_bytes(int64_t n, int64_t maxkey, size_t* bytes) {
    *bytes = 0;
    if (n == 0) return 0;
    return rocprim::radix_sort_pairs(nullptr, *bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)n, 0, bits_for(maxkey)) ==
                   hipSuccess
               ? 0
               : -1;
}

int elem_incidence(int64_t ne, int nv, int64_t ndofs, const int32_t* dofs, uint32_t* key, uint32_t* key2, int32_t* idx,
                   int32_t* inc_t, void* temp, size_t temp_bytes, int32_t* inc_ptr, hipStream_t s) {
    const int64_t n = ne * nv;
    if (n > 0) {
        hipLaunchKernelGGL(k_incidence_keys, dim3(blocks(n, 256)), dim3(256), 0, s, n, dofs, key, idx);
        size_t tb = temp_bytes;
        if (rocprim::radix_sort_pairs(temp, tb, key, key2, idx, inc_t, (size_t)n, 0, bits_for(ndofs), s) !=
            hipSuccess)
            return -1;
    }
    hipLaunchKernelGGL(k_ptr_from_sorted, dim3(blocks(n + 1, 256)), dim3(256), 0, s, n, key2, ndofs, inc_ptr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int elem_colour_round(int64_t ne, int nv, const int32_t* dofs, const int32_t* inc_ptr, const int32_t* inc_t,
                      int32_t* colour, int* coloured, int* overflow, hipStream_t s) {
    if (ne == 0) return 0;
    hipLaunchKernelGGL(k_colour_round, dim3(blocks(ne, 256)), dim3(256), 0, s, ne, nv, dofs, inc_ptr, inc_t, colour,
                       coloured, overflow);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int elem_colour_order(int64_t ne, int ncolours, const int32_t* colour, uint32_t* key, uint32_t* key2, int32_t* idx,
                      int32_t* order, void* temp, size_t temp_bytes, int32_t* colour_ptr, hipStream_t s) {
    if (ne > 0) {
        hipLaunchKernelGGL(k_colour_keys, dim3(blocks(ne, 256)), dim3(256), 0, s, ne, colour, key, idx);
        size_t tb = temp_bytes;
        if (rocprim::radix_sort_pairs(temp, tb, key, key2, idx, order, (size_t)ne, 0, bits_for(ncolours), s) !=
            hipSuccess)
            return -1;
    }
    hipLaunchKernelGGL(k_ptr_from_sorted, dim3(blocks(ne + 1, 256)), dim3(256), 0, s, ne, key2, (int64_t)ncolours,
                       colour_ptr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int elem_slots_by_position(int64_t ne, int nv, const int32_t* order, const int32_t* slot_local, int32_t* slot,
                           hipStream_t s) {
    if (ne == 0) return 0;
    hipLaunchKernelGGL(k_slots_by_position, dim3(blocks(ne, 256)), dim3(256), 0, s, ne, nv * nv, order, slot_local,
                       slot);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---- assembly ---------------------------------------------------------------------

int elem_assemble_colour(int elem_type, const ElemLaunch& L, int64_t pos0, int64_t cnt, hipStream_t s) {
    if (cnt <= 0) return 0;
    ElemArgs A;
    A.ne = L.ne;
    A.order = L.order;
    A.dofs = L.dofs;
    A.coords = L.coords;
    A.kel = L.kel;
    A.f_const = L.f_const;
    A.f_dof = L.f_dof;
    for (int c = 0; c < 3; ++c) A.vel[c] = L.vel[c];
    A.conv = L.conv;
    A.slot = L.slot;
    A.val = L.val;
    A.rhs = L.rhs;
    A.u = L.u;
    A.d = L.d;
    if (elem_type == 0)
        hipLaunchKernelGGL(k_elem_tri_fv1, dim3(blocks(cnt, 128)), dim3(128), 0, s, A, pos0, cnt);
    else if (elem_type == 1)
        hipLaunchKernelGGL(k_elem_tet_fv1, dim3(blocks(cnt, 128)), dim3(128), 0, s, A, pos0, cnt);
    else if (elem_type == 2)
        hipLaunchKernelGGL(k_elem_hex_q1, dim3(blocks(cnt, 128)), dim3(128), 0, s, A, pos0, cnt);
    else
        return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int zero_listed(int64_t n, const int32_t* idx, double* x, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_zero_listed, dim3(blocks(n, 256)), dim3(256), 0, s, n, idx, x);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int lattice_connectivity(int kind, int n, const int32_t* dof_of_lat, const uint16_t* cell_tets, const double* theta,
                         int32_t* dofs, double* kel, hipStream_t s) {
    if (kind == 0) {
        const int64_t ne = (int64_t)n * n * n;
        hipLaunchKernelGGL(k_lattice_hex_dofs, dim3(blocks(ne, 256)), dim3(256), 0, s, n, dof_of_lat, theta, dofs, kel);
    } else if (kind == 1) {
        const int64_t ne = 2 * (int64_t)n * n;
        hipLaunchKernelGGL(k_lattice_tri_dofs, dim3(blocks(ne, 256)), dim3(256), 0, s, n, dof_of_lat, theta, dofs, kel);
    } else if (kind == 2) {
        const int64_t ne = 6 * (int64_t)n * n * n;
        hipLaunchKernelGGL(k_lattice_tet_dofs, dim3(blocks(ne, 256)), dim3(256), 0, s, n, cell_tets, dof_of_lat, theta,
                           dofs, kel);
    } else {
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugamd
