// Internal launcher declarations (kernels.hip). All are stream ordered, return
// 0 on success and -1 when the launch failed; none allocates device memory.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ugamd {

template <typename TM, typename TV>
int spmv_csr(int64_t nrows, const int32_t* rowptr, const int32_t* col, const TM* val, double alpha, const TV* v1,
             double beta, const TV* w, TV* dest, int mode, int lanes, hipStream_t s);

// sliced ELL (64-row slices) SpMV layout; order (nullable) is the row
// permutation that fills the slices
int sell_widths(int64_t nrows, const int32_t* rowptr, const int32_t* order, int32_t* width, hipStream_t s);
template <typename TM>
int sell_fill(int64_t nrows, int64_t ncols, const int32_t* rowptr, const int32_t* col, const TM* val,
              const int32_t* order, const int64_t* slice_ptr, const int32_t* width, int32_t* scol, TM* sval,
              hipStream_t s, bool vals_only = false);
template <typename TM, typename TV>
int sell_spmv(int64_t nrows, const int64_t* slice_ptr, const int32_t* width, const int32_t* scol, const TM* sval,
              double alpha, const TV* v1, double beta, const TV* w, TV* dest, int mode, const int32_t* rowptr,
              const int32_t* order, hipStream_t s);

template <typename T>
int vec_set(int64_t n, T* x, double v, hipStream_t s);
template <typename T>
int vec_scale_add(int64_t n, T* dest, double a1, const T* v1, double a2, const T* v2, double a3, const T* v3,
                  hipStream_t s);
template <typename TD, typename TV>
int diag_mult(int64_t n, TV* dest, const TD* d, const TV* x, hipStream_t s);
template <typename TM, typename TD>
int jacobi_diag_csr(int64_t n, const int32_t* rowptr, const int32_t* col, const TM* val, double omega, TD* dinv,
                    int* bad, hipStream_t s);
template <typename T>
int dot(int64_t n, const T* a, const T* b, double* partial, double* out, bool take_sqrt, hipStream_t s);
// out[0] = number of nonzero entries of a[0, n)
int count_nonzero(int64_t n, const double* a, double* partial, double* out, hipStream_t s);
// number of partial sums dot() writes for n entries (one per workgroup of its
// grid, at most 65536)
int dot_partials(int64_t n);
// CG with device scalars sc[0..6] = rho, lambda, alpha, beta, rho_new, |r|^2,
// breakdown flag
int sum_partials(int np, const double* partial, double* out, hipStream_t s);
int cg_beta(double* sc, hipStream_t s);
// alpha = sc[0] / sc[1] into sc[2] (0 and sc[6] = 1 when sc[1] == 0), then the
// update; rf (nullable): also writes (float) r, the FP32 cycle's copy of the defect
int cg_update(int64_t n, double* sc, const double* p, const double* q, double* x, double* r, double* partial,
              hipStream_t s,
              float* rf = nullptr);
int cg_direction(int64_t n, const double* sc, double* p, const double* z, hipStream_t s);
template <typename T>
int convert_dot(int64_t n, const T* c, double* z, const double* r, double* partial, double* out, hipStream_t s);
template <typename TS, typename TL>
int permute_to(int64_t n, const int32_t* dof_of_lat, const TS* src, TL* dst, hipStream_t s);
// dst = src permuted into lattice order, dst32 = (float) dst
int permute_to_dual(int64_t n, const int32_t* dof_of_lat, const double* src, double* dst, float* dst32, hipStream_t s);
template <typename TL, typename TS>
int permute_add_from(int64_t n, const int32_t* lat_of_dof, const TL* src, TS* dst, hipStream_t s);

template <typename TM, typename TV>
int dia_apply_sub(int64_t n, int nd, const int32_t* off, const TM* A, const TV* x, const TV* d_in, TV* d_out,
                  hipStream_t s);
template <typename TM, typename TV>
int dia_axpy(int64_t n, int nd, const int32_t* off, const TM* A, double alpha, const TV* v, double beta, const TV* x,
             TV* y, hipStream_t s);
template <typename TM, typename TV>
int dia_jacobi_step(int64_t n, int nd, const int32_t* off, const TM* A, const TV* dinv, const TV* sd_in, TV* sd_out,
                    TV* sc, hipStream_t s);
// residency queries of the matrix-free kernels at lattice width W, made
// before any cycle is captured into a graph
int q1mf_prepare(int W);
// the first Jacobi step of a cycle from the FP64 defect d64 and a zero
// correction (sd_out = (T)d64 - A st, sc = st, st = dinv .* (T)d64), when
// q1mf_first_fits(W, sizeof(T)) (the z-marching plan applies)
bool q1mf_first_fits(int W, size_t elem);
template <typename T>
int q1mf_jacobi_first(int W, const T* cell, const T* dinv, const double* d64, T* sd_out, T* sc, hipStream_t s);
// the CG direction update fused with the operator: p_new = fma(*beta_dev,
// p_old, z) (as cg_direction), y = A p_new, (y, p_new) partials; p_new and
// p_old distinct, both with the halo
template <typename T>
int q1mf_direction_axpy(int W, const T* cell, const T* p_old, const T* z, const double* beta_dev, T* p_new, T* y,
                        hipStream_t s, double* dot_partial, int* n_partial);
// the same from a defect already in T (with the halo)
template <typename T>
int q1mf_jacobi_first_t(int W, const T* cell, const T* dinv, const T* sd_in, T* sd_out, T* sc, hipStream_t s);
template <typename TV>
int jacobi_accumulate(int64_t n, const TV* dinv, const TV* sd, TV* sc, hipStream_t s);
// z = (double)(sc + dinv .* sd) (arithmetic of jacobi_accumulate); with out,
// also (z, r) into *out through partial (as convert_dot)
template <typename TV>
int jacobi_convert_dot(int64_t n, const TV* dinv, const TV* sd, const TV* sc, double* z, const double* r,
                       double* partial, double* out, hipStream_t s);
// symmetric (Dirichlet-eliminated) storage: nd/2+1 tiled upper diagonals
template <typename TM>
int dia_symmetrize(int64_t n, int nd, const int32_t* off, const uint8_t* dir, const TM* full, TM* up, hipStream_t s);
template <typename TM, typename TV>
int dias_apply_sub(int64_t n, int nd, const int32_t* off, const TM* U, const TV* x, const TV* d_in, TV* d_out,
                   hipStream_t s);
template <typename TM, typename TV>
int dias_axpy(int64_t n, int nd, const int32_t* off, const TM* U, double alpha, const TV* v, double beta, const TV* x,
              TV* y, hipStream_t s);
// dst_dof[i] = src_lat[lat_of_dof[i]]
template <typename TL, typename TS>
int permute_from(int64_t n, const int32_t* lat_of_dof, const TL* src, TS* dst, hipStream_t s);
// dst_dof[dof_of_lat[l]] = src_lat[l]: the same copy when dof_of_lat is a
// bijection of [0, n) (every DoF numbers exactly one lattice point)
template <typename TL, typename TS>
int permute_scatter(int64_t n, const int32_t* dof_of_lat, const TL* src, TS* dst, hipStream_t s);
template <typename TM, typename TV>
int dias_jacobi_step(int64_t n, int nd, const int32_t* off, const TM* U, const TV* dinv, const TV* sd_in, TV* sd_out,
                     TV* sc, hipStream_t s);
// matrix-free Q1 operator on a W^3 lattice with per-element coefficients
// c_e = h k_e ((W-1)^3, element lattice order); vectors carry the lattice halo
template <typename T>
int q1mf_jacobi_step(int W, const T* cell, const T* dinv, const T* sd_in, T* sd_out, T* sc, hipStream_t s);
template <typename T>
int q1mf_apply_sub(int W, const T* cell, const T* x, const T* d_in, T* d_out, hipStream_t s);
// the last post-smoothing step fused with the final correction update and
// its conversion: z = (double)(sc + st + dinv .* (sd_in - A st)), st =
// dinv .* sd_in (q1mf_jacobi_step then jacobi_convert_dot, without storing the
// step's outputs); with dot_partial also the (z, r) partials (n_partial in:
// capacity, out: number written). q1mf_last_fits(W, sizeof(T)) tells whether
// the plan applies.
template <typename T>
int q1mf_jacobi_last(int W, const T* cell, const T* dinv, const T* sd_in, const T* sc, double* z, const double* r,
                     hipStream_t s, double* dot_partial, int* n_partial);
bool q1mf_last_fits(int W, size_t elem);
// dot_partial (optional): per-block partial sums of y.x (n_partial in: capacity,
// out: number written); needs the z-marching kernel
template <typename T>
int q1mf_axpy(int W, const T* cell, double alpha, const T* v, double beta, const T* x, T* y, hipStream_t s,
              double* dot_partial = nullptr, int* n_partial = nullptr);
template <typename T>
int q1mf_coeff(int64_t ne, double h, const double* theta, T* c, hipStream_t s);
template <typename TM, typename TV>
int dia_jacobi_diag(int64_t n, int nd, const TM* dia, double omega, TV* dinv, int* bad, hipStream_t s);

// tiled diagonal storage of level operators: the 64 values of diagonal d for
// rows 64t..64t+63 are contiguous and the nd diagonals of one tile follow
// each other
__host__ __device__ inline int64_t dia_tiled_index(int d, int64_t i, int nd) {
    return ((i >> 6) * nd + d) * 64 + (i & 63);
}
inline int64_t dia_tiled_size(int64_t n, int nd) { return ((n + 63) / 64) * 64 * nd; }
// P1 transfers between lattice levels. In 3D with Wf = 2 Wc - 1 the Dirichlet
// flags must mark exactly the lattice boundary (as the multigrid levels' flags
// do): the tiled kernels take the Dirichlet rows from the lattice geometry.
// prolong_add with dinv_c and sd_c reads the coarse correction as
// sc + dinv_c .* sd_c (the coarse level's last post-smoothing update, with the
// arithmetic of jacobi_accumulate, which then need not run); restrict_defect
// with zero_c also sets zero_c (the coarse correction) to 0
template <typename TV>
int prolong_add(int dim, int Wc, int Wf, int64_t n_f, const TV* scc, const uint8_t* dir_f, TV* st, TV* sc,
                hipStream_t s, const TV* dinv_c = nullptr, const TV* sd_c = nullptr);
template <typename TV>
int restrict_defect(int dim, int Wc, int Wf, int64_t n_c, const TV* sdf, const uint8_t* dir_c, const uint8_t* dir_f,
                    TV* sdc, hipStream_t s, TV* zero_c = nullptr);
// x = U^{-1} L^{-1} P b with row-major triangular inverses; perm[i] = row of b
template <typename TV>
int lu_solve(int n, const double* LinvT, const double* UinvT, const int32_t* perm, const TV* b, TV* x, hipStream_t s);

int invert_map(int64_t n, const int32_t* dof_of_lat, int32_t* lat_of_dof, hipStream_t s);
int pattern_count(int dim, int W, int64_t n, const int32_t* lat_of_dof, int32_t* cnt, hipStream_t s);
template <typename TM>
int assemble_fv1_tri(int n, double h, const double* theta, int nfine, const double* f_at_dof, const int32_t* dof_of_lat,
                     const int32_t* rowptr, const int32_t* col, TM* csr_val, TM* dia, int64_t dia_n, double* rhs,
                     hipStream_t s);
template <typename TM>
int dirichlet_csr(int64_t n, int64_t n_inner, const int32_t* rowptr, const int32_t* col, TM* val, double* rhs,
                  hipStream_t s);
template <typename TM>
int dirichlet_dia(int64_t n, int nd, int centre, const uint8_t* dir, TM* dia, hipStream_t s);
template <typename TA, typename TB>
int convert(int64_t n, const TA* a, TB* b, hipStream_t s);
int boundary_flags(int dim, int W, int64_t n, uint8_t* dir, hipStream_t s);
int adjoint_grad_q1_hex(int n, double h, const double* theta, const int32_t* dof_of_lat, int64_t n_inner,
                        const double* u, const double* lambda, double* grad, hipStream_t s);
int zero_range(int64_t lo, int64_t hi, double* x, hipStream_t s);

// ---- Q2 hexahedron element matrices (q2_kernels.hip) ----
// X [n][8][3] corners in the reference order, k [n] (or null: 1), Ke
// [n][27][27] (row-major), fe [n][27] (or null); device pointers
int q2_hex_elements(int64_t n, const double* X, const double* k, double f_const, double* Ke, double* fe,
                    hipStream_t s);
// Q2 on the unit-cube lattice of n^3 elements, nodes on the (2n+1)^3 lattice
// (x fastest): row lengths, sorted columns, colour-ordered element scatter
// (k = exp(theta) per element, theta null: 1; f constant) and Dirichlet rows
int q2_lattice_count(int n, int32_t* cnt, hipStream_t s);
int q2_lattice_fill(int n, const int32_t* rowptr, int32_t* col, hipStream_t s);
int q2_lattice_assemble(int n, const double* theta, double f_const, const int32_t* rowptr, const int32_t* col,
                        double* val, double* rhs, hipStream_t s);

// ---- tetrahedral hierarchy (tet_kernels.hip) ----
struct TetDisc {
    int n = 0;                    // cells per side of the level
    int shift = 0;                // fine lattice units per level lattice unit: 1 << shift
    int n_fine = 0;               // fine cells per side
    double scale = 0.0;           // coordinate per fine lattice unit (1 / n_fine)
    const double* theta = nullptr;   // per fine cell (lattice order); null: k = 1
    uint64_t jseed = 0;
    double jamp = 0.0;            // jitter amplitude in units of the fine h
    int jitter = 0;               // jitter the interior vertices (top level only)
    double vel[3] = {0.0, 0.0, 0.0};
    double f = 0.0;               // constant source
};
int tet_mask(int n, const uint16_t* cell_tets, uint32_t* mask, hipStream_t s);
// DoF coordinates of a level in DoF order (dim per DoF); kind 2 uses the
// (jittered) tetrahedral vertex positions of P
int dof_positions(const TetDisc& P, int kind, int dim, int W, int64_t N, const int32_t* lat_of_dof, double* out,
                  hipStream_t s);
int pattern_count_mask(int64_t N, const int32_t* lat_of_dof, const uint32_t* mask, int32_t* cnt, hipStream_t s);
int pattern_fill_mask(int W, int64_t N, const int32_t* lat_of_dof, const int32_t* dof_of_lat, const uint32_t* mask,
                      const int32_t* rowptr, int32_t* col, hipStream_t s);
template <typename TM>
int assemble_fv1_tet(const TetDisc& P, const uint16_t* cell_tets, const int32_t* dof_of_lat, const int32_t* rowptr,
                     const int32_t* col, TM* csr_val, TM* dia, double* rhs, hipStream_t s);
int adjoint_grad_fv1_tet(const TetDisc& P, const uint16_t* cell_tets, const int32_t* dof_of_lat, int64_t n_inner,
                         const double* u, const double* lambda, double* grad, hipStream_t s);
template <typename TV>
int prolong_add_tet(int Wc, int Wf, int64_t n_f, const TV* scc, const uint8_t* pcode, const uint8_t* dir_f, TV* st,
                    TV* sc, hipStream_t s);
template <typename TV>
int restrict_tet(int Wc, int Wf, int64_t n_c, const TV* sdf, const uint8_t* pcode, const uint8_t* dir_c,
                 const uint8_t* dir_f, TV* sdc, hipStream_t s);
// Galerkin coarse operator R A_f P of the P1 lattice transfer (dim 3: 27
// diagonals, dim 2: 9), tiled DIA storage, FP64; Wf = 2 Wc - 1
int rap_dia(int dim, int Wf, int Wc, const uint8_t* dir_f, const uint8_t* dir_c, const double* Af, double* Ac,
            hipStream_t s);
template <typename T>
int dia_transpose_elim(int W, int64_t n, const uint8_t* dir, const T* A, T* AT, hipStream_t s);
int csr_transpose_map(int64_t n, const int32_t* rowptr, const int32_t* col, int32_t* tp, int* bad, hipStream_t s);
int csr_adjoint_values(int64_t n, int64_t n_inner, const int32_t* rowptr, const int32_t* col, const int32_t* tp,
                       const double* val, double* valT, hipStream_t s);

// ---- SparseMatrix surface (matrix_kernels.hip) ----
// temporary bytes of the transpose's radix sort for nnz entries
int transpose_sort_bytes(int64_t nnz, int32_t ncols, size_t* bytes);
// stable sort of (column, entry) pairs; drop_zeros sends zero entries past the
// last column; rowptr_t (ncols + 1) is read off the sorted keys (key2), the
// sorted entry indices are in idx2; key, key2, idx, idx2, row_of hold nnz
int transpose_csr(int64_t nrows, int32_t ncols, int64_t nnz, const int32_t* rowptr, const int32_t* col,
                  const double* val, int drop_zeros, uint32_t* key, uint32_t* key2, int32_t* idx, int32_t* idx2,
                  int32_t* row_of, void* temp, size_t temp_bytes, int32_t* rowptr_t, hipStream_t s);
// col_t[k] = source row, val_t[k] = scale * value of sorted entry k < n
int transpose_gather(int64_t n, const int32_t* idx_sorted, const int32_t* row_of, const double* val, double scale,
                     int32_t* col_t, double* val_t, hipStream_t s);
// dest = a1*v1 + b1*A^T w over the transpose (rp, cl, vl) of A (n = A's columns)
int axpy_transposed(int64_t n, const int32_t* rp, const int32_t* cl, const double* vl, double a1, const double* v1,
                    bool v1_is_dest, double b1, const double* w, double* dest, hipStream_t s);
// *missing_flag = 1 if a listed row stores no diagonal entry
int missing_diagonals(int64_t nlist, const int32_t* rows, const int32_t* rowptr, const int32_t* col,
                      int* missing_flag, hipStream_t s);
// SetDirichletRow for the listed rows (every one stores its diagonal)
int dirichlet_rows(int64_t nlist, const int32_t* rows, const int32_t* rowptr, const int32_t* col, double* val,
                   hipStream_t s);
int scale_values(int64_t n, const double* src, double a, double* dst, hipStream_t s);

// ---- element-loop assembly from connectivity (elem_assembly.hip) ----
// temporary bytes of the pattern's sort and scan for ne elements of nv DoFs
// rowptr[0] = 0, rowptr[i+1] = cnt[0] + ... + cnt[i], scanned in 64 bits into
// incl (n entries; incl[n-1] is the total, which the caller checks against
// the int32 range of rowptr)
int rowptr_scan_temp_bytes(int64_t n, size_t* bytes);
int rowptr_from_counts(int64_t n, const int32_t* cnt, int64_t* incl, void* temp, size_t bytes, int32_t* rowptr,
                       hipStream_t s);
int elem_pattern_temp_bytes(int64_t ne, int nv, int64_t ndofs, size_t* sort_bytes, size_t* scan_bytes);
// sorts the element couplings (keys/idx into key2/idx2), marks run heads,
// scans them (run_incl), fills rowptr (ndofs + 1) and *nnz_dev; buffers of
// ne*nv*nv entries, temp of max(sort_bytes, scan_bytes)
int elem_pattern(int64_t ne, int nv, int64_t ndofs, const int32_t* dofs, uint64_t* key, uint64_t* key2, int32_t* idx,
                 int32_t* idx2, int32_t* head, int32_t* run_incl, void* temp, size_t sort_bytes, size_t scan_bytes,
                 int32_t* rowptr, int32_t* nnz_dev, hipStream_t s);
// the columns (nnz) and the CSR slot of every local entry (ne*nv*nv, index e*nv*nv + a*nv + b)
int elem_pattern_fill(int64_t ne, int nv, const uint64_t* key2, const int32_t* idx2, const int32_t* head,
                      const int32_t* run_incl, int32_t* col, int32_t* slot_local, hipStream_t s);
int elem_incidence_temp_bytes(int64_t n, int64_t maxkey, size_t* bytes);
// DoF -> incidences (t = e*nv + a) sorted by DoF: inc_ptr (ndofs + 1), inc_t (ne*nv)
int elem_incidence(int64_t ne, int nv, int64_t ndofs, const int32_t* dofs, uint32_t* key, uint32_t* key2, int32_t* idx,
                   int32_t* inc_t, void* temp, size_t temp_bytes, int32_t* inc_ptr, hipStream_t s);
// one colouring round; *coloured += elements coloured, *overflow = 1 past 64 colours
int elem_colour_round(int64_t ne, int nv, const int32_t* dofs, const int32_t* inc_ptr, const int32_t* inc_t,
                      int32_t* colour, int* coloured, int* overflow, hipStream_t s);
// elements sorted by colour (order) and colour_ptr (ncolours + 1)
int elem_colour_order(int64_t ne, int ncolours, const int32_t* colour, uint32_t* key, uint32_t* key2, int32_t* idx,
                      int32_t* order, void* temp, size_t temp_bytes, int32_t* colour_ptr, hipStream_t s);
int elem_slots_by_position(int64_t ne, int nv, const int32_t* order, const int32_t* slot_local, int32_t* slot,
                           hipStream_t s);
struct ElemLaunch {
    int64_t ne = 0;
    const int32_t* order = nullptr;
    const int32_t* dofs = nullptr;
    const double* coords = nullptr;
    const double* kel = nullptr;
    double f_const = 0.0;
    const double* f_dof = nullptr;
    double vel[3] = {0.0, 0.0, 0.0};
    int conv = 0;
    const int32_t* slot = nullptr;
    double* val = nullptr;
    double* rhs = nullptr;
    const double* u = nullptr;
    double* d = nullptr;
};
// one colour's elements (positions pos0 .. pos0+cnt-1 of the colour order);
// elem_type 0 P1 triangle FV1, 1 P1 tetrahedron FV1, 2 Q1 hexahedron FE
int elem_assemble_colour(int elem_type, const ElemLaunch& L, int64_t pos0, int64_t cnt, hipStream_t s);
int zero_listed(int64_t n, const int32_t* idx, double* x, hipStream_t s);
// Q2 hexahedra (27 DoFs in the local shape order of LagrangeLSFS<Hexahedron,
// 2>): one colour, one element per wavefront (q2_kernels.hip); the
// coefficient per element (kel), per element and Gauss point (kq) or 1;
// val (through the slots) and rhs may be null
int elem_assemble_q2_colour(int64_t ne, const int32_t* order, const int32_t* dofs, const double* coords,
                            const double* kel, const double* kq, double f_const, const int32_t* slot, double* val,
                            double* rhs, int64_t pos0, int64_t cnt, hipStream_t s);
// element -> DoF table and k = exp(theta) per element of a lattice domain's
// top level (kind 0 hexahedra, 1 triangles, 2 tetrahedra)
int lattice_connectivity(int kind, int n, const int32_t* dof_of_lat, const uint16_t* cell_tets, const double* theta,
                         int32_t* dofs, double* kel, hipStream_t s);

// ---- Q2 on the hexahedral lattice hierarchy (q2_hierarchy.hip) ----
// numbering of level l from the P1 numbering of level l+1 (N1 vertices,
// n_inner1 Inner ones, lattice width W1 = 2 n_l + 1): node_of_dof and
// dof_of_node over the level-(l+1) lattice, Dirichlet flag per DoF; key,
// key2, node: N1 entries of scratch, temp: q2_numbering_temp_bytes
int q2_numbering_temp_bytes(int64_t N1, size_t* bytes);
int q2_numbering(int64_t N1, int64_t n_inner1, int W1, const int32_t* lat_of_dof1, uint64_t* key, uint64_t* key2,
                 int32_t* node, int32_t* node_of_dof, void* temp, size_t temp_bytes, int32_t* dof_of_node,
                 uint8_t* dir, hipStream_t s);
// [n^3][27] DoFs of the level's hexahedra (lattice order) in the Q2 shape order
int q2_connectivity(int n, const int32_t* dof_of_node, int32_t* conn, hipStream_t s);
int q2_coords(int64_t N, int n, const int32_t* node_of_dof, double* X, hipStream_t s);
// exp(theta) of the fine element (nfine per axis) at every Gauss point
int q2_coef_points(int n, int nfine, const double* theta, double* kq, hipStream_t s);
// prolongation from level l-1 to level l (nf elements per axis): row counts,
// then columns (ascending) and values incl. the Dirichlet rows
int q2_prolong_count(int64_t Nf, int nf, const int32_t* node_of_dof_f, int32_t* cnt, hipStream_t s);
int q2_prolong_fill(int64_t Nf, int nf, const int32_t* node_of_dof_f, const uint8_t* dir_f,
                    const int32_t* dof_of_node_c, const int32_t* rowptr, int32_t* col, double* val, hipStream_t s);
// the Dirichlet rows of R = P^T (last: Nc entries, all -1 on entry)
int q2_restriction_adjust(int64_t Nf, int nf, int64_t Nc, const int32_t* node_of_dof_f, const uint8_t* dir_f,
                          const int32_t* dof_of_node_c, int32_t* last, const int32_t* rowptr, const int32_t* col,
                          double* val, hipStream_t s);

// ---- lattice operators written row by row (lattice_csr.hip) ----
// Q1 unit-cube tables, 2x2x2 Gauss points (host): G_q[a][b] =
// w_q grad phi_a . grad phi_b at 64q + 8a + b, K = sum_q G_q at 512 + 8a + b,
// F_a = sum_q w_q phi_a at 576 + a
constexpr int kQ1Tab = 584;
void q1_lattice_tables(double* tab);
// rp_lat[p] = rowptr[dof_of_lat[p]] (row start of every lattice point)
int rowptr_of_lattice(int64_t npts, const int32_t* dof_of_lat, const int32_t* rowptr, int32_t* rp_lat, hipStream_t s);
// col of the lattice pattern (rows = DoFs; 3D: 27-point, 2D: the triangles'
// 7-point pattern); rp_lat from rowptr_of_lattice
int lattice_pattern(int dim, int W, const int32_t* dof_of_lat, const int32_t* rp_lat, int32_t* col, hipStream_t s);
int exp_values(int64_t n, const double* theta, double* out, hipStream_t s);
// Q1 top-level operator into the lattice pattern's CSR (every entry written)
// and the load vector; kel: exp(theta) per element or null
int q1_lattice_csr(int n, double h, double f_const, const double* tab, const double* kel,
                   const int32_t* dof_of_lat, const int32_t* rp_lat, double* val, double* rhs, hipStream_t s);
// Q1 operator of a level (n elements per axis) into its 27 tiled diagonals
// (every diagonal of every row written); kfine: exp(theta) per fine element
// (nfine per axis, a multiple of n) or null
int q1_lattice_dia(int n, int nfine, double h, const double* tab, const double* kfine, double* dia, hipStream_t s);

// ---- sliced ELL with a column dictionary per slice (spmv_dict.hip) ----
// slot s of the slice order holds row order[s] or -1; a slice is 64 slots
int dict_widths(int64_t nslices, int64_t nrows, const int32_t* rowptr, const int32_t* order, int32_t* width,
                hipStream_t s);
// distinct columns per slice (INT_MAX: the slice is wider than the layout takes)
int dict_count(int64_t nslices, int64_t nrows, const int32_t* rowptr, const int32_t* col, const int32_t* order,
               const int32_t* width, int32_t* ucount, hipStream_t s);
int dict_fill(int64_t nslices, int64_t nrows, const int32_t* rowptr, const int32_t* col, const double* val,
              const int32_t* order, const int32_t* width, const int32_t* dptr, const int64_t* vptr,
              const int64_t* iptr, int32_t* dict, double* sval, uint32_t* sidx, hipStream_t s);
// the values alone into an existing dictionary layout of the same pattern
int dict_values(int64_t nslices, int64_t nrows, const int32_t* rowptr, const double* val, const int32_t* order,
                const int32_t* width, const int64_t* vptr, double* sval, hipStream_t s);
// y = alpha*v1 + beta*A*x (mode 0) or y = beta*A*x on rows with entries (mode 1)
int dict_spmv(int64_t nrows, int64_t nslices, const int32_t* width, const int32_t* dptr, const int32_t* dict,
              const int64_t* vptr, const int64_t* iptr, const double* sval, const uint32_t* sidx,
              const int32_t* order, double alpha, const double* v1, double beta, const double* x, double* y, int mode,
              const int32_t* rowptr, hipStream_t s);
// the slots of a lattice's rows in blocks of 64 points (4x4x4 or 8x8)
int64_t lattice_block_slots(int dim, int W);
int lattice_block_order(int dim, int W, const int32_t* dof_of_lat, int32_t* order, hipStream_t s);

}  // namespace ugamd
