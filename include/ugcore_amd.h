/*
 * ugcore_amd C ABI: the drop-in boundary of the MI355X discretize-and-solve
 * path for ugcore. Plain pointers, sizes and opaque handles; no C++ or torch
 * types. Every function returns 0 (or a non-negative count/handle) on success
 * and a negative value on failure; ugx_last_error() then describes the error.
 *
 * Two layers:
 *  1. Kernel entry points (device pointers, stream ordered) in the style of the
 *     reference's C wrappers ugbase/lib_algebra/gpu_algebra/cuda/common_cuda.h:36-49
 *     (extern "C" bool CUDA_VecAdd2(int len, double a, double* x, double b, const double* y)).
 *  2. Object entry points (handles) over the GPUAlgebra C++ classes in
 *     include/ug_gpu/gpu_algebra.hpp, which the ugcore bridge registers as the
 *     "GPU1" algebra; the Python bindings (ugcore_amd/__init__.py) bind these.
 *
 * `stream` arguments may be NULL, meaning the library's own stream.
 */
#ifndef UGCORE_AMD_H
#define UGCORE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library ------------------------------------------------------------ */
int ugx_init(void);                                  /* selects/initialises the device */
const char* ugx_last_error(void);
void* ugx_stream(void);                              /* the library's hipStream_t */
int ugx_synchronize(void);

/* ---- device buffers and timing (plumbing for callers without a GPU framework) */
void* ugx_malloc(int64_t bytes);
int ugx_free(void* p);
int ugx_memcpy_htod(void* dst_dev, const void* src_host, int64_t bytes);   /* synchronous */
int ugx_memcpy_dtoh(void* dst_host, const void* src_dev, int64_t bytes);   /* synchronous */
int ugx_memcpy_dtod(void* dst_dev, const void* src_dev, int64_t bytes);    /* synchronous */
void* ugx_event_create(void);
int ugx_event_destroy(void* ev);
int ugx_event_record(void* ev, void* stream);
/* waits for `end` and returns the elapsed milliseconds between the two events */
int ugx_event_elapsed_ms(void* begin, void* end, double* ms);

/* ---- kernel entry points: CSR (int32 row pointers/columns, sorted rows) ---- */

/* dest = alpha*v1 + beta*A*w. Replaces SparseMatrix::axpy
 * (lib_algebra/cpu_algebra/sparsematrix_impl.h:291-339) and the dead
 * cusparseDcsrmv call (lib_algebra/gpu_algebra/gpusparsematrix_impl.h:156-164).
 * v1 may alias dest; w must not. alpha == 0 ignores v1 (empty rows give 0). */
int ugx_spmv_csr_f64(int64_t nrows, const int32_t* rowptr, const int32_t* col, const double* val, double alpha,
                     const double* v1, double beta, const double* w, double* dest, void* stream);
int ugx_spmv_csr_f32(int64_t nrows, const int32_t* rowptr, const int32_t* col, const float* val, double alpha,
                     const float* v1, double beta, const float* w, float* dest, void* stream);
/* dest[i] = beta*(A*w)[i] for rows with entries; empty rows untouched.
 * Replaces SparseMatrix::apply_ignore_zero_rows (sparsematrix_impl.h:271-288). */
int ugx_spmv_csr_ignore_zero_rows_f64(int64_t nrows, const int32_t* rowptr, const int32_t* col, const double* val,
                                      double beta, const double* w, double* dest, void* stream);

/* ---- kernel entry points: vectors ---------------------------------------- */
/* VecScaleAdd 2-/3-term (lib_algebra/common/operations_vec.h:55-63, 162-175),
 * replacing CUDA_VecAdd_2/CUDA_VecAdd_3 (gpu_algebra/cuda/common_cuda.cu:86-100). */
int ugx_vec_scale_add2_f64(int64_t n, double* dest, double a1, const double* v1, double a2, const double* v2,
                           void* stream);
int ugx_vec_scale_add3_f64(int64_t n, double* dest, double a1, const double* v1, double a2, const double* v2,
                           double a3, const double* v3, void* stream);
int ugx_vec_set_f64(int64_t n, double* x, double value, void* stream);
/* Vector::dotprod / norm (cpu_algebra/vector_impl.h:70-79, 323-329), replacing
 * cublasDdot/cublasDnrm2 (gpu_algebra/gpuvector.h:200-214). Deterministic
 * two-pass reduction; synchronises and writes the result to host memory. */
int ugx_dot_f64(int64_t n, const double* a, const double* b, double* result_host, void* stream);
int ugx_norm_f64(int64_t n, const double* a, double* result_host, void* stream);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ugx_dot_f64 without any allocation. The partial sums go
 * to a fixed device array of the library (65536 doubles = 512 KiB per device,
 * held until the process ends; the reduction grid never needs more, so n is
 * not limited further), the result to result_dev (device memory, written on
 * the stream, no synchronisation). Same kernels and grid as ugx_dot_f64, so
 * the same value bit for bit. stream NULL means the library's stream, which
 * exists only after ugx_init() (the call fails otherwise). Every call on one
 * device shares that device's array: calls that overlap on one device race. */
int ugx_dot_f64_noalloc(int64_t n, const double* a, const double* b, double* result_dev, void* stream);

/* ---- kernel entry points: Jacobi (lib_algebra/operator/preconditioner/jacobi.h:155-249,
 * replacing GPUJacobi::preprocess + CUDA_JacobiApply, gpujacobi.h:113-144, common_cuda.cu:46-52) */
/* dinv_i = 1/(a_ii*(1/omega)); returns -2 if a diagonal entry is zero */
int ugx_jacobi_diag_inv_f64(int64_t nrows, const int32_t* rowptr, const int32_t* col, const double* val, double omega,
                            double* dinv, void* stream);
/* c = dinv .* d */
int ugx_jacobi_apply_f64(int64_t n, double* c, const double* dinv, const double* d, void* stream);

/* ---- object entry points --------------------------------------------------- */

/* Host-only (no device use): LEVEL DoF numbering of a structured hierarchy as
 * produced by lib_grid's GlobalMultiGridRefiner + DoFDistribution
 * (global_multi_grid_refiner.cpp:135-455, dof_distribution.cpp:1401-1421).
 * kind 0: unit cube from one hexahedron, kind 1: unit square from two
 * triangles, kind 2: unit cube as 3x3x3 hexahedra x 6 tetrahedra (regular
 * tetrahedron rule, tetrahedron_rules.cpp:487-609).
 * out_host receives W^dim entries, W = 2^level+1 (kind 2: 3*2^level+1):
 * lattice point (x fastest) -> DoF.
 * n_inner receives the number of DoFs of subset "Inner" (the rest are Dirichlet). */
int ugx_hierarchy_numbering(int kind, int num_refs, int level, int32_t* out_host, int64_t* n_inner);
/* Host-only: the tetrahedra of level `level` of the kind-2 hierarchy, per
 * lattice cell (x fastest) 6 entries in creation order, each packing the 4
 * cell-corner indices (bit 0 x, bit 1 y, bit 2 z) of the tetrahedron's local
 * vertex order as c0 | c1<<3 | c2<<6 | c3<<9; parent_edge (level >= 1, may be
 * NULL) receives per lattice vertex 0 (copy of a coarse vertex) or the
 * direction code 1..13 of the coarse edge it halves. No reference counterpart:
 * it exposes the refinement result the reference keeps inside lib_grid. */
int ugx_hierarchy_tets(int num_refs, int level, uint16_t* cell_tets, uint8_t* parent_edge);

/* structured domain + DoF distribution + DomainDiscretization
 * (lib_disc/spatial_disc/domain_disc_impl.h:469-948 for the assembly; the DoF
 * numbering follows lib_disc/dof_manager/dof_distribution.cpp:1401-1421).
 * kind 0: unit cube from one hexahedron, Q1 FE for -div(k grad u) = f;
 * kind 1: unit square from two triangles, FV1 box method;
 * kind 2: unit cube as 3x3x3 hexahedra x 6 tetrahedra, FV1 for
 * -div(k grad u) + div(v u) = f with full upwind convection
 * (ConvectionDiffusion FV1 with ConvectionShapesFullUpwind, conv_shape.h:278-330),
 * theta per fine lattice cell. */
void* ugx_disc_create(int kind, int num_refs);
/* ordering 0: the reference's numbering (as ugx_disc_create); 1: lexicographic
 * lattice order with Dirichlet DoFs last (OrderLex-style renumbering) */
void* ugx_disc_create_ordered(int kind, int num_refs, int ordering);
/* setup 0: the multigrid hierarchy and its DoF numbering are built on the
 * device (GlobalMultiGridRefiner::perform_refinement,
 * global_multi_grid_refiner.cpp:135-455, and DoFDistribution::reinit,
 * dof_distribution.cpp:1401-1421, replayed in parallel); setup 1: built on the
 * host and uploaded. Both give the same numbering; ugx_disc_create and
 * ugx_disc_create_ordered use setup 0. */
void* ugx_disc_create_setup(int kind, int num_refs, int ordering, int setup);
/* kind 2: the level's tetrahedra per lattice cell (6 packed corner codes per
 * cell, n^3 cells, creation order) and, for levels >= 1, the parent-edge
 * direction code of every lattice point ((n+1)^3 entries, 0 for copies of
 * coarse vertices); either pointer may be NULL */
int ugx_disc_level_tets(void* disc, int level, uint16_t* cell_tets_host, uint8_t* parent_codes_host);
void ugx_disc_destroy(void* disc);
int ugx_disc_num_levels(void* disc);
int64_t ugx_disc_num_dofs(void* disc, int level);
/* DoFs [0, num_inner) of the level belong to subset "Inner", the rest are
 * Dirichlet DoFs; -1 for a level out of range */
int64_t ugx_disc_num_inner(void* disc, int level);
int ugx_disc_dof_of_lattice(void* disc, int level, int32_t* out_host);   /* (n+1)^dim entries */
/* DoF coordinates of a level in DoF order, dim (2 or 3) per DoF: what the
 * approximation space hands to SaveMatrixForConnectionViewer
 * (grid_function_util.h); tetrahedra: the jittered top-level positions */
int ugx_disc_dof_positions(void* disc, int level, double* out_host);
int ugx_disc_set_theta(void* disc, const double* theta_host);            /* k = exp(theta) per fine element */
int ugx_disc_set_source_const(void* disc, double f);
int ugx_disc_set_source_dofs(void* disc, const double* f_host);          /* per top-level DoF (FV1) */
int ugx_disc_set_theta_device(void* disc, const double* theta_dev);
/* zero the Dirichlet entries of a top-level vector handle (DirichletBoundary::adjust_defect,
 * lagrange_dirichlet_boundary_impl.h:920-1035) */
int ugx_disc_adjust_defect(void* disc, void* vec);
/* Adjoint gradient (no counterpart in the reference): for J(u(theta)) with
 * A(theta) u = b and the adjoint solution lambda of A^T lambda = dJ/du,
 * grad_dev[e] = -lambda_e^T (dA_e/dtheta_e) u_e for every fine element e
 * (lattice order, n_fine^3 doubles, device memory). Q1 hexahedra (kind 0). */
int ugx_disc_adjoint_gradient(void* disc, void* u, void* lambda, double* grad_dev);
/* kind 2: ConvectionDiffusion::set_velocity with a constant vector */
int ugx_disc_set_velocity(void* disc, double vx, double vy, double vz);
/* kind 2: jitter of the top-level interior vertices, amp * h * (2U-1) per
 * component, U uniform from the seeded counter generator (counter 3 * lattice
 * index + component); amp in [0, 0.5). No counterpart in the reference (a
 * grid modification a script would make before assembling). */
int ugx_disc_set_jitter(void* disc, uint64_t seed, double amp);
/* operator of the adjoint solve (no counterpart in the reference; replaces
 * SparseMatrix::set_as_transpose_of, sparsematrix_impl.h:148-183, applied to
 * the Dirichlet-eliminated matrix): at = (A with Dirichlet columns of interior
 * rows zeroed)^T with identity Dirichlet rows, on A's structurally symmetric
 * pattern; a from ugx_disc_assemble_linear */
int ugx_disc_adjoint_operator(void* disc, void* a, void* at);
/* Element matrices of n Q2 (triquadratic Lagrange) hexahedra, the local
 * stiffness k * int grad phi_a . grad phi_b and load f * int phi_a an IElemDisc
 * adds per element (add_jac_A_elem / add_rhs_elem, elem_disc_interface.h:
 * 470-752) in the local shape order of LagrangeLSFS<ReferenceHexahedron, 2>
 * (lagrange.cpp:236-253, lagrange.h:2106-2190), trilinear geometry, 3x3x3
 * Gauss points. Device pointers: corners [n][8][3] in the reference
 * hexahedron order, k [n] (NULL: 1; must hold n entries), ke [n][27][27]
 * row-major, fe [n][27] (NULL: not formed). Enqueued on the library stream. */
int ugx_q2_hex_element_matrices(int64_t n, const double* corners, const double* k, double f_const, double* ke,
                                double* fe);
/* assemble_linear (domain_disc_impl.h:835-948) of -div(k grad u) = f with
 * Q2 Lagrange elements on the unit cube split into n^3 hexahedra, u = 0 on the
 * boundary (identity rows on the kept pattern, SetDirichletRow,
 * sparsematrix_util.h:850-861): k = exp(theta[e]) per element (theta_dev
 * device pointer to n^3 entries, x fastest, or NULL: 1), f constant;
 * 1 <= n <= 161 so that nnz = (8n+1)^3 fits int32 (checked before any
 * allocation). Nodes are the
 * (2n+1)^3 lattice points numbered x fastest (not the reference's DoF order,
 * which lib_grid fixes); element matrices as ugx_q2_hex_element_matrices,
 * scattered colour by colour without atomics into a pattern built on the
 * device. Fills the matrix and vector handles. */
int ugx_q2_assemble_lattice(int n, const double* theta_dev, double f_const, void* matrix, void* rhs);
/* DomainDiscretization::assemble_linear(A, b) (domain_disc_impl.h:835-948) incl. Dirichlet rows */
int ugx_disc_assemble_linear(void* disc, void* matrix, void* rhs);
double ugx_disc_assemble_seconds(void* disc);

/* GPUSparseMatrix (matrix_type) wrapped as MatrixOperator (matrix_operator.h:47) */
/* DomainDiscretization::assemble_jacobian (domain_disc_impl.h:469-600): the
 * discretisations here are linear in u, so J is the matrix of
 * ugx_disc_assemble_linear with its Dirichlet rows */
int ugx_disc_assemble_jacobian(void* disc, void* matrix, void* u);
/* DomainDiscretization::assemble_defect (domain_disc_impl.h:646-788;
 * AssembleDefect, elem_disc_assemble_util.h:577-687): d = sum_e (A_e u_e - f_e)
 * by the element loop over the top level, Dirichlet entries 0
 * (DirichletBoundary::adjust_defect, lagrange_dirichlet_boundary_impl.h:920-1035) */
int ugx_disc_assemble_defect(void* disc, void* d, void* u);

/* ---- element-loop assembly from caller-supplied connectivity -------------
 * The element loop of DomainDiscretization::assemble_linear / _jacobian /
 * _defect (domain_disc_impl.h:469-948; StdGlobAssembler,
 * elem_disc_assemble_util.h:326-413, 577-687, 902-1003) over one element type.
 * elem_dofs: num_elems x nv DoF indices per element in the reference element's
 * local vertex order (what DoFDistribution::indices gives per element,
 * dof_distribution.cpp:1117-1206); dof_coords: num_dofs x dim coordinates.
 * elem_type 0: P1 triangles, FV1 (nv 3, dim 2); 1: P1 tetrahedra, FV1 with
 * optional full-upwind convection (nv 4, dim 3); 2: Q1 hexahedra, FE with
 * 2x2x2 Gauss (nv 8, dim 3); 3: Q2 hexahedra, FE with 3x3x3 Gauss (nv 27 in
 * the shape order of LagrangeLSFS<ReferenceHexahedron, 2>, lagrange.cpp:236-253:
 * the 8 vertices, the 12 edge midpoints, the 6 face centres, the centre;
 * dim 3; element matrices on the matrix cores, one element per wavefront;
 * its defect is A u - b of the assembled operator). Host arrays.
 * num_elems * nv^2 < 2^31.
 * The CSR pattern holds every element coupling (rows sorted by column, zero
 * couplings kept, as AddLocalMatrixToGlobal's insertion, local_algebra.h:766-789);
 * it, the CSR slot of every local entry and an element colouring are built
 * once on the device here; each assembly is one atomic-free launch per colour. */
void* ugx_elemdisc_create(int elem_type, int64_t num_elems, const int32_t* elem_dofs, int64_t num_dofs,
                          const double* dof_coords);
void ugx_elemdisc_destroy(void* e);
int64_t ugx_elemdisc_nnz(void* e);
int ugx_elemdisc_num_colours(void* e);
/* diffusion k per element (host array of num_elems; NULL: 1) */
int ugx_elemdisc_set_diffusion(void* e, const double* k_host);
/* Q2 only: k at the 3x3x3 Gauss points of every element (host array of
 * num_elems x 27, x fastest; NULL: back to ugx_elemdisc_set_diffusion), the
 * coefficient function evaluated where the reference's quadrature evaluates it */
int ugx_elemdisc_set_diffusion_at_points(void* e, const double* kq_host);
/* source f: a constant, or per DoF (FV1 types); FV1 takes f |scv|, Q1 f int phi */
int ugx_elemdisc_set_source_const(void* e, double f);
int ugx_elemdisc_set_source_dofs(void* e, const double* f_host);
/* convection velocity (tetrahedra only; full upwind) */
int ugx_elemdisc_set_velocity(void* e, double vx, double vy, double vz);
/* homogeneous Dirichlet DoFs: identity rows (SetDirichletRow), rhs and defect 0 */
int ugx_elemdisc_set_dirichlet(void* e, const int32_t* dofs_host, int64_t n);
int ugx_elemdisc_assemble_linear(void* e, void* matrix, void* rhs);
int ugx_elemdisc_assemble_jacobian(void* e, void* matrix, void* u);
int ugx_elemdisc_assemble_defect(void* e, void* d, void* u);

void* ugx_matrix_create(void);
void ugx_matrix_destroy(void* m);
int64_t ugx_matrix_num_rows(void* m);
int64_t ugx_matrix_nnz(void* m);
int ugx_matrix_upload_csr(void* m, int64_t rows, int64_t cols, const int32_t* rowptr, const int32_t* col,
                          const double* val);
int ugx_matrix_download_csr(void* m, int32_t* rowptr, int32_t* col, double* val);
/* Raw device CSR pointers (write access): from this call on the matrix no
 * longer counts as the discretization's assembly, so no operator derived from
 * that assembly (matrix-free top level, lattice-order CG) is used with it, and
 * SpMV reads the CSR until ugx_matrix_finalize rebuilds the sliced-ELL layout
 * from the edited values. */
int ugx_matrix_device_csr(void* m, int32_t** rowptr, int32_t** col, double** val);
int ugx_matrix_axpy(void* m, void* dest, double alpha, void* v1, double beta, void* w);   /* vectors are handles */
/* tuning: lanes of a wavefront that share one CSR row in this matrix's SpMV (4..64, power of 2) */
int ugx_matrix_set_spmv_lanes(void* m, int lanes);
/* SpMV layout: 2 = sliced ELL with a column dictionary per slice where the
 * matrix has a slice order (operators the lattice discretisations assemble:
 * 64 rows of a 4x4x4 or 8x8 block per slice, one-byte column positions, x
 * staged once per slice; default), else sliced ELL; 1 = sliced ELL (64-row
 * slices); 0 = CSR directly. Every layout is built from the CSR and sums each
 * row in the same order as the sliced ELL. Call ugx_matrix_finalize after
 * changing values through the raw CSR pointers of ugx_matrix_device_csr. */
int ugx_matrix_set_spmv_layout(void* m, int layout);
/* the layout SpMV runs on now: 0 CSR, 1 sliced ELL, 2 dictionary */
int ugx_matrix_spmv_layout(void* m);
/* Row permutation that fills the sliced-ELL slices (host array of n = rows
 * entries, checked to be a permutation; order = NULL clears it). Rows whose
 * columns lie close together should be neighbours in it, so that the column
 * gathers of one 64-row slice stay local; SpMV results do not depend on it.
 * Assembly by a discretization sets its lattice walk. No counterpart in the
 * reference (SparseMatrix has one storage, sparsematrix.h). */
int ugx_matrix_set_row_order(void* m, const int32_t* order, int64_t n);
int ugx_matrix_has_row_order(void* m);
/* Slices of the dictionary layout (host array of nslots entries, a multiple of
 * 64): slice t holds rows order[64t .. 64t+63], -1 marks an empty slot, every
 * row appears exactly once (checked). The layout is used when every slice's
 * entries touch at most 255 distinct columns and no row of a slice is longer
 * than 64; otherwise SpMV keeps the sliced ELL. order = NULL clears it.
 * Assembly by a discretization sets 4x4x4 (8x8) lattice blocks. No
 * counterpart in the reference. */
int ugx_matrix_set_slice_order(void* m, const int32_t* order, int64_t nslots);
/* after editing values through ugx_matrix_device_csr: rebuilds the SpMV layout */
int ugx_matrix_finalize(void* m);
int ugx_matrix_apply_sub(void* m, void* d, void* x);                                      /* d -= A*x */

/* ---- the SparseMatrix surface of the reference's callers
 * (lib_algebra/cpu_algebra/sparsematrix.h:98-752, sparsematrix_impl.h) ---- */
int64_t ugx_matrix_num_cols(void* m);
/* the device CSR for reading only (the provenance is kept, unlike
 * ugx_matrix_device_csr): for kernels of callers that read the matrix */
int ugx_matrix_device_csr_readonly(void* m, const int32_t** rowptr, const int32_t** col, const double** val);
/* dest[i] = beta*(A w)[i] for rows with entries, other rows untouched
 * (apply_ignore_zero_rows, sparsematrix_impl.h:271-288) */
int ugx_matrix_apply_ignore_zero_rows(void* m, void* dest, double beta, void* w);
/* resize_and_clear (sparsematrix_impl.h:90-109): drops the device CSR and
 * starts a host-side pattern build; entries are created by ugx_matrix_entry
 * and moved to the device by ugx_matrix_finalize (defragment,
 * sparsematrix_impl.h:703-757). Sizes below 2^31. */
int ugx_matrix_resize_and_clear(void* m, int64_t rows, int64_t cols);
/* resize_and_keep_values (sparsematrix_impl.h:111-133): a host build that
 * keeps the entries (a device CSR is downloaded into it), dropping columns >= cols */
int ugx_matrix_resize_and_keep_values(void* m, int64_t rows, int64_t cols);
/* clear_retain_structure (sparsematrix_impl.h:136-140): values 0, pattern kept */
int ugx_matrix_clear_retain_structure(void* m);
/* 1 between resize_and_clear and finalize */
int ugx_matrix_is_building(void* m);
/* operator()(r, c) (sparsematrix.h:293-300; get_index, sparsematrix_impl.h:553-700):
 * the address of the entry's host-side value, the entry created with value 0
 * on first access (rows stay sorted by column); valid until the next
 * insertion. Host build only: NULL once the matrix is finalised on the device. */
double* ugx_matrix_entry(void* m, int64_t r, int64_t c);
/* const operator()(r, c): *value = the stored value or 0; returns 1 if the
 * entry is stored, 0 if not (host build or device CSR) */
int ugx_matrix_get(void* m, int64_t r, int64_t c, double* value);
/* begin_row/end_row (sparsematrix.h:484-487): the number of stored entries of
 * row r, and a host copy of its columns and values (either pointer may be NULL) */
int64_t ugx_matrix_row_length(void* m, int64_t r);
int ugx_matrix_row(void* m, int64_t r, int32_t* cols, double* vals);
/* m = scale * src^T on the device: keep_zeros 1 keeps every stored entry, as
 * set_as_transpose_of (sparsematrix_impl.h:148-183); 0 drops zero entries, as
 * set_as_transpose_of2 (:185-240). Rows sorted by column. */
int ugx_matrix_set_as_transpose_of(void* m, void* src, double scale, int keep_zeros);
/* set_as_copy_of (sparsematrix_impl.h:471-481) and scale (sparsematrix_impl.h:486-496) */
int ugx_matrix_set_as_copy_of(void* m, void* src, double scale);
int ugx_matrix_scale(void* m, double a);
/* dest = alpha*v1 + beta*A^T*w (axpy_transposed, sparsematrix_impl.h:341-370):
 * per column in ascending row, zero entries skipped; v1 may be dest */
int ugx_matrix_axpy_transposed(void* m, void* dest, double alpha, void* v1, double beta, void* w);
/* SetDirichletRow (algebra_common/sparsematrix_util.h:850-861) for every
 * listed row: entries 0, diagonal 1. On the device every listed row must
 * store its diagonal (the pattern is fixed there). */
int ugx_matrix_set_dirichlet_rows(void* m, const int32_t* rows_host, int64_t n);

/* GPUVector (vector_type) */
void* ugx_vector_create(int64_t n);
void ugx_vector_destroy(void* v);
int64_t ugx_vector_size(void* v);
double* ugx_vector_device_ptr(void* v);
int ugx_vector_upload(void* v, const double* host);
int ugx_vector_download(void* v, double* host);
int ugx_vector_set(void* v, double value);

/* Q2 Lagrange on the unit cube (one hexahedron refined num_refs <= 6 times):
 * -div(k grad u) = f, u = 0 on the boundary, k = exp(theta) per top-level
 * element taken at the Gauss points of every level's elements, f constant.
 * Every level in the reference's LEVEL DoF order (DoFDistribution::reinit,
 * dof_distribution.cpp:1401-1444: vertices, edges, faces, volumes, each
 * subset in creation order); operators by the element loop (elem_type 3 of
 * ugx_elemdisc_*); transfers as StdTransfer builds them for Lagrange elements
 * (std_transfer_impl.h:214-388) with the Dirichlet adjustments of
 * lagrange_dirichlet_boundary_impl.h:485-740. No counterpart as one class in
 * the reference (there: ApproximationSpace + DomainDiscretization +
 * StdTransfer). */
void* ugx_q2disc_create(int num_refs);
void ugx_q2disc_destroy(void* q);
int ugx_q2disc_num_levels(void* q);
int64_t ugx_q2disc_num_dofs(void* q, int level);
/* DoF coordinates (num_dofs x 3) and Dirichlet flags (num_dofs) of a level, host arrays */
int ugx_q2disc_dof_positions(void* q, int level, double* out_host);
int ugx_q2disc_dirichlet(void* q, int level, uint8_t* out_host);
/* theta per top-level element ((2^num_refs)^3, x fastest; host; NULL: k = 1) */
int ugx_q2disc_set_theta(void* q, const double* theta_host);
int ugx_q2disc_set_source(void* q, double f);
/* level operator and rhs (assemble_linear, domain_disc_impl.h:835-948) */
int ugx_q2disc_assemble_linear(void* q, int level, void* matrix, void* rhs);
/* StdTransfer::prolongation (level-1 -> level, rows on level) and restriction
 * (R = P^T of that P with adjust_restriction), std_transfer_impl.h:602-717 */
int ugx_q2disc_prolongation(void* q, int level, void* P);
int ugx_q2disc_restriction(void* q, int level, void* P, void* R);
/* AssembledMultiGridCycle over level operators and transfers given as
 * matrices (mg_solver_impl.hpp:175-2136): A[l] for l = 0..num_levels-1, P[l]
 * and R[l] for l >= 1 (P[0], R[0] ignored); damped Jacobi, dense LU on the
 * base level; the top level uses the operator given to ugx_mgmg_init or the
 * solver. The multigrid shares the matrices (they stay alive while it does,
 * whatever happens to the handles). As a preconditioner:
 * ugx_solver_set_preconditioner(kind 3). */
void* ugx_mgmg_create(int num_levels, void* const* A, void* const* P, void* const* R, int base_level, int nu1,
                      int nu2, int cycle, double omega);
void ugx_mgmg_destroy(void* g);
int ugx_mgmg_init(void* g, void* matrix);
int ugx_mgmg_apply(void* g, void* c, void* d);

/* AssembledMultiGridCycle (mg_solver_impl.hpp:175-2136): damped Jacobi
 * smoother, P1 transfer, dense LU on the base level.
 * cycle: 1 = V, 2 = W, gamma > 2, or -1 = F (one F-cycle then one V-cycle on
 * each coarser level, mg_solver_impl.hpp:2115-2118; mg_solver.h:339-341);
 * nu1, nu2 >= 0 smoothing steps; 0 <= base_level <= top level (checked at
 * ugx_gmg_init). precision: 0 = FP64 level operators, 1 = FP32 level
 * operators (base LU FP64). Anything else returns NULL with a message. */
void* ugx_gmg_create(void* disc, int base_level, int nu1, int nu2, int cycle, double omega, int precision);
void ugx_gmg_destroy(void* g);
/* before ugx_gmg_init: 1 = precondition the adjoint operator (level operators
 * A_l,elim^T, transfer unchanged); no counterpart in the reference */
int ugx_gmg_set_transposed(void* g, int transposed);
/* before ugx_gmg_init: 1 (default) = apply the top-level operator of the
 * hexahedral discretisation from its per-element coefficients (matrix-free,
 * same operator up to rounding), 0 = from the stored level matrix. No
 * counterpart in the reference (a storage choice behind apply/apply_sub). */
int ugx_gmg_set_matrix_free(void* g, int on);
/* before ugx_gmg_init: 1 = Galerkin coarse operators A_{l-1} = R_l A_l P_l
 * from the assembled top level (hexahedral and triangular hierarchies);
 * replaces AssembledMultiGridCycle::set_rap (mg_solver.h:149) */
int ugx_gmg_set_rap(void* g, int on);
/* before ugx_gmg_init: 1 (default) = the lattice-order cycle of the CG fast
 * path replays captured HIP graphs, 0 = launches its kernels one by one (same
 * kernels and results). No counterpart in the reference (launch mechanics). */
int ugx_gmg_set_graphs(void* g, int on);
/* before ugx_gmg_init: 1 (default) = on a matrix-free top level, the CG
 * fast path's cycle runs its last full post-smoothing step together with the
 * correction update that ends it (mg_solver_impl.hpp:1919-1943, the last
 * apply_sub and the final c += B d), the conversion of the correction and the
 * (c, d) product as one pass; 0 = separate passes. The correction is the
 * same; (c, d) is summed in another order. The switch has no counterpart in
 * the reference (kernel fusion). */
int ugx_gmg_set_fused_smoothing(void* g, int on);
int ugx_gmg_init(void* g, void* matrix);
int ugx_gmg_apply(void* g, void* c, void* d);                 /* c = B*d */
/* nsteps fused damped-Jacobi smoothing steps on one level's current defect
 * (st = dinv.*sd; sd -= A*st; sc += st), as in presmooth_and_restriction
 * (mg_solver_impl.hpp:1702-1730); used to time the smoother in isolation */
int ugx_gmg_smooth_level(void* g, int level, int nsteps);
/* y = A*x with the cycle's surface operator in lattice order (FP64, device
 * pointers to (n+1)^dim entries; x needs ugx_gmg_lattice_halo zero entries on
 * each side and must vanish on Dirichlet DoFs) */
int ugx_gmg_surface_apply_lattice(void* g, double* y_dev, const double* x_dev);
/* the lattice CG's direction update fused with the next operator application
 * (cg.h:234 p := beta p + z, then cg.h:166-169 q := A p, lambda = (q, p)):
 * p_new = z + (*beta_dev) p_old, q = A p_new, (q, p_new) into *qp_dev
 * (device); p_old and p_new are distinct lattice vectors with the halo,
 * partial_dev device scratch of 65536 doubles. Matrix-free top level only:
 * an error when it is not available. */
int ugx_gmg_surface_direction_lattice(void* g, double* q_dev, const double* p_old_dev, const double* z_dev,
                                      const double* beta_dev, double* p_new_dev, double* qp_dev,
                                      double* partial_dev);
/* c = B*d with the cycle applied in lattice order (device pointers to (n+1)^dim
 * FP64 entries, each with ugx_gmg_lattice_halo zero entries on either side; d
 * must vanish on Dirichlet DoFs): the preconditioner of the lattice CG, equal
 * to ugx_gmg_apply on the permuted vectors. Symmetric discretisations only.
 * d_f32_dev (may be NULL; same layout and halo): (float) d, which an FP32 cycle
 * reads instead of d when ugx_gmg_lattice_uses_f32_defect says so (the lattice
 * CG writes it along with its residual); the result is the same.
 * Replaces AssembledMultiGridCycle::apply (mg_solver_impl.hpp:175-275) for
 * callers that keep their vectors in lattice order. */
int ugx_gmg_apply_lattice(void* g, double* c_dev, const double* d_dev, const float* d_f32_dev);
int ugx_gmg_lattice_uses_f32_defect(void* g);
int64_t ugx_gmg_lattice_halo(void* g);
int ugx_gmg_level_num_diags(void* g, int level);
int ugx_gmg_level_dia(void* g, int level, double* out_host);  /* ndiag x (n+1)^dim values, lattice numbering */

/* Jacobi as a stand-alone preconditioner on a matrix handle */
void* ugx_jacobi_create(double omega);
void ugx_jacobi_destroy(void* j);

/* CG (cg.h:103-242) = 0, BiCGStab (bicgstab.h:112-385) = 1, with StdConvCheck */
void* ugx_solver_create(int type);
void ugx_solver_destroy(void* s);
/* precond_kind: 0 none, 1 GMG handle, 2 Jacobi handle, 3 ugx_mgmg handle */
int ugx_solver_set_preconditioner(void* s, int precond_kind, void* precond);
int ugx_solver_set_conv_check(void* s, int max_steps, double min_defect, double rel_reduction, int verbose);
int ugx_solver_init(void* s, void* matrix);
/* returns 1 if converged, 0 if not, <0 on error */
int ugx_solver_apply(void* s, void* x, void* b);
/* ILinearOperatorInverse::apply_return_defect (operator/interface/linear_operator_inverse.h:188):
 * solves A x = b from the initial x and leaves the final defect b - A x in b;
 * 1 converged, 0 not converged, -1 error */
int ugx_solver_apply_return_defect(void* s, void* x, void* b);
int ugx_solver_steps(void* s);
/* CG only: run on lattice-ordered copies of the vectors when the
 * preconditioner is a GMG built on the same assembly (1 = on, default) */
int ugx_solver_set_lattice_path(void* s, int on);
/* 1 if the last CG apply ran on the lattice path */
int ugx_solver_used_lattice_path(void* s);
int ugx_solver_defects(void* s, double* out_host, int max_n);
/* BiCGStab only: BiCGStab::set_restart (restart every n steps, 0 = never) and
 * set_min_orthogonality (restart when |(r0,r)| <= m ||r|| ||r0||),
 * bicgstab.h:386-390 */
int ugx_solver_set_restart(void* s, int n);
int ugx_solver_set_min_orthogonality(void* s, double m);

/* ---- file formats (host only, no GPU work) -------------------------------
 * ConnectionViewer text files: ConnectionViewer::WriteMatrix / WriteVector
 * (lib_algebra/common/connection_viewer_output.h:74-141, 402-422) with one
 * position (dim coordinates) per row; rowptr/col/val a host CSR. */
int ugx_cv_write_matrix(const char* path, int64_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                        const double* positions, int dim);
int ugx_cv_write_vector(const char* path, int64_t n, const double* values, const double* positions, int dim);
/* the serial SaveToFile / ReadFromFile vector file (bridge/algebra_bridges/
 * restart_bridge.cpp:49-75, 146-158; lib_algebra/cpu_algebra/vector_impl.h:
 * 347-361): host arrays; ugx_vector_file_length returns the stored length or
 * -1 (message in ugx_last_error) */
int ugx_vector_file_write(const char* path, int64_t n, const double* values);
int64_t ugx_vector_file_length(const char* path);
int ugx_vector_file_read(const char* path, int64_t n, double* values);
/* SaveToFile(v, filename) / ReadFromFile(v, filename) on a GPUVector handle
 * (restart_bridge.cpp:146-158); ReadFromFile resizes v */
int ugx_vector_save(void* v, const char* path);
int ugx_vector_read(void* v, const char* path);

#ifdef __cplusplus
}
#endif

#endif /* UGCORE_AMD_H */
