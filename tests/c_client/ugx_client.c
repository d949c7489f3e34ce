/* A C client of the drop-in boundary: it uses nothing but include/ugcore_amd.h
 * and libugcore_amd.so, the way a ugcore-side binding would (no Python, no
 * torch). Q1 diffusion on the unit cube refined num_refs times with
 * k = exp(theta), theta a smooth deterministic field; CG + GMG V(2,2) with FP32
 * level operators; checks ||b - A x|| <= 1e-9 ||b|| through the ABI's own
 * apply_sub and prints one summary line. Exit code 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "ugcore_amd.h"

static int fail(const char* what) {
    fprintf(stderr, "ugx_client: %s: %s\n", what, ugx_last_error());
    return 1;
}

int main(int argc, char** argv) {
    const int refs = argc > 1 ? atoi(argv[1]) : 5;
    if (ugx_init() != 0) return fail("ugx_init");
    void* disc = ugx_disc_create(0, refs);
    if (!disc) return fail("ugx_disc_create");
    const int n = 1 << refs;
    const int64_t ne = (int64_t)n * n * n;
    double* theta = (double*)malloc(sizeof(double) * (size_t)ne);
    for (int64_t e = 0; e < ne; ++e) {
        const double x = (e % n + 0.5) / n, y = ((e / n) % n + 0.5) / n, z = (e / ((int64_t)n * n) + 0.5) / n;
        theta[e] = sin(6.0 * x) * cos(4.0 * y) + 0.5 * z;
    }
    if (ugx_disc_set_theta(disc, theta) != 0) return fail("ugx_disc_set_theta");
    const int64_t N = ugx_disc_num_dofs(disc, refs);
    void* A = ugx_matrix_create();
    void* b = ugx_vector_create(N);
    void* x = ugx_vector_create(N);
    void* r = ugx_vector_create(N);
    if (!A || !b || !x || !r) return fail("create");
    if (ugx_disc_assemble_linear(disc, A, b) != 0) return fail("ugx_disc_assemble_linear");
    void* gmg = ugx_gmg_create(disc, 0, 2, 2, 1, 0.66, 1);
    void* cg = ugx_solver_create(0);
    if (!gmg || !cg) return fail("solver create");
    if (ugx_solver_set_preconditioner(cg, 1, gmg) != 0) return fail("set_preconditioner");
    if (ugx_solver_set_conv_check(cg, 200, 1e-50, 1e-10, 0) != 0) return fail("set_conv_check");
    if (ugx_solver_init(cg, A) != 0) return fail("ugx_solver_init");
    if (ugx_vector_set(x, 0.0) != 0) return fail("ugx_vector_set");
    const int ok = ugx_solver_apply(cg, x, b);
    if (ok < 0) return fail("ugx_solver_apply");
    /* r = b - A x */
    double* hb = (double*)malloc(sizeof(double) * (size_t)N);
    double* hr = (double*)malloc(sizeof(double) * (size_t)N);
    if (ugx_disc_assemble_linear(disc, A, b) != 0) return fail("re-assemble");
    if (ugx_vector_download(b, hb) != 0 || ugx_vector_upload(r, hb) != 0) return fail("copy b");
    if (ugx_matrix_apply_sub(A, r, x) != 0) return fail("ugx_matrix_apply_sub");
    if (ugx_vector_download(r, hr) != 0) return fail("download r");
    double nb = 0.0, nr = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        nb += hb[i] * hb[i];
        nr += hr[i] * hr[i];
    }
    const double rel = sqrt(nr) / sqrt(nb);
    printf("ugx_client refs=%d dofs=%lld converged=%d steps=%d lattice_path=%d relres=%.3e\n", refs, (long long)N,
           ok, ugx_solver_steps(cg), ugx_solver_used_lattice_path(cg), rel);
    ugx_solver_destroy(cg);
    ugx_gmg_destroy(gmg);
    ugx_vector_destroy(r);
    ugx_vector_destroy(x);
    ugx_vector_destroy(b);
    ugx_matrix_destroy(A);
    ugx_disc_destroy(disc);
    free(theta);
    free(hb);
    free(hr);
    return (ok == 1 && rel <= 1e-9) ? 0 : 2;
}
