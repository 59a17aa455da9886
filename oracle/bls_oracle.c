/*
 * bls_oracle.c -- CPU oracle for the elimination hot path of BlockLinearSystems.jl.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker; never by the product library. Every function restates the reference function cited beside it,
 * operation by operation, in the same loop order (compiled with -ffp-contract=off, so no fused multiply-adds
 * are introduced that the Julia code does not have).
 *
 * Parity pinning (what checks each function, beyond re-reading the Julia):
 *   inv_sym_block ......... numpy inverse + the Cholesky acceptance rule on random SPD / singular blocks, the
 *                           known-answer behaviour of test/kernels.jl:11-49 (closed forms accept exactly what
 *                           the Cholesky path accepts)                          -> tests/test_oracle_numerics.py
 *   reduce_schur_{dense,sparse}, reset_schur_{dense,sparse}
 *                         dense brute force S = R + lambda D_r - C'(E + lambda D_e)^-1 C, g likewise; full solve
 *                           equals dense Cholesky of the damped H (test/fixtures/SchurFixtures.jl:114-117);
 *                           rank-1 E at lambda = 0 fails at the lowest block (test/schur.jl:405-425);
 *                           nthreads 1 vs 4 bitwise (test/schur.jl:255-256)      -> tests/test_oracle_numerics.py
 *   backsubstitute ........ full solve vs dense Cholesky                      -> tests/test_oracle_numerics.py
 *   copyto ................ literal fixture out2 / symmetrifysparse of test/SparseBlockLinearSystem.jl:140-238
 *   reduce_qr ............. U'U = E + lambda diag, U'B = C, U'q = b_e (test/qr_reduction.jl:20-58); reduced
 *                           normal equations equal the Schur reduction (test/schur_qr.jl:13-57); dense least
 *                           squares (test/fixtures/QRFixtures.jl:164-169)       -> tests/test_oracle_qr.py
 *   reduced_{mul,tmul,residual,normal,diagonal_blocks}, qr backsubstitute: dense J~ products / solve
 *   factor/solve_diagonal_blocks: numpy Cholesky per block (test/diagonal_blocks.jl:16-34)
 *   pcg ................... no reference counterpart (the reference has no PCG): pinned by the true residual
 *                           ||S x - g|| / ||g|| and by scipy's direct solve   -> tests/test_oracle_numerics.py
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "bls_oracle.h"

#define T double
#define S f64
#include "bls_oracle_t.h"
#undef T
#undef S

#define T float
#define S f32
#include "bls_oracle_t.h"
#undef T
#undef S

/* y = S x for the symmetric reduced system stored as its upper block pattern (build_schur_pattern,
 * src/schur_structure.jl:241-323): block (r1, r2), r1 <= r2, at nzval[colptr[ro2] + (col-1)*col_heights[r2]
 * + block_row_offsets[idx] + row - 1]; a diagonal block is read from its upper triangle (the reduction writes
 * only that, src/schur_reduction.jl:283). */
void orc_sym_spmv_f64(int64_t nr, int64_t nrblocks, const int64_t *r_dims, const int64_t *boffsets_r,
                      const orc_pattern *pat, const double *nzval, const double *x, double *y)
{
    for (int64_t i = 0; i < nr; i++) y[i] = 0.0;
    for (int64_t r2 = 1; r2 <= nrblocks; r2++) {
        const int64_t ro2 = A1(boffsets_r, r2), dr2 = A1(r_dims, r2);
        const int64_t base = A1(pat->S_colptr, ro2), stride = A1(pat->col_heights, r2);
        for (int64_t idx = A1(pat->block_colptr, r2); idx < A1(pat->block_colptr, r2 + 1); idx++) {
            const int64_t r1 = A1(pat->block_rowval, idx);
            const int64_t ro1 = A1(boffsets_r, r1), dr1 = A1(r_dims, r1);
            const int64_t off = base + A1(pat->block_row_offsets, idx);
            for (int64_t col = 1; col <= dr2; col++)
                for (int64_t row = 1; row <= dr1; row++) {
                    if (r1 == r2 && row > col) continue;
                    const double v = A1(nzval, off + (col - 1) * stride + row - 1);
                    A1(y, ro1 + row - 1) += v * A1(x, ro2 + col - 1);
                    if (!(r1 == r2 && row == col)) A1(y, ro2 + col - 1) += v * A1(x, ro1 + row - 1);
                }
        }
    }
}

void orc_diag_blocks_from_sparse_f64(int64_t nrblocks, const int64_t *r_dims, const int64_t *boffsets_r,
                                     const orc_pattern *pat, const double *nzval, double *D)
{
    int64_t o = 0;
    for (int64_t r = 1; r <= nrblocks; r++) {
        const int64_t d = A1(r_dims, r), ro = A1(boffsets_r, r);
        const int64_t base = A1(pat->S_colptr, ro), stride = A1(pat->col_heights, r);
        int64_t idx = A1(pat->block_colptr, r);
        while (A1(pat->block_rowval, idx) != r) idx++;
        const int64_t off = base + A1(pat->block_row_offsets, idx);
        for (int64_t col = 1; col <= d; col++)
            for (int64_t row = 1; row <= d; row++) {
                const int64_t a = row <= col ? row : col, b = row <= col ? col : row;
                D[o + (col - 1) * d + row - 1] = A1(nzval, off + (b - 1) * stride + a - 1);
            }
        o += d * d;
    }
}

static double dot64(int64_t n, const double *a, const double *b)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int64_t orc_pcg_f64(int64_t nr, int64_t nrblocks, const int64_t *r_dims, const int64_t *boffsets_r,
                    const orc_pattern *pat, const double *nzval, const double *g, double *x, double tol,
                    int64_t maxit, const double *precond_blocks, double *relres_out)
{
    int64_t nd = 0;
    for (int64_t r = 0; r < nrblocks; r++) nd += r_dims[r] * r_dims[r];
    double *L = (double *)malloc(sizeof(double) * nd);
    double *r = (double *)malloc(sizeof(double) * nr), *z = (double *)malloc(sizeof(double) * nr);
    double *p = (double *)malloc(sizeof(double) * nr), *q = (double *)malloc(sizeof(double) * nr);
    int64_t it = 0;
    if (precond_blocks) memcpy(L, precond_blocks, sizeof(double) * nd);
    else orc_diag_blocks_from_sparse_f64(nrblocks, r_dims, boffsets_r, pat, nzval, L);
    if (!orc_factor_diagonal_blocks_f64(nrblocks, r_dims, L)) { it = -1; goto done; }
    const double gn = sqrt(dot64(nr, g, g));
    /* r0 = g - S x0 */
    orc_sym_spmv_f64(nr, nrblocks, r_dims, boffsets_r, pat, nzval, x, q);
    for (int64_t i = 0; i < nr; i++) r[i] = g[i] - q[i];
    double rel = gn > 0 ? sqrt(dot64(nr, r, r)) / gn : 0.0;
    orc_solve_diagonal_blocks_f64(nrblocks, r_dims, z, L, r);
    for (int64_t i = 0; i < nr; i++) p[i] = z[i];
    double rz = dot64(nr, r, z);
    while (rel > tol && it < maxit) {
        orc_sym_spmv_f64(nr, nrblocks, r_dims, boffsets_r, pat, nzval, p, q);
        const double alpha = rz / dot64(nr, p, q);
        for (int64_t i = 0; i < nr; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        it++;
        rel = gn > 0 ? sqrt(dot64(nr, r, r)) / gn : 0.0;
        if (rel <= tol) { /* confirm with the true residual */
            orc_sym_spmv_f64(nr, nrblocks, r_dims, boffsets_r, pat, nzval, x, q);
            for (int64_t i = 0; i < nr; i++) r[i] = g[i] - q[i];
            rel = gn > 0 ? sqrt(dot64(nr, r, r)) / gn : 0.0;
            if (rel <= tol) break;
        }
        orc_solve_diagonal_blocks_f64(nrblocks, r_dims, z, L, r);
        const double rz_new = dot64(nr, r, z);
        const double beta = rz_new / rz;
        rz = rz_new;
        for (int64_t i = 0; i < nr; i++) p[i] = z[i] + beta * p[i];
    }
    if (relres_out) *relres_out = rel;
done:
    free(L); free(r); free(z); free(p); free(q);
    return it;
}
