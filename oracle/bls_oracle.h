/*
 * bls_oracle.h -- CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement of the elimination kernels of BlockLinearSystems.jl, used by the tests, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker. Nothing in the product
 * (the HIP library under blocklinearsystems.jl_amd/) links, loads or calls it.
 *
 * Conventions are the reference's, verbatim: every index array holds 1-based Int64 values exactly as the
 * Julia structures store them; blocks are column-major; C_er is stored d_e x d_r (src/schur_assembly.jl:236-246).
 * Element type T is double ("_f64") or float ("_f32"); b, z, g_S, damping and scales are always double, as in the
 * reference (src/schur_reduction.jl:116-120).
 *
 * Parity status per function is recorded in the header of bls_oracle.c.
 */
#ifndef BLS_ORACLE_H
#define BLS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The fields of SchurStructure (src/schur_structure.jl:50-68) and of SchurReductionWorkspace's index
 * (src/schur_reduction.jl:16-27) that the kernels read. Phase-1 chunks are flattened: ranges of all runs,
 * in run order, `chunk_run[i]` naming the run of range i (ranges of one run are consecutive). */
typedef struct {
    int64_t nr, nrblocks, neblocks, ncoup;
    const int64_t *r_dims, *e_dims;
    const int64_t *boffsets_r, *boffsets_e, *eoffsets;
    const int64_t *eliminated_ptr, *entry_retained, *coffsets;
    int64_t nranges;
    const int64_t *chunk_lo, *chunk_hi, *chunk_run;
    const int64_t *retained_ptr, *retained_entry, *entry_eliminated;
    int64_t ncol_chunks;
    const int64_t *col_lo, *col_hi;
} orc_schur;

/* The integer part of build_schur_pattern's result (src/schur_structure.jl:241-323) plus S.colptr. */
typedef struct {
    const int64_t *S_colptr, *S_rowval;
    const int64_t *block_colptr, *block_rowval, *block_row_offsets, *col_heights;
} orc_pattern;

/* SchurJacobianStructure (src/schur_jacobian.jl:30-57); `schur` supplies dims, offsets and chunks. */
typedef struct {
    int64_t ntables;
    const int64_t *table_nrows, *table_ncols, *table_top;
    const int64_t *raw_offsets, *residual_offsets, *reduced_offsets, *doffsets;
    const int64_t *table_ptr, *entry_retained, *entry_col, *entry_table;
    const int64_t *retained_ptr, *retained_entry;
    int64_t ncol_chunks;
    const int64_t *col_lo, *col_hi;
} orc_jacobian;

/* nthreads <= 1 runs serially; >1 runs the reference's chunked loops on OpenMP threads (the Polyester
 * structure of ext/PolyesterExt.jl:12-23). Results do not depend on nthreads. */

#define ORC_DECL(T, S)                                                                                          \
    int orc_inv_sym_block_##S(int64_t d, const T *H, const T *damp, T *invH);                                  \
    int orc_inv_sym_block_chol_##S(int64_t d, const T *H, const T *damp, T *invH);                             \
    int orc_reset_schur_dense_##S(int64_t nr, T *S, const T *R, const double *scaling_r, double lambda,        \
                                  const double *s_r);                                                           \
    int orc_reset_schur_sparse_##S(int64_t nnz, T *nzval, int64_t nR, const T *R_nzval,                       \
                                   const int64_t *r_to_s_map, int64_t nr, const int64_t *diag_inds,            \
                                   const double *scaling_r, double lambda, const double *s_r,                  \
                                   const int64_t *S_colptr, const int64_t *S_rowval);                          \
    int orc_reduce_schur_##S(int sparse, const orc_schur *st, const orc_pattern *pat, T *S, double *g_S,       \
                             const double *b_r, const double *b_e, const double *scaling_e, double lambda,     \
                             const T *E_data, const T *C_data, T *E_inv, double *z, T *W,                      \
                             const double *s_r, const double *s_e, int64_t *failed_block, int nthreads);       \
    void orc_backsubstitute_schur_##S(const orc_schur *st, double *x_e, const double *x_r, const T *C_data,    \
                                      const T *E_inv, const double *z, const double *s_r, const double *s_e);  \
    void orc_copyto_##S(int64_t n, T *dest, const T *src, const int64_t *scatter_map);                        \
    int orc_reduce_qr_##S(const orc_schur *st, const orc_jacobian *jst, T *F, double *Fr, const T *Rw,         \
                          const double *Rr, const double *scaling_e, double lambda, const double *cs,          \
                          int64_t *failed_block, int nthreads);                                                 \
    void orc_qr_backsubstitute_##S(const orc_schur *st, const orc_jacobian *jst, double *x_e,                  \
                                   const double *x_r, const T *F, const double *Fr);                           \
    void orc_reduced_normal_##S(int sparse, const orc_schur *st, const orc_jacobian *jst,                      \
                                const orc_pattern *pat, T *S, double *g_S, const T *F, const double *Fr);       \
    void orc_reduced_diagonal_blocks_##S(const orc_schur *st, const orc_jacobian *jst, T *D, const T *F);      \
    void orc_reduced_mul_##S(const orc_schur *st, const orc_jacobian *jst, double *y, const T *F,              \
                             const double *v);                                                                  \
    void orc_reduced_tmul_##S(const orc_schur *st, const orc_jacobian *jst, double *g, const T *F,             \
                              const double *w);                                                                 \
    void orc_reduced_residual_##S(const orc_jacobian *jst, double *rt, const double *Fr);                      \
    int orc_factor_diagonal_blocks_##S(int64_t nblocks, const int64_t *dims, T *D);                            \
    void orc_solve_diagonal_blocks_##S(int64_t nblocks, const int64_t *dims, double *y, const T *D,            \
                                       const double *x);

ORC_DECL(double, f64)
ORC_DECL(float, f32)

/* Symmetric block-sparse S (upper block CSC of build_schur_pattern, diagonal blocks read from their upper
 * triangle) times x, in double: y = S x. */
void orc_sym_spmv_f64(int64_t nr, int64_t nrblocks, const int64_t *r_dims, const int64_t *boffsets_r,
                      const orc_pattern *pat, const double *nzval, const double *x, double *y);
/* Textbook preconditioned CG (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 9.1) on the
 * sparse reduced system, block-Jacobi preconditioner from the upper triangles of S's diagonal blocks.
 * Stops when ||S x - g|| / ||g|| <= tol (true residual, double) or after maxit iterations.
 * Returns the iteration count (negative if the preconditioner could not be factored). */
int64_t orc_pcg_f64(int64_t nr, int64_t nrblocks, const int64_t *r_dims, const int64_t *boffsets_r,
                    const orc_pattern *pat, const double *nzval, const double *g, double *x, double tol,
                    int64_t maxit, const double *precond_blocks, double *relres_out);
/* The block-Jacobi blocks of S: full d x d blocks (lower triangle mirrored from upper), flat, column-major. */
void orc_diag_blocks_from_sparse_f64(int64_t nrblocks, const int64_t *r_dims, const int64_t *boffsets_r,
                                     const orc_pattern *pat, const double *nzval, double *D);

#ifdef __cplusplus
}
#endif
#endif
